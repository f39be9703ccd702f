#!/usr/bin/env python3
"""Python entry points of the SPME engine -- a drop-in for the reference's
ewald.py (frobnitzem/EwaldCorrel ewald.py:1-141).

Same names and call shapes: ``describe_sfac``, ``Sfac(L, K, order=4)`` with
``.sfac(w, x)``, ``.get_S()``, ``.set_A(f, max_m)``, ``.set_L(L)`` and the
forwarded ``.en()``, ``.de1(vir)``, ``.de2(N, w, x, dx)``, ``.potl(N, x,
phi)``; module-level ``eta``, ``ewald_f``, ``LofS``, ``SofL``, ``array``.
``libsfac.so`` is loaded from this file's directory (ewald.py:52-53); it is
the HIP library, and there is no fallback when it is missing.

Additions (keyword arguments and new methods only):
  * ``Sfac(..., precision="fp64"|"fp32", fixed=False, fp32_influence=False)``
  * ``set_A`` recognises the closed-form Ewald kernel (``ewald_f`` or a
    kernel from ``ewald_kernel(eta)``) and builds the table natively,
    instead of one Python callback per lattice point
  * ``evaluate(w, x)`` -> (E, vir, dx, phi): one evaluation in one call
  * ``*_dev`` methods taking device pointers (ints), asynchronous on
    ``stream``; ``sync()``
"""
import os
from functools import reduce
from math import pi, exp, erfc  # noqa: F401  (reference module namespace)

from ctypes import CDLL, c_void_p, c_int, c_double, c_char_p, c_longlong, \
    POINTER, PYFUNCTYPE, cast, create_string_buffer

import numpy as np
import numpy.ctypeslib as ct

array = np.array

SFAC_FP64 = 0
SFAC_FP32 = 1
SFAC_FIXED = 2
SFAC_FP32_INFLUENCE = 4

_LIB = None
# double (*rad_fn)(double r2, double *dA, void *info), called with the GIL held
_RAD_FN = PYFUNCTYPE(c_double, c_double, POINTER(c_double), c_void_p)


def _libpath():
    return os.path.join(os.path.dirname(os.path.abspath(__file__)), "libsfac.so")


def describe_sfac():
    """Load libsfac.so and declare the C signatures (ewald.py:15-68).

    Returns (sfac_p, rad_fn, lib) like the reference."""
    global _LIB

    def decl_fn(a, *args):
        a.argtypes = args[:-1]
        a.restype = args[-1]

    def void_fn(a, *args):
        decl_fn(a, *(args + (None,)))

    def int_fn(a, *args):
        decl_fn(a, *(args + (c_int,)))

    def dbl_fn(a, *args):
        decl_fn(a, *(args + (c_double,)))

    nparr = np.ctypeslib.ndpointer(dtype=np.float64, flags='C_CONTIGUOUS')
    dblarr = POINTER(c_double)

    def nparr_t(*shape):
        return np.ctypeslib.ndpointer(dtype=np.float64, shape=shape, flags='C_CONTIGUOUS')

    def np_intarr_t(*shape):
        return np.ctypeslib.ndpointer(dtype=np.int32, shape=shape, flags='C_CONTIGUOUS')

    sfac_p = c_void_p
    rad_fn = _RAD_FN  # one type for the cached library's argtypes

    if _LIB is None:
        path = _libpath()
        if not os.path.exists(path):
            raise ImportError(
                "libsfac.so (the HIP SPME library) is not built; run "
                "`python -m ewaldcorrel_amd._build` (expected at %s)" % path)
        lib = CDLL(path)
        decl_fn(lib.sfac_ctor, nparr_t(6), np_intarr_t(3), c_int, sfac_p)
        void_fn(lib.sfac_dtor, sfac_p)
        void_fn(lib.sfac, sfac_p, c_int, nparr, nparr)
        decl_fn(lib.get_S, sfac_p, dblarr)
        void_fn(lib.set_L, sfac_p, nparr_t(6))
        int_fn(lib.set_A, sfac_p, rad_fn, c_double, c_void_p)
        dbl_fn(lib.en, sfac_p)
        dbl_fn(lib.de1, sfac_p, nparr_t(6))
        void_fn(lib.de2, sfac_p, c_int, nparr, nparr, nparr)
        void_fn(lib.potl, sfac_p, c_int, nparr, nparr)
        # extensions (sfac_ext.h)
        decl_fn(lib.sfac_ctor_ex, nparr_t(6), np_intarr_t(3), c_int, c_int, sfac_p)
        decl_fn(lib.sfac_last_error, c_char_p)
        int_fn(lib.sfac_set_A_ewald, sfac_p, c_double, c_double)
        int_fn(lib.sfac_table, sfac_p, c_void_p, c_void_p)
        int_fn(lib.sfac_accum_S2, sfac_p, c_void_p)
        int_fn(lib.sfac_accum_cross, sfac_p, sfac_p, c_void_p)
        int_fn(lib.sfac_radial, sfac_p, c_void_p, c_int, c_double, c_void_p, c_void_p)
        int_fn(lib.sfac_radial_plot, sfac_p, c_void_p, c_double, c_double, c_int, c_void_p,
               c_void_p)
        int_fn(lib.sfac_real_space, sfac_p, c_int, nparr, nparr, c_double, c_double,
               nparr_t(10), c_void_p, c_void_p)
        int_fn(lib.sfac_real_space_dev, sfac_p, c_int, c_void_p, c_void_p, c_double, c_double,
               c_void_p, c_void_p, c_void_p)
        int_fn(lib.sfac_dev, sfac_p, c_int, c_void_p, c_void_p)
        int_fn(lib.en_dev, sfac_p, c_void_p)
        int_fn(lib.de1_dev, sfac_p, c_void_p)
        int_fn(lib.de2_dev, sfac_p, c_int, c_void_p, c_void_p, c_void_p)
        int_fn(lib.potl_dev, sfac_p, c_int, c_void_p, c_void_p)
        int_fn(lib.sfac_eval_dev, sfac_p, c_int, c_void_p, c_void_p, c_void_p,
               c_void_p, c_void_p)
        int_fn(lib.sfac_eval_dev_noalloc, sfac_p, c_int, c_void_p, c_void_p, c_void_p,
               c_void_p, c_void_p)
        dbl_fn(lib.sfac_eval, sfac_p, c_int, nparr, nparr, nparr_t(6), nparr, nparr)
        decl_fn(lib.sfac_stream, sfac_p, c_void_p)
        int_fn(lib.sfac_sync, sfac_p)
        int_fn(lib.sfac_info, sfac_p, POINTER(c_longlong))
        int_fn(lib.sfac_set_graph, sfac_p, c_int)
        int_fn(lib.sfac_set_transform, sfac_p, c_int)
        decl_fn(lib.sfac_batch_create, POINTER(c_void_p), c_int, c_void_p)
        int_fn(lib.sfac_batch_set, c_void_p, c_int, c_int, c_void_p, c_void_p, c_void_p,
               c_void_p, c_void_p)
        int_fn(lib.sfac_batch_eval, c_void_p, c_void_p)
        int_fn(lib.sfac_batch_sync, c_void_p)
        int_fn(lib.sfac_batch_set_graph, c_void_p, c_int)
        int_fn(lib.sfac_batch_set_lanes, c_void_p, c_int)
        int_fn(lib.sfac_batch_lanes, c_void_p)
        int_fn(lib.sfac_batch_destroy, c_void_p)
        decl_fn(lib.sfac_server_create, c_int, c_double, c_int, c_int, c_void_p)
        decl_fn(lib.sfac_server_submit, c_void_p, c_void_p, c_void_p, c_int, c_int, c_double,
                c_double, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_longlong)
        int_fn(lib.sfac_server_wait, c_void_p, c_longlong, c_void_p)
        int_fn(lib.sfac_server_poll, c_void_p, c_longlong)
        int_fn(lib.sfac_server_stats, c_void_p, c_void_p)
        int_fn(lib.sfac_server_reset_stats, c_void_p)
        int_fn(lib.sfac_server_destroy, c_void_p)
        int_fn(lib.sfac_autotune, sfac_p, c_int, c_void_p, c_void_p, c_int, c_char_p, c_int)
        int_fn(lib.sfac_batch_autotune, c_void_p, c_int, c_char_p, c_int)
        int_fn(lib.sfac_tune_export, c_char_p, c_int)
        int_fn(lib.sfac_tune_import, c_char_p)
        int_fn(lib.sfac_tune_clear)
        int_fn(lib.sfac_tune_current, sfac_p, c_char_p, c_int)
        int_fn(lib.sfac_profile, sfac_p, c_int)
        int_fn(lib.sfac_profile_read, sfac_p, c_char_p, c_int, POINTER(c_double),
               POINTER(c_int), c_int)
        _LIB = lib
    return sfac_p, rad_fn, _LIB


def last_error():
    _, _, lib = describe_sfac()
    return lib.sfac_last_error().decode()


def _check(rc, what):
    if rc != 0:
        raise RuntimeError("%s failed (%d): %s" % (what, rc, last_error()))


class Sfac:
    """Handle on one SPME system (ewald.py:70-122)."""

    def __init__(self, L, K, order=4, precision="fp64", fixed=False, fp32_influence=False):
        sfac_p, rad_fn, libsfac = describe_sfac()
        self._libsfac = libsfac
        self.rad_fn = rad_fn
        K = np.asarray(K)
        self.K = tuple(K.tolist())
        self.ldim = K[2] // 2 + 1
        flags = 0
        if precision == "fp32":
            flags |= SFAC_FP32
        elif precision != "fp64":
            raise ValueError("precision must be 'fp64' or 'fp32'")
        if fixed:
            flags |= SFAC_FIXED
        if fp32_influence:
            if precision != "fp32":
                raise ValueError("fp32_influence applies to precision='fp32' only")
            flags |= SFAC_FP32_INFLUENCE
        self.precision = precision
        L = np.ascontiguousarray(np.asarray(L, dtype=np.float64))
        self._L6 = L.copy()
        self._data = libsfac.sfac_ctor_ex(L, np.ascontiguousarray(K.astype(np.int32)),
                                          order, flags)
        if not self._data:
            raise RuntimeError("Error intializing Sfac. (%s)" % last_error())

    def __getattr__(self, name):
        """Pass the handle as the first argument to the forwarded C
        functions (ewald.py:83-95)."""
        names = ["sfac_dtor", "en", "de1", "de2", "potl"]
        if name not in names:
            raise AttributeError(name)
        try:
            f = getattr(self._libsfac, name)
        except AttributeError:
            return None
        return lambda *args, **kws: f(self._data, *args, **kws)

    def get_S(self):
        """Complex (K0, K1, ldim) copy of the grid (ewald.py:97-101)."""
        Q = self._libsfac.get_S(self._data)
        if not Q:
            raise RuntimeError("get_S failed: %s" % last_error())
        shape = (self.K[0], self.K[1], self.ldim, 2)
        s = ct.as_array(Q, shape=shape)
        return s[:, :, :, 0] + 1j * s[:, :, :, 1]

    def sfac(self, w, x):
        assert len(w.shape) == 1
        assert x.shape == (len(w), 3)
        self._libsfac.sfac(self._data, len(w), np.ascontiguousarray(w, dtype=np.float64),
                           np.ascontiguousarray(x, dtype=np.float64))

    def __del__(self):
        if getattr(self, "_data", None):
            self._libsfac.sfac_dtor(self._data)
            self._data = None

    def set_A(self, f, max_m):
        self.f = f
        eta_native = _native_eta(f)
        if eta_native is not None:
            n = self._libsfac.sfac_set_A_ewald(self._data, eta_native, max_m)
        else:
            n = self._libsfac.set_A(self._data, self.rad_fn(f), max_m, cast(0, c_void_p))
        if n != 0:
            raise RuntimeError("Error (%d) setting A-array." % n)

    def set_L(self, L):
        self._libsfac.set_L(self._data, np.ascontiguousarray(L, dtype=np.float64))
        self._L6 = np.array(L, dtype=np.float64)
        if hasattr(self, "f"):
            self.set_A(self.f, 1.0)

    # ---- additions -------------------------------------------------------
    def evaluate(self, w, x):
        """One evaluation (sfac + de1 + de2 + potl); returns E, vir, dx, phi."""
        w = np.ascontiguousarray(w, dtype=np.float64)
        x = np.ascontiguousarray(x, dtype=np.float64)
        n = len(w)
        vir = np.zeros(6)
        dx = np.zeros((n, 3))
        phi = np.zeros(n)
        E = self._libsfac.sfac_eval(self._data, n, w, x, vir, dx, phi)
        err = last_error()
        if err:
            raise RuntimeError("evaluate failed: " + err)
        return E, vir, dx, phi

    @property
    def stream(self):
        return self._libsfac.sfac_stream(self._data)

    def sync(self):
        _check(self._libsfac.sfac_sync(self._data), "sync")

    def sfac_dev(self, n, w_ptr, x_ptr):
        _check(self._libsfac.sfac_dev(self._data, n, w_ptr, x_ptr), "sfac_dev")

    def en_dev(self, out_ptr):
        _check(self._libsfac.en_dev(self._data, out_ptr), "en_dev")

    def de1_dev(self, out_ptr):
        _check(self._libsfac.de1_dev(self._data, out_ptr), "de1_dev")

    def de2_dev(self, n, w_ptr, x_ptr, dx_ptr):
        _check(self._libsfac.de2_dev(self._data, n, w_ptr, x_ptr, dx_ptr), "de2_dev")

    def potl_dev(self, n, x_ptr, phi_ptr):
        _check(self._libsfac.potl_dev(self._data, n, x_ptr, phi_ptr), "potl_dev")

    def eval_dev(self, n, w_ptr, x_ptr, out_ptr, dx_ptr, phi_ptr):
        _check(self._libsfac.sfac_eval_dev(self._data, n, w_ptr, x_ptr, out_ptr,
                                           dx_ptr, phi_ptr), "sfac_eval_dev")

    def eval_dev_noalloc(self, n, w_ptr, x_ptr, out_ptr, dx_ptr, phi_ptr):
        """eval_dev with no device allocation (workspace in the library's
        static scratch of this GPU; overlapping calls on one GPU race)."""
        _check(self._libsfac.sfac_eval_dev_noalloc(self._data, n, w_ptr, x_ptr, out_ptr,
                                                   dx_ptr, phi_ptr), "sfac_eval_dev_noalloc")

    def profile(self, enable=True):
        """Bracket every kernel launch with HIP events on ``stream``."""
        _check(self._libsfac.sfac_profile(self._data, int(bool(enable))), "profile")

    def profile_read(self):
        """{kernel: (total_ms, launches)} recorded since profiling was enabled."""
        import ctypes
        mx = 64
        names = ctypes.create_string_buffer(4096)
        ms = (c_double * mx)()
        cnt = (c_int * mx)()
        n = self._libsfac.sfac_profile_read(self._data, names, 4096, ms, cnt, mx)
        if n < 0:
            raise RuntimeError("profile_read failed: " + last_error())
        keys = names.value.decode().split("\n") if n else []
        return {k: (ms[i], cnt[i]) for i, k in enumerate(keys)}

    def set_graph(self, enable=True):
        """Replay eval_dev/evaluate from a captured hipGraph (default on)."""
        _check(self._libsfac.sfac_set_graph(self._data, int(bool(enable))), "set_graph")

    def table(self):
        """(A[H], dA[H, 6]) of the current influence function, as the
        reference's set_A builds them (sfac.cpp:213-292)."""
        K0, K1, K2 = (int(k) for k in self.K)
        H = K0 * K1 * (K2 // 2 + 1)
        A = np.zeros(H)
        dA = np.zeros((H, 6))
        _check(self._libsfac.sfac_table(self._data, A.ctypes.data, dA.ctypes.data), "table")
        return A, dA

    # ---- structure-factor accumulation (device pointers, FP64) -----------
    def accum_S2(self, acc_ptr):
        """acc[H] += |S_hat|^2 after sfac (run.py:45), on the device."""
        _check(self._libsfac.sfac_accum_S2(self._data, acc_ptr), "accum_S2")

    def accum_cross(self, other, acc_ptr):
        """acc[2H] += S_hat conj(S_hat of other) (gmx_ewald.py:101-104)."""
        _check(self._libsfac.sfac_accum_cross(self._data, other._data, acc_ptr),
               "accum_cross")

    def radial(self, val_ptr, nbins, dk, sum_ptr, cnt_ptr):
        """|k| shell sums and counts of a per-point array (plot.py:19-32)."""
        _check(self._libsfac.sfac_radial(self._data, val_ptr, int(nbins), float(dk),
                                         sum_ptr, cnt_ptr), "radial")

    def radial_plot(self, val_ptr, dk, clip, nbins, sum_ptr, cnt_ptr):
        """|k| shell sums and counts with plot.py's k-vectors and clipping
        (plot.py:9-32), on the device; see sfac_radial_plot."""
        _check(self._libsfac.sfac_radial_plot(self._data, val_ptr, float(dk), float(clip),
                                              int(nbins), sum_ptr, cnt_ptr), "radial_plot")

    def plot_spectrum(self, val):
        """plot.py's output for a (K0, K1, ldim) array of this handle's grid
        (a device tensor or host array, e.g. the real part of a sum of
        |S|^2): rows (k, mean, count) of its non-empty |k| shells, binned on
        the device.  plot.py's box length is taken from the handle (L6[0])."""
        import math
        import torch
        v = torch.as_tensor(val, dtype=torch.float64, device="cuda").reshape(-1).contiguous()
        K0 = int(self.K[0])
        L = float(self._L6[0])
        kmax = 0.25 * K0 * math.pi / L
        dk = kmax / 5000
        clip = kmax / dk + 1
        nbins = int(clip) + 1
        sm = torch.zeros(nbins, dtype=torch.float64, device="cuda")
        ct = torch.zeros(nbins, dtype=torch.float64, device="cuda")
        torch.cuda.synchronize()
        self.radial_plot(v.data_ptr(), dk, clip, nbins, sm.data_ptr(), ct.data_ptr())
        self.sync()
        s, c = sm.cpu().numpy(), ct.cpu().numpy()
        nz = np.nonzero(c)[0]
        m = int(nz[-1]) if len(nz) else 0
        return [(i * dk, s[i] / c[i], c[i]) for i in range(1, m) if c[i] > 0]

    # ---- real-space Ewald ------------------------------------------------
    def real_space(self, w, x, eta, rc):
        """Real-space Ewald sum over the images within rc plus the self and
        neutralising-background terms, in the current box.  Returns a dict
        with E (pairs), vir (6), dx (= dE/dx of the pairs), phi (dE/dw of
        all terms), E_self, E_bg and vir_bg."""
        w = np.ascontiguousarray(w, dtype=np.float64)
        x = np.ascontiguousarray(x, dtype=np.float64)
        n = len(w)
        out = np.zeros(10)
        dx = np.zeros((n, 3))
        phi = np.zeros(n)
        _check(self._libsfac.sfac_real_space(self._data, n, w, x, float(eta), float(rc), out,
                                             dx.ctypes.data, phi.ctypes.data), "real_space")
        return dict(E=out[0], vir=out[1:7].copy(), E_self=out[7], E_bg=out[8], vir_bg=out[9],
                    dx=dx, phi=phi)

    def real_space_dev(self, n, w_ptr, x_ptr, eta, rc, out_ptr, dx_ptr=None, phi_ptr=None):
        _check(self._libsfac.sfac_real_space_dev(self._data, n, w_ptr, x_ptr, float(eta),
                                                 float(rc), out_ptr, dx_ptr, phi_ptr),
               "real_space_dev")

    def autotune(self, n=None, w_ptr=None, x_ptr=None, reps=10):
        """Time and choose launch shapes (transform path, FFT radix orders,
        spreading and gather tiles) for this handle; returns the report.
        Without pointers, n synthetic atoms (default: 0.1 per cubic A of the
        box).  The grid is overwritten (call sfac again afterwards)."""
        if n is None:
            Lm = LofS(np.asarray(self._L6))
            n = max(1000, int(0.1 * abs(np.linalg.det(Lm))))
        buf = create_string_buffer(1 << 16)
        _check(self._libsfac.sfac_autotune(self._data, int(n), w_ptr, x_ptr, int(reps), buf,
                                           len(buf)), "sfac_autotune")
        return buf.value.decode()

    def launch_shapes(self):
        """The handle's current launch shapes as a tune_export() line."""
        buf = create_string_buffer(512)
        rc = self._libsfac.sfac_tune_current(self._data, buf, len(buf))
        if rc < 0:
            raise RuntimeError("sfac_tune_current failed: " + last_error())
        return buf.value.decode()

    def set_transform(self, mode):
        """0: y-z slab transforms where available, the spreading fused into the
        forward one where its plane of sums fits in LDS (default); 1: z and y
        line passes; 2: y-z slab transforms after separate spreading tiles.
        info()["slab"] reports the path: 0 lines, 1 slab, 2 slab with fused
        spreading."""
        _check(self._libsfac.sfac_set_transform(self._data, int(mode)), "set_transform")

    def info(self):
        out = (c_longlong * 12)()
        _check(self._libsfac.sfac_info(self._data, out), "info")
        keys = ("K0", "K1", "K2", "ldim", "order", "flags", "device_bytes", "state",
                "table_mode", "graph", "slab", "slab_available")
        return dict(zip(keys, list(out)))


class SfacBatch:
    """Frames (Sfac handles with one precision, accumulation mode and spline
    order, influence functions set) evaluated together: one launch per stage
    for all frames, replayed from a captured hipGraph while the per-frame
    device pointers stay the same."""

    def __init__(self, frames):
        _, _, lib = describe_sfac()
        self._lib = lib
        self.frames = list(frames)
        arr = (c_void_p * len(self.frames))(*[f._data for f in self.frames])
        self._b = lib.sfac_batch_create(arr, len(self.frames))
        if not self._b:
            raise RuntimeError("sfac_batch_create failed: " + last_error())

    def set(self, f, n, w_ptr, x_ptr, out_ptr, dx_ptr=None, phi_ptr=None):
        _check(self._lib.sfac_batch_set(self._b, f, n, w_ptr, x_ptr, out_ptr, dx_ptr,
                                        phi_ptr), "sfac_batch_set")

    def eval(self, stream=None):
        """One evaluation of every frame on `stream` (None: the first
        frame's stream, waited for by sync())."""
        _check(self._lib.sfac_batch_eval(self._b, stream), "sfac_batch_eval")

    def sync(self):
        _check(self._lib.sfac_batch_sync(self._b), "sfac_batch_sync")

    def set_graph(self, enable=True):
        _check(self._lib.sfac_batch_set_graph(self._b, int(bool(enable))), "batch set_graph")

    def set_lanes(self, k):
        """Evaluate the frames as `k` groups of consecutive frames (1-4),
        each on its own stream (the first on the evaluation's stream, group g
        on its first frame's stream, forked and joined by events)."""
        _check(self._lib.sfac_batch_set_lanes(self._b, int(k)), "sfac_batch_set_lanes")

    def lanes(self):
        k = self._lib.sfac_batch_lanes(self._b)
        _check(min(k, 0), "sfac_batch_lanes")
        return k

    def autotune(self, reps=10):
        """Time and choose launch shapes for the frames (after set() of every
        frame); returns the report."""
        buf = create_string_buffer(1 << 16)
        _check(self._lib.sfac_batch_autotune(self._b, int(reps), buf, len(buf)),
               "sfac_batch_autotune")
        return buf.value.decode()

    def close(self):
        if getattr(self, "_b", None):
            self._lib.sfac_batch_destroy(self._b)
            self._b = None

    def __del__(self):
        self.close()


def tune_export():
    """Launch shapes remembered by the autotuner in this process (text)."""
    _, _, lib = describe_sfac()
    need = lib.sfac_tune_export(None, 0)
    buf = create_string_buffer(max(need, 1))
    lib.sfac_tune_export(buf, len(buf))
    return buf.value.decode()


def tune_import(text):
    """Add remembered launch shapes from tune_export() text; returns lines read."""
    _, _, lib = describe_sfac()
    k = lib.sfac_tune_import(text.encode())
    if k < 0:
        raise RuntimeError("sfac_tune_import failed: " + last_error())
    return k


def tune_clear():
    _, _, lib = describe_sfac()
    _check(lib.sfac_tune_clear(), "sfac_tune_clear")


class ServeRequest:
    """One submitted evaluation: holds the borrowed input arrays and the
    output arrays until ``SfacServer.wait`` returns them."""

    def __init__(self, rid, w, x, out, dx, phi):
        self.id = rid
        self._keep = (w, x)
        self.out, self.dx, self.phi = out, dx, phi
        self.timing = None


def _ptr(a):
    return None if a is None else a.ctypes.data


class SfacServer:
    """Continuous-batching serving loop over resident engines (native, in
    libsfac.so): requests of mixed atom counts, boxes, grids, orders and
    precisions are admitted by free worker slot and HBM budget, overlap on
    the device, and report their latency.  Each request is one evaluation
    (E, virial, dE/dx, phi) with the closed-form Ewald kernel of width
    ``eta`` and cutoff ``max_m``, its influence table built on the device."""

    _STATS = ("completed", "failed", "latency_p50_s", "latency_p99_s", "latency_mean_s",
              "latency_max_s", "wait_p50_s", "wait_p99_s", "resident_engines",
              "resident_bytes", "budget_bytes", "engines_created", "evictions",
              "batched_launches", "batched_frames", "max_running")

    def __init__(self, slots=4, hbm_budget=0, max_batch=8, batch_atoms=32768):
        _, _, lib = describe_sfac()
        self._lib = lib
        self._s = lib.sfac_server_create(int(slots), float(hbm_budget), int(max_batch),
                                         int(batch_atoms))
        if not self._s:
            raise RuntimeError("sfac_server_create failed: " + last_error())

    def submit(self, L, K, order, w, x, eta, max_m=1.0, precision="fp64", fixed=False,
               forces=True, potentials=True, out=None):
        """Queue one evaluation.  ``w`` (n) and ``x`` (n x 3) are float64 and
        stay referenced by the returned request; pinned host arrays are
        copied by DMA.  ``out`` = optional (E/virial[7], dx[n,3], phi[n])
        arrays to fill."""
        w = np.ascontiguousarray(w, dtype=np.float64)
        x = np.ascontiguousarray(x, dtype=np.float64)
        n = len(w)
        if x.shape != (n, 3):
            raise ValueError("x must have shape (n, 3)")
        flags = (SFAC_FP32 if precision == "fp32" else 0) | (SFAC_FIXED if fixed else 0)
        if precision not in ("fp32", "fp64"):
            raise ValueError("precision must be 'fp64' or 'fp32'")
        if out is None:
            o = np.zeros(7)
            dx = np.zeros((n, 3)) if forces else None
            phi = np.zeros(n) if potentials else None
        else:
            o, dx, phi = out
        L = np.ascontiguousarray(np.asarray(L, dtype=np.float64))
        Ki = np.ascontiguousarray(np.asarray(K, dtype=np.int32))
        rid = self._lib.sfac_server_submit(self._s, L.ctypes.data, Ki.ctypes.data, int(order),
                                           flags, float(eta), float(max_m), n, _ptr(w),
                                           _ptr(x), _ptr(o), _ptr(dx), _ptr(phi))
        if rid <= 0:
            raise RuntimeError("sfac_server_submit failed (%d): %s" % (rid, last_error()))
        return ServeRequest(rid, w, x, o, dx, phi)

    def wait(self, req):
        """Block until ``req`` ends; returns (E, vir, dx, phi)."""
        t = np.zeros(3)
        rc = self._lib.sfac_server_wait(self._s, req.id, t.ctypes.data)
        req.timing = {"wait_s": t[0], "service_s": t[1], "latency_s": t[2]}
        if rc != 0:
            raise RuntimeError("request %d failed (%d): %s" % (req.id, rc, last_error()))
        return req.out[0], req.out[1:7].copy(), req.dx, req.phi

    def poll(self, req):
        return self._lib.sfac_server_poll(self._s, req.id)

    def stats(self):
        out = np.zeros(16)
        _check(self._lib.sfac_server_stats(self._s, out.ctypes.data), "sfac_server_stats")
        return dict(zip(self._STATS, out.tolist()))

    def reset_stats(self):
        _check(self._lib.sfac_server_reset_stats(self._s), "sfac_server_reset_stats")

    def close(self):
        if getattr(self, "_s", None):
            self._lib.sfac_server_destroy(self._s)
            self._s = None

    def __del__(self):
        self.close()


# ---- the Ewald kernel (ewald.py:124-132) ------------------------------------
eta = 0.5
_fac = -(pi / eta) ** 2


def ewald_f(r2, ptr, info):
    """f(r2) = exp(-pi^2 r2/eta^2)/(pi r2); *ptr = f'(r2); 0 for r2 < 1e-10.

    Like the reference it reads the module-level ``_fac`` (set from ``eta``
    at import).  ``Sfac.set_A`` recognises this function and builds the same
    table natively with eta = pi/sqrt(-_fac)."""
    if r2 < 1e-10:
        ptr.contents.value = 0.0
        return 0.0
    f = exp(r2 * _fac) / (r2 * pi)
    ptr.contents.value = (_fac - 1. / r2) * f
    return f


def ewald_kernel(eta_):
    """rad_fn-compatible Ewald kernel with its own eta (e.g. eta = beta);
    recognised by ``Sfac.set_A`` for the native table build."""
    fac = -(pi / eta_) ** 2

    def f(r2, ptr, info):
        if r2 < 1e-10:
            ptr.contents.value = 0.0
            return 0.0
        v = exp(r2 * fac) / (r2 * pi)
        ptr.contents.value = (fac - 1. / r2) * v
        return v

    f._ewald_eta = float(eta_)
    return f


def _native_eta(f):
    if f is ewald_f:
        return pi / (-_fac) ** 0.5
    return getattr(f, "_ewald_eta", None)


def LofS(L):
    return np.array([[L[0], 0., 0.],
                     [L[3], L[1], 0.],
                     [L[4], L[5], L[2]]])


def SofL(L):
    return np.array([L[0, 0], L[1, 1], L[2, 2], L[1, 0], L[2, 0], L[2, 1]])
