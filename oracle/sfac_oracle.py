"""FP64 CPU oracle for the SPME reciprocal-space path.

TEST INFRASTRUCTURE ONLY.  This module is the checker that the HIP engine is
compared against.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import it; the product library never
calls it and has no CPU fallback.

It is a clean-room numpy restatement of the reference algorithm
(frobnitzem/EwaldCorrel, ``sfac.hpp`` / ``sfac.cpp`` / ``ewald.py``); every
function cites the reference lines it follows.  Index conventions, the ½
factors of the half-complex influence table, the ×2 planes of the
convolution and the virial assembly are reproduced exactly.  The 3-D FFT is
numpy's pocketfft, which computes the same unnormalised DFT as the
reference's FFTW r2c/c2r calls (``sfac.cpp:78-81,125,392``).

Pinning (see ``tests/test_oracle.py``):
  * cell / scale            -- closed-form inverse, L @ iL == I, wrap range
  * bspline_coeffs / eval   -- exact rational derivation vs M_4, M_6 closed
                               forms, partition of unity, derivative by FD
  * bspl_dft                -- direct complex DFT of the node values
  * spread + forward        -- golden Re S(0,0,kz) and line errors of
                               ``test.log.ref:2-6``; S(0) == sum(w) exactly;
                               brute-force O(G^2) DFT on a tiny prime grid
  * set_A                   -- brute-force per-point image sum
  * en / de1                -- golden E of ``test.log.ref:7-8``; en == de1
  * de2                     -- golden dx of ``test.log.ref:9-12``; central FD
  * virial                  -- golden virial ``test.log.ref:17-18``; FD of E(L)
  * potl                    -- phi == dE/dw exactly (E quadratic in w);
                               E == 1/2 w.phi
"""
from __future__ import annotations

from fractions import Fraction
from math import comb, factorial, pi

import numpy as np

__all__ = [
    "Cell", "bspline_coeffs", "bspl_eval", "bspl_deval", "bspl_dft",
    "OracleSfac", "ewald_rad", "LofS", "SofL",
]


# --------------------------------------------------------------------------
# Geometry  (sfac.hpp:33-63)
# --------------------------------------------------------------------------
class Cell:
    """Lower-triangular box from L6 = [a_x, b_y, c_z, b_x, c_x, c_y].

    Rows of ``L`` are the box vectors; ``iL`` is its explicit inverse and
    ``iV = 1/(a_x b_y c_z)`` (sfac.hpp:37-52).
    """

    def __init__(self, L6):
        L6 = np.asarray(L6, dtype=np.float64)
        self.L6 = L6.copy()
        a, b, c, bx, cx, cy = L6
        L = np.zeros((3, 3))
        L[0, 0], L[1, 1], L[2, 2] = a, b, c
        L[1, 0], L[2, 0], L[2, 1] = bx, cx, cy
        iV = 1.0 / (a * b * c)
        iL = np.zeros((3, 3))
        iL[0, 0], iL[1, 1], iL[2, 2] = 1.0 / a, 1.0 / b, 1.0 / c
        iL[1, 0] = -bx * c * iV
        iL[2, 0] = (bx * cy - b * cx) * iV
        iL[2, 1] = -a * cy * iV
        self.L, self.iL, self.iV = L, iL, iV

    def scale(self, x):
        """Fractional coordinates wrapped to [0,1) (sfac.hpp:56-62).

        s_0 = x.iL[:,0], s_1 = x1 iL11 + x2 iL21, s_2 = x2 iL22, each
        reduced by ``s - floor(s)``; the expression order follows the
        reference so that the rounding is the same.
        """
        x = np.asarray(x, dtype=np.float64).reshape(-1, 3)
        iL = self.iL
        s0 = x[:, 0] * iL[0, 0] + x[:, 1] * iL[1, 0] + x[:, 2] * iL[2, 0]
        s1 = x[:, 1] * iL[1, 1] + x[:, 2] * iL[2, 1]
        s2 = x[:, 2] * iL[2, 2]
        s = np.stack([s0, s1, s2], axis=1)
        return s - np.floor(s)


def LofS(L):
    """6-vector -> lower-triangular 3x3 (ewald.py:134-138)."""
    return np.array([[L[0], 0.0, 0.0], [L[3], L[1], 0.0], [L[4], L[5], L[2]]])


def SofL(L):
    """lower-triangular 3x3 -> 6-vector (ewald.py:139-140)."""
    return np.array([L[0, 0], L[1, 1], L[2, 2], L[1, 0], L[2, 0], L[2, 1]])


# --------------------------------------------------------------------------
# Cardinal B-splines  (sfac.hpp:65-158)
# --------------------------------------------------------------------------
def bspline_coeffs(n: int) -> np.ndarray:
    """n x n matrix; row j holds the coefficients of M_n(x+j), x in [0,1),
    highest power first (the layout of sfac.hpp:81-109).

    Derived here from M_n(y) = 1/(n-1)! sum_k (-1)^k C(n,k) (y-k)_+^(n-1):
    for y = x + j with x in [0,1) only k <= j contributes.  Exact rational
    arithmetic, rounded once to double.
    """
    M = [[Fraction(0)] * n for _ in range(n)]
    for j in range(n):
        for k in range(j + 1):
            c = Fraction((-1) ** k * comb(n, k), factorial(n - 1))
            a = j - k
            for p in range(n):  # (x + a)^(n-1) = sum_p C(n-1,p) x^p a^(n-1-p)
                M[j][n - 1 - p] += c * comb(n - 1, p) * Fraction(a) ** (n - 1 - p)
    return np.array([[float(v) for v in row] for row in M])


def bspl_eval(M: np.ndarray, x, j):
    """Horner value of M_n(x+j) (sfac.hpp:124-131); x, j broadcast."""
    n = M.shape[0]
    x = np.asarray(x, dtype=np.float64)
    b = M[j, 0] * np.ones_like(x)
    for i in range(1, n):
        b = b * x + M[j, i]
    return b


def bspl_deval(M: np.ndarray, x, j):
    """Horner value and derivative of M_n(x+j) (sfac.hpp:146-157)."""
    n = M.shape[0]
    x = np.asarray(x, dtype=np.float64)
    db = M[j, 0] * np.ones_like(x)
    b = M[j, 0] * x + M[j, 1]
    for i in range(2, n):
        db = db * x + b
        b = b * x + M[j, i]
    return b, db


def bspl_dft(K: int, order: int, M: np.ndarray) -> np.ndarray:
    """iB[i] = 1 / sum_{j=1}^{n-1} cos(2 pi i (j - n/2) / K) M_n(j)
    (sfac.cpp:21-35, node values = last column of M, sfac.cpp:65-68)."""
    i = np.arange(K, dtype=np.float64)
    om = -(2.0 * pi * i) / K
    x = np.zeros(K)
    j0 = order // 2
    for j in range(1, order):
        x += np.cos(om * (j - j0)) * M[j, order - 1]
    return 1.0 / x


# --------------------------------------------------------------------------
# Radial kernels  (ewald.py:124-132)
# --------------------------------------------------------------------------
def ewald_rad(eta: float):
    """Vectorised Ewald kernel f(z) = exp(-pi^2 z/eta^2)/(pi z) and
    f'(z) = (-pi^2/eta^2 - 1/z) f, both 0 for z < 1e-10 (ewald.py:124-132)."""
    fac = -(pi / eta) ** 2

    def f(r2):
        r2 = np.asarray(r2, dtype=np.float64)
        ok = r2 >= 1e-10
        z = np.where(ok, r2, 1.0)
        val = np.exp(z * fac) / (z * pi)
        der = (fac - 1.0 / z) * val
        return np.where(ok, val, 0.0), np.where(ok, der, 0.0)

    return f


# --------------------------------------------------------------------------
# Engine  (sfac.hpp:163-192, sfac.cpp:39-545)
# --------------------------------------------------------------------------
class OracleSfac:
    """FP64 restatement of ``struct SFac`` and its C interface.

    ``Q`` is kept as a real K0 x K1 x K2 array and ``FQ`` as the half-complex
    K0 x K1 x ldim array, instead of the reference's in-place padded buffer.
    The state machine (0 = S-hat, 1 = real-space convolution) follows
    sfac.cpp:126,393 but raises instead of printing.
    """

    def __init__(self, L6, K, order=4):
        # sfac.cpp:39-56
        if order % 2 != 0:
            raise ValueError("invalid order")
        K = [int(k) for k in K]
        if min(K) < 1:
            raise ValueError("invalid K")
        self.order = order
        self.K = K
        self.ldim = K[2] // 2 + 1
        self.cell = Cell(L6)
        self.M = bspline_coeffs(order)
        self.iB = [bspl_dft(K[d], order, self.M) for d in range(3)]
        self.A = None
        self.dA = None
        self.state = None
        self.FQ = None
        self.Q = None

    # ---- set_L / set_A --------------------------------------------------
    def set_L(self, L6):
        """Replace the cell only (sfac.cpp:199-202)."""
        self.cell = Cell(L6)

    def set_A(self, rad, max_m):
        """Influence table over the ball |m| <= max_m (sfac.cpp:213-292).

        ``rad(r2) -> (f, f')`` is vectorised.  Bounds of the i/j/k loops,
        the aliasing of every image onto the grid and the folding of the
        k >= ldim half onto the conjugate point are those of the reference;
        A receives 1/2 f and dA the full f' m_a m_b.
        """
        K0, K1, K2 = self.K
        ldim = self.ldim
        iL = self.cell.iL
        R2 = max_m * max_m
        i0 = int(max_m / iL[0, 0])
        idx_all, m2_all, mm_all = [], [], []
        for i in range(-i0, i0 + 1):
            mx = i * iL[0, 0]
            m20 = mx * mx
            dy = np.sqrt(R2 - m20)
            my0 = i * iL[1, 0]
            j0 = int(np.ceil((-dy - my0) / iL[1, 1]))
            j1 = int(np.floor((dy - my0) / iL[1, 1]))
            if j1 < j0:
                continue
            js = np.arange(j0, j1 + 1)
            my = my0 + js * iL[1, 1]
            m21 = m20 + my * my
            dz = np.sqrt(R2 - m21)
            mz0 = i * iL[2, 0] + js * iL[2, 1]
            k0 = np.ceil((-dz - mz0) / iL[2, 2]).astype(np.int64)
            k1 = np.floor((dz - mz0) / iL[2, 2]).astype(np.int64)
            cnt = np.maximum(k1 - k0 + 1, 0)
            tot = int(cnt.sum())
            if tot == 0:
                continue
            row = np.repeat(np.arange(len(js)), cnt)
            start = np.repeat(np.cumsum(cnt) - cnt, cnt)
            ks = k0[row] + (np.arange(tot) - start)
            jr = js[row]
            myr = my[row]
            mz = mz0[row] + ks * iL[2, 2]
            m2 = m21[row] + mz * mz
            n0 = i % K0
            n1 = jr % K1
            n2 = ks % K2
            fold = n2 >= ldim
            ix = np.where(fold, (-i) % K0, n0)
            iy = np.where(fold, (-jr) % K1, n1)
            iz = np.where(fold, K2 - n2, n2)
            idx_all.append((ix * K1 + iy) * ldim + iz)
            m2_all.append(m2)
            mm_all.append(np.stack([np.full(tot, mx), myr, mz], axis=1))
        H = K0 * K1 * ldim
        A = np.zeros(H)
        dA = np.zeros((H, 6))
        if idx_all:
            idx = np.concatenate(idx_all)
            m2 = np.concatenate(m2_all)
            mm = np.concatenate(mm_all)
            f, fp = rad(m2)
            A += np.bincount(idx, weights=0.5 * f, minlength=H)
            mx, my, mz = mm[:, 0], mm[:, 1], mm[:, 2]
            for c, prod in enumerate((mx * mx, my * my, mz * mz,
                                      my * mx, mz * mx, mz * my)):
                dA[:, c] += np.bincount(idx, weights=fp * prod, minlength=H)
        self.A = A.reshape(K0, K1, ldim)
        self.dA = dA.reshape(K0, K1, ldim, 6)
        return 0

    # ---- sfac: sort + spread + r2c + /B  (sfac.cpp:92-139) ---------------
    def cells(self, x):
        """u = scale(x) K, c = (int)u, frac = u - c (sfac.cpp:97-104).

        A coordinate that wraps to exactly 1.0 gives c = K; the reference
        then drops or mis-bins the atom.  Here c is reduced mod K (the
        periodic image), which is also what the gather path (ceil) sees.
        """
        u = self.cell.scale(x) * np.asarray(self.K, dtype=np.float64)
        c = u.astype(np.int64)
        frac = u - c
        c = c % np.asarray(self.K)
        return c, frac

    def spread(self, w, x):
        """Real grid Q (K0,K1,K2).  Scatter form of accum_spline
        (sfac.cpp:141-168): an atom in cell c with fraction f adds
        w prod_d M(f_d + i_d) to grid (c_d + n/2 - i_d) mod K_d."""
        K0, K1, K2 = self.K
        n = self.order
        w = np.asarray(w, dtype=np.float64)
        c, frac = self.cells(x)
        Mx = [bspl_eval(self.M, frac[:, d][:, None], np.arange(n)[None, :])
              for d in range(3)]  # (N, n) each
        gx = (c[:, 0][:, None] + n // 2 - np.arange(n)[None, :]) % K0
        gy = (c[:, 1][:, None] + n // 2 - np.arange(n)[None, :]) % K1
        gz = (c[:, 2][:, None] + n // 2 - np.arange(n)[None, :]) % K2
        val = (w[:, None, None, None] * Mx[0][:, :, None, None]
               * Mx[1][:, None, :, None] * Mx[2][:, None, None, :])
        idx = ((gx[:, :, None, None] * K1 + gy[:, None, :, None]) * K2
               + gz[:, None, None, :])
        Q = np.bincount(idx.ravel(), weights=val.ravel(), minlength=K0 * K1 * K2)
        return Q.reshape(K0, K1, K2)

    def iB3(self):
        K0, K1, K2 = self.K
        return (self.iB[0][:, None, None] * self.iB[1][None, :, None]
                * self.iB[2][None, None, :self.ldim])

    def sfac(self, w, x):
        """S-hat = r2c(Q) / (b0 b1 b2) (sfac.cpp:113-139); state 0."""
        Q = self.spread(w, x)
        FQ = np.fft.rfftn(Q)  # sign -1, unnormalised (FFTW r2c convention)
        self.FQ = FQ * self.iB3()
        self.Q = None
        self.state = 0

    def get_S(self):
        """Complex K0 x K1 x ldim copy of S-hat (ewald.py:97-101)."""
        if self.state != 0:
            raise RuntimeError("get_S: grid holds the real-space convolution")
        return self.FQ.copy()

    # ---- en / de1  (sfac.cpp:304-396) -------------------------------------
    def en(self):
        if self.state != 0:
            raise RuntimeError("You must call sfac before this routine.")
        if self.A is None:
            raise RuntimeError("No convolution kernel has been set.")
        s = self.FQ.real ** 2 + self.FQ.imag ** 2
        return float(np.sum(self.A * s) * self.cell.iV)

    def de1(self):
        """Energy, virial (xx,yy,zz,yx,zx,zy) and the convolved real grid.

        FQ <- S-hat A iB0 iB1 iB2, doubled on kz=0 and (even K2) kz=K2/2
        (sfac.cpp:361-382); vir = (sum |S|^2 dA + delta sum A|S|^2) iV^2
        (sfac.cpp:384-390); unnormalised c2r (sfac.cpp:392); state 1.
        """
        if self.state != 0:
            raise RuntimeError("You must call sfac before this routine.")
        if self.A is None:
            raise RuntimeError("No convolution kernel has been set.")
        K0, K1, K2 = self.K
        s = self.FQ.real ** 2 + self.FQ.imag ** 2
        e_raw = float(np.sum(self.A * s))
        v = np.einsum("ijk,ijkc->c", s, self.dA)
        fac = self.cell.iV ** 2
        vir = np.array([v[0] + e_raw, v[1] + e_raw, v[2] + e_raw,
                        v[3], v[4], v[5]]) * fac
        t = self.A * self.iB3()
        t[:, :, 0] *= 2.0
        if K2 % 2 == 0:
            t[:, :, self.ldim - 1] *= 2.0
        conv = self.FQ * t
        self.Q = np.fft.irfftn(conv, s=(K0, K1, K2), axes=(0, 1, 2), norm="forward")
        self.FQ = None
        self.state = 1
        return e_raw * self.cell.iV, vir

    # ---- de2 / potl  (sfac.cpp:409-545) -----------------------------------
    def _stencil(self, x):
        """k1 = ceil(u), t = k1 - u, grid (k1 - n/2 + i) mod K
        (sfac.cpp:428-439,514-525)."""
        u = self.cell.scale(x) * np.asarray(self.K, dtype=np.float64)
        k1 = np.ceil(u)
        t = k1 - u
        n = self.order
        ar = np.arange(n)
        g = [((k1[:, d].astype(np.int64) - n // 2)[:, None] + ar[None, :])
             % self.K[d] for d in range(3)]
        return t, g

    def de2(self, w, x):
        """dE/dx, N x 3 (the gradient; force = -dx) (sfac.cpp:409-483)."""
        if self.state != 1:
            raise RuntimeError("You must call de1 before this routine.")
        w = np.asarray(w, dtype=np.float64)
        n = self.order
        t, g = self._stencil(x)
        ar = np.arange(n)[None, :]
        b = []
        db = []
        for d in range(3):
            bv, dbv = bspl_deval(self.M, t[:, d][:, None], ar)
            b.append(bv)
            db.append(dbv)
        K0, K1, K2 = self.K
        Qg = self.Q[g[0][:, :, None, None], g[1][:, None, :, None],
                    g[2][:, None, None, :]]  # (N,n,n,n)
        d0 = np.einsum("ai,aj,ak,aijk->a", db[0], b[1], b[2], Qg)
        d1 = np.einsum("ai,aj,ak,aijk->a", b[0], db[1], b[2], Qg)
        d2 = np.einsum("ai,aj,ak,aijk->a", b[0], b[1], db[2], Qg)
        iV = self.cell.iV
        d0 = d0 * (-w * iV) * K0
        d1 = d1 * (-w * iV) * K1
        d2 = d2 * (-w * iV) * K2
        iL = self.cell.iL
        dx = np.empty((len(w), 3))
        dx[:, 0] = iL[0, 0] * d0
        dx[:, 1] = iL[1, 0] * d0 + iL[1, 1] * d1
        dx[:, 2] = iL[2, 0] * d0 + iL[2, 1] * d1 + iL[2, 2] * d2
        return dx

    def potl(self, x):
        """phi_a = iV sum prod M Q = dE/dw_a (sfac.cpp:497-545)."""
        if self.state != 1:
            raise RuntimeError("You must call de1 before this routine.")
        n = self.order
        t, g = self._stencil(x)
        ar = np.arange(n)[None, :]
        b = [bspl_eval(self.M, t[:, d][:, None], ar) for d in range(3)]
        Qg = self.Q[g[0][:, :, None, None], g[1][:, None, :, None],
                    g[2][:, None, None, :]]
        return np.einsum("ai,aj,ak,aijk->a", b[0], b[1], b[2], Qg) * self.cell.iV

    # ---- convenience: one evaluation = sfac + de1 + de2 + potl ------------
    def evaluate(self, w, x):
        self.sfac(w, x)
        E, vir = self.de1()
        dx = self.de2(w, x)
        phi = self.potl(x)
        return E, vir, dx, phi


# ---- real-space Ewald and corrections (notes.tex:73-79, test.py:83,126-132) --------------
def real_space(L6, w, x, eta, rc):
    """Real-space Ewald sum over every periodic image within the cutoff rc,
    with the self and neutralising-background corrections, by brute force
    (small systems; the checker of the device real-space kernel).

    The reciprocal part with f(z) = exp(-pi^2 z/eta^2)/(pi z) is completed
    by (CGS units, notes.tex:66-79)
      E_pair = 1/2 sum_i sum_{j, n}' w_i w_j erfc(eta r)/r,  r = |x_i - x_j + L^T n|
      E_self = -eta/sqrt(pi) sum_i w_i^2                    (test.py:83)
      E_bg   = -pi (sum_i w_i)^2 / (2 V eta^2)
    Returns dict(E=E_pair, vir (xx,yy,zz,yx,zx,zy) of the pairs with the
    convention of de1 (-(1/V) sum U'(r)/r r_a r_b), dx = dE_pair/dx,
    phi = dE/dw_i of all three terms, E_self, E_bg, vir_bg = E_bg / V on the
    diagonal)."""
    from math import erfc, exp, sqrt
    cell = Cell(L6)
    Lm = LofS(np.asarray(L6, dtype=float))
    iL = cell.iL
    V = 1.0 / cell.iV
    n = len(w)
    x = np.asarray(x, dtype=float)
    w = np.asarray(w, dtype=float)
    # image range: |s_d + n_d| <= rc |iL[:, d]| for every image within rc
    ext = [int(np.ceil(rc * np.linalg.norm(iL[:, d]))) + 1 for d in range(3)]
    E = 0.0
    vir = np.zeros((3, 3))
    dx = np.zeros((n, 3))
    phi = np.zeros(n)
    c2 = 2.0 * eta / sqrt(pi)
    for i in range(n):
        d = x[i] - x
        s = d @ iL
        s -= np.round(s)
        for a in range(-ext[0], ext[0] + 1):
            for b in range(-ext[1], ext[1] + 1):
                for c in range(-ext[2], ext[2] + 1):
                    r = (s + np.array([a, b, c])) @ Lm
                    r2 = np.einsum("ij,ij->i", r, r)
                    for j in np.nonzero((r2 < rc * rc) & (r2 > 1e-20))[0]:
                        rr = sqrt(r2[j])
                        g = erfc(eta * rr) / rr
                        phi[i] += w[j] * g
                        u = w[i] * w[j] * g
                        up = w[i] * w[j] * (-g / rr - c2 * exp(-eta * eta * r2[j]) / rr)
                        E += 0.5 * u
                        dx[i] += up / rr * r[j]
                        vir += 0.5 * (up / rr) * np.outer(r[j], r[j])
    vir = -vir / V
    Q = w.sum()
    E_self = -eta / sqrt(pi) * np.sum(w * w)
    E_bg = -pi * Q * Q / (2.0 * V * eta * eta)
    phi += -c2 * w - pi * Q / (V * eta * eta)
    return dict(E=E, vir=np.array([vir[0, 0], vir[1, 1], vir[2, 2], vir[1, 0], vir[2, 0],
                                   vir[2, 1]]),
                dx=dx, phi=phi, E_self=E_self, E_bg=E_bg, vir_bg=E_bg / V)


def plot_radial(f, L):
    """|k|-shell averages of a (K0, K1, ldim) array, restating plot.py:9-32:
    along each axis of the array (whatever its length n, ldim for z) index i
    stands for i - n when i >= n // 2 (plot.py:3-7), k_d = 2 pi i / L with one
    box length L for all axes, kmax = K0 pi / (4 L) and dk = kmax / 5000; a
    point's bin is the integer part of |k| / dk clipped to kmax / dk + 1.
    Returns (k, mean, count) of the non-empty bins from 1 to one below the
    last non-empty bin (plot.py drops bin 0 and the last bin of its count)."""
    f = np.asarray(f, dtype=np.float64)

    def axis(n):
        i = np.arange(n)
        i = np.where(i >= n // 2, i - n, i)
        return i * np.pi * 2 / L

    kmax = 0.25 * f.shape[0] * np.pi / L
    dk = kmax / 5000
    kx, ky, kz = (axis(n) for n in f.shape)
    s2 = (kx * kx)[:, None, None] + (ky * ky)[None, :, None] + (kz * kz)[None, None, :]
    b = np.minimum(np.sqrt(s2) / dk, kmax / dk + 1).astype(int).ravel()
    cnt = np.bincount(b)
    tot = np.bincount(b, weights=f.ravel())
    return [(i * dk, tot[i] / cnt[i], float(cnt[i]))
            for i in range(1, len(cnt) - 1) if cnt[i] > 0]
