/* sfac_ext.h -- extensions of the libsfac.so C ABI (plain C, no torch or
 * HIP types in the signatures).  None of these has a counterpart in the
 * reference; each states the reference call(s) it stands beside.
 *
 * Conventions: handles come from sfac_ctor / sfac_ctor_ex.  Functions that
 * return int return 0 on success and a negative code on failure, with the
 * message available from sfac_last_error().  "_dev" functions take device
 * (HBM) pointers, run asynchronously on the handle's HIP stream and return
 * without synchronising; results are read after sfac_sync().
 */
#ifndef EWALDCORREL_AMD_SFAC_EXT_H
#define EWALDCORREL_AMD_SFAC_EXT_H

#ifdef __cplusplus
extern "C" {
#endif

/* precision / accumulation flags for sfac_ctor_ex */
#define SFAC_FP64 0  /* everything in FP64 (the reference's arithmetic)        */
#define SFAC_FP32 1  /* spread/FFT/gather FP32; table, reductions FP64; int64   */
                     /* fixed-point grid accumulation                          */
#define SFAC_FIXED 2 /* deterministic int64 fixed-point spreading (FP64 path)   */
#define SFAC_FP32_INFLUENCE 4 /* with SFAC_FP32: the closed-form Ewald influence   */
                              /* and the per-thread energy/virial sums in FP32    */
                              /* (faster; by default they are FP64, as are /B,    */
                              /* the table and every reduction)                   */

#define SFAC_ERR_USAGE (-1)  /* call out of order (reference prints an error) */
#define SFAC_ERR_ARG (-2)    /* invalid argument                                */
#define SFAC_ERR_DEVICE (-3) /* HIP runtime error / out of device memory        */

/* sfac_ctor (sfac.hpp:196) with a precision flag; NULL on failure. */
void *sfac_ctor_ex(const double L[6], const int K[3], int order, int flags);

/* message of the last failed call on this thread ("" if none) */
const char *sfac_last_error(void);

/* set_A (sfac.cpp:213-292) for the closed-form Ewald kernel
 * f(z) = exp(-pi^2 z/eta^2)/(pi z), f'(z) = (-pi^2/eta^2 - 1/z) f
 * (ewald.py:124-132), built without a per-point callback. */
int sfac_set_A_ewald(void *h, double eta, double max_m);

/* Copy of the current influence table (the reference keeps it in SFac::A,
 * SFac::dA, sfac.hpp:184): A[H] and dA[6H], H = K0 K1 (K2/2+1); either
 * pointer may be NULL.  After sfac_set_A_ewald the table is built on the
 * device and SFac::A is NULL. */
int sfac_table(void *h, double *A, double *dA);

/* Device-pointer forms of sfac / en / de1 / de2 / potl (sfac.cpp:190-545).
 * en_dev writes E to out_d[0]; de1_dev writes E to out_d[0] and the virial
 * (xx,yy,zz,yx,zx,zy) to out_d[1..6]. */
int sfac_dev(void *h, int n, const double *w_d, const double *x_d);
int en_dev(void *h, double *out_d);
int de1_dev(void *h, double *out_d);
int de2_dev(void *h, int n, const double *w_d, const double *x_d, double *dx_d);
int potl_dev(void *h, int n, const double *x_d, double *phi_d);

/* One evaluation = sfac + de1 + de2 + potl on the same atoms.
 * dx_d or phi_d may be NULL to skip that output. */
int sfac_eval_dev(void *h, int n, const double *w_d, const double *x_d, double *out_d,
                  double *dx_d, double *phi_d);
/* sfac_eval_dev with no device allocation, for callers that cannot allocate
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc); no counterpart in the reference.  The results are those of
 * sfac_eval_dev.  The per-atom sort workspace lives in the library's static
 * scratch array of SFAC_NOALLOC_SCRATCH_BYTES per GPU (48 B per atom in FP32,
 * 96 B in FP64, plus K0 + 1 ints per 3,072 FP32 or 1,024 FP64 atoms: about
 * 1.39 M atoms in FP32 or 0.69 M in FP64 on a 128-point x axis), shared by
 * every handle on that GPU:
 * calls that overlap on one GPU race.  Kernels are launched one by one (no
 * graph capture or instantiation).  Returns SFAC_ERR_ARG, with the limit in
 * the message, when n atoms need more scratch than that. */
#define SFAC_NOALLOC_SCRATCH_BYTES (64ull << 20)
int sfac_eval_dev_noalloc(void *h, int n, const double *w_d, const double *x_d, double *out_d,
                          double *dx_d, double *phi_d);
/* host-pointer form; returns E, fills vir[6], dx[3n], phi[n] */
double sfac_eval(void *h, int n, const double *w, const double *x, double vir[6],
                 double *dx, double *phi);

/* the handle's HIP stream (hipStream_t) and a blocking wait on it */
void *sfac_stream(void *h);
int sfac_sync(void *h);

/* out[0..11] = K0, K1, K2, ldim, order, flags, device bytes held, state,
 * influence-table mode (0 table in HBM, 1 closed-form Ewald from the
 * indices), graph replay enabled, y-z slab transforms in use (0 line
 * passes, 1 slab transforms after spreading tiles, 2 slab transforms with
 * the spreading fused into the forward one), slab transforms available for
 * this grid */
int sfac_info(void *h, long long out[12]);

/* Transform path: 0 (default) the y-z slab transforms when the plane fits
 * in LDS and K1, K2 factor into 2, 3, 5, with the spreading fused into the
 * forward slab transform where a plane of fixed-point sums fits in LDS as
 * well; 1 separate z and y line passes; 2 the slab transforms after separate
 * spreading tiles. */
int sfac_set_transform(void *h, int mode);

/* sfac_eval_dev captures its launches in a hipGraph and replays it while
 * the arguments stay the same (default on); 0 launches kernel by kernel. */
int sfac_set_graph(void *h, int enable);

/* Batches: frames (handles with one precision, accumulation mode and spline
 * order, each with its influence function set) evaluated together, one
 * launch per stage for all frames, replayed from a captured hipGraph while
 * the per-frame arguments stay the same.  sfac_batch_set gives frame f its
 * device pointers (out_d receives E and the virial as in de1_dev);
 * sfac_batch_eval runs one evaluation of every frame on `stream` (a
 * hipStream_t; NULL selects the stream of the first frame's handle, which
 * sfac_batch_sync waits for).  No counterpart in the reference, which
 * evaluates one system per call (sfac.cpp:190-545). */
void *sfac_batch_create(void **handles, int nframes);
int sfac_batch_set(void *b, int f, int n, const double *w_d, const double *x_d, double *out_d,
                   double *dx_d, double *phi_d);
int sfac_batch_eval(void *b, void *stream);
int sfac_batch_sync(void *b);
int sfac_batch_set_graph(void *b, int enable);
int sfac_batch_destroy(void *b);

/* sfac_batch_set_lanes: the batch's frames as `lanes` groups of consecutive
 * frames (1 to 4, at most one group per frame), each replayed on its own
 * stream: the first group on the evaluation's stream, group g on the stream
 * of its first frame's handle, forked from and joined back into the
 * evaluation's stream by events, so that one group's kernels fill the CUs
 * the other leaves idle.  Each frame's results are those of one group; 1
 * (the default) launches each stage once for every frame.
 * sfac_batch_autotune tries the other of 1 and 2 and keeps it when faster.
 * sfac_batch_lanes returns the current count (or an error code).
 * sfac_batch_sync waits for every group's stream.  No counterpart in the
 * reference. */
int sfac_batch_set_lanes(void *b, int lanes);
int sfac_batch_lanes(void *b);

/* Structure-factor accumulation on the device, asynchronous on the
 * handle's stream (the callers of get_S: run.py:36-47 sums |S|^2 over
 * frames, gmx_ewald.py:78-107 sums S_i conj(S_j) over species pairs,
 * plot.py:9-32 averages over |k| shells).  Arrays are FP64 device memory
 * owned by the caller, indexed like the half-complex grid (H points).
 *   sfac_accum_S2:    after sfac, acc_d[h] += |S_hat_h|^2
 *   sfac_accum_cross: after sfac on both handles (same grid and precision),
 *                     acc_d[2h], acc_d[2h+1] += Re, Im of S_a conj(S_b)
 *   sfac_radial:      sum_d[b] += sum of val_d[h] over the points with
 *                     b = floor(|k| / dk) < nbins, cnt_d[b] += their number;
 *                     k = 2 pi m, m from the signed indices (K/2 positive)
 *                     and kz in [0, ldim), in the handle's current box */
int sfac_accum_S2(void *h, double *acc_d);
int sfac_accum_cross(void *ha, void *hb, double *acc_d);
int sfac_radial(void *h, const double *val_d, int nbins, double dk, double *sum_d,
                double *cnt_d);
/* sfac_radial with plot.py's |k| shells (plot.py:9-32), for the array run.py
 * saves (K0 x K1 x ldim): along each axis of that array index i stands for
 * i - n when i >= n/2 (n = K0, K1 or ldim), k_d = 2 pi i / L_d with the box's
 * axis lengths L6[0..2] (plot.py assumes a cubic box of side L), and point h
 * goes to bin b = (int) min(|k| / dk, clip); points with b >= nbins are left
 * out.  plot.py's own output takes kmax = K0 pi / (4 L), dk = kmax / 5000,
 * clip = kmax / dk + 1, nbins = (int) clip + 1, and reports the non-empty
 * bins 1 .. m - 1 (m the last non-empty one) as k = b dk, sum / count, count. */
int sfac_radial_plot(void *h, const double *val_d, double dk, double clip, int nbins,
                     double *sum_d, double *cnt_d);

/* Real-space Ewald part and corrections (notes.tex:66-79, test.py:83,
 * 126-132), which with en/de1 for f(z) = exp(-pi^2 z/eta^2)/(pi z) give the
 * total Coulomb energy of the periodic system, independent of eta:
 *   out[0]    E_pair = 1/2 sum_i sum_{j,n}' w_i w_j erfc(eta r)/r over every
 *             periodic image with r < rc
 *   out[1..6] its virial (xx,yy,zz,yx,zx,zy) in de1's convention
 *   out[7]    E_self = -eta/sqrt(pi) sum w^2
 *   out[8]    E_bg = -pi (sum w)^2 / (2 V eta^2) (neutralising background)
 *   out[9]    E_bg / V, the background's virial on each diagonal component
 *   dx        dE_pair/dx (NULL: not computed);  phi  dE/dw of all three terms
 * FP64 pair sum in the handle's current box, O(n^2) (for checks and moderate
 * n).  sfac_real_space takes host pointers and returns when done; the _dev
 * form takes device pointers and is asynchronous on the handle's stream. */
int sfac_real_space(void *h, int n, const double *w, const double *x, double eta, double rc,
                    double out[10], double *dx, double *phi);
int sfac_real_space_dev(void *h, int n, const double *w_d, const double *x_d, double eta,
                        double rc, double *out_d, double *dx_d, double *phi_d);

/* Serving loop: continuous batching of independent evaluations (each like
 * sfac + de1 + de2 + potl after set_A with the closed-form Ewald kernel, on
 * its own box, grid, order and precision) on one GPU.  No counterpart in the
 * reference, whose callers evaluate one system at a time
 * (run.py:34-45, gmx_ewald.py:76-89).
 *   sfac_server_create: `slots` worker threads, each with its own HIP
 *     streams, admit queued requests as soon as they are free and the
 *     request's engine fits `hbm_budget` bytes of device memory (0: 80 % of
 *     the free HBM).  Engines (grid, FFT plans, table, sort workspace) stay
 *     resident per (K, order, flags) and are evicted least recently used.
 *     A worker that admits a request of at most `batch_atoms` atoms (0:
 *     32768) also takes up to `max_batch` - 1 more queued small requests
 *     whose idle engines run the same kernel variants, and evaluates them as
 *     one batch; each result is the one the request gets alone.
 *   sfac_server_submit: queue one request (flags: SFAC_FP32 or SFAC_FIXED as
 *     for sfac_ctor_ex; eta, max_m: the Ewald kernel and cutoff of
 *     sfac_set_A_ewald).  Host buffers are borrowed until sfac_server_wait
 *     returns: w[n], x[3n] read; out[7] = E and the virial
 *     (xx,yy,zz,yx,zx,zy), dx[3n] = dE/dx and phi[n] written (dx, phi may
 *     be NULL).  Pinned host memory is copied by DMA directly.  Returns the
 *     request id (> 0) or a negative error code.
 *   sfac_server_wait: block until the request ends and release it; timing
 *     (may be NULL) = queue wait, service and total seconds.  Returns 0 or
 *     the request's error code (message in sfac_last_error).
 *   sfac_server_poll: 1 done, 0 pending, < 0 failed.
 *   sfac_server_stats: out[16] = completed, failed, latency p50, p99, mean,
 *     max (s), queue wait p50, p99 (s), resident engines, HBM bytes they
 *     hold, budget, engines created, evictions, batched launches, frames in
 *     them, most requests running at once (since create or reset). */
void *sfac_server_create(int slots, double hbm_budget, int max_batch, int batch_atoms);
long long sfac_server_submit(void *s, const double L6[6], const int K[3], int order, int flags,
                             double eta, double max_m, int n, const double *w, const double *x,
                             double *out, double *dx, double *phi);
int sfac_server_wait(void *s, long long id, double timing[3]);
int sfac_server_poll(void *s, long long id);
int sfac_server_stats(void *s, double out[16]);
int sfac_server_reset_stats(void *s);
int sfac_server_destroy(void *s);

/* Autotuner of launch shapes.  No counterpart in the reference, whose FFT
 * plans are chosen by FFTW_PATIENT planning in the constructor
 * (sfac.cpp:78-81); this tunes the transform path (y-z slabs or line
 * passes), the radix order of each axis's FFT, the spreading tile and the
 * gather tile by timing whole evaluations on the device, one knob at a time,
 * keeping a change only if it is 1 % faster.  The result applies to the
 * handle(s) and is remembered for the process: handles (and serving engines)
 * made later with the same grid, order and flags start from it.  Outputs of
 * the deterministic modes are unchanged by the spreading and gather tiles;
 * a radix order changes only the FFT's rounding.
 *   sfac_autotune: times single-handle evaluations of n atoms at device
 *     pointers w_d, x_d (both NULL: n synthetic atoms uniform in the box,
 *     alternating unit charges); the influence function must be set.  The
 *     grid is overwritten: call sfac again before en/de1 (state reset).
 *   sfac_batch_autotune: times evaluations of the whole batch (after
 *     sfac_batch_set of every frame); frames with the same grid, order and
 *     flags share one shape.
 *   Both write a report (one line per candidate) into report[report_len]
 *   when not NULL; `reps` timed evaluations per candidate (median).
 *   sfac_tune_export: the remembered shapes as text (one line per key);
 *     returns the bytes needed including the terminating NUL.
 *   sfac_tune_import: adds/replaces remembered shapes from such text;
 *     returns the number of lines read.  sfac_tune_clear forgets them.
 *   sfac_tune_current: the handle's current shapes as one such line
 *     (returns the bytes needed including the NUL).
 * Line format: "K0xK1xK2/order/flags R0 R1 R2 TXxTYxTZ GTXxGTYxGTZ slab" with
 * each axis's radices comma-separated ("-": the default plan), TZ and GTZ
 * the z chunk of the spreading and gather tiles (0: whole columns), a gather
 * tile 0x0x0 for the per-atom gather, slab 1 for the y-z slab transforms. */
int sfac_autotune(void *h, int n, const double *w_d, const double *x_d, int reps, char *report,
                  int report_len);
int sfac_batch_autotune(void *b, int reps, char *report, int report_len);
int sfac_tune_export(char *buf, int len);
int sfac_tune_import(const char *text);
int sfac_tune_current(void *h, char *buf, int len);
int sfac_tune_clear(void);

/* Per-kernel timing: while enabled, every launch on the handle's stream is
 * bracketed by HIP events.  sfac_profile_read synchronises, writes up to
 * `max` entries (kernel names joined by '\n' into `names`, summed
 * milliseconds, launch counts), clears the record and returns the number of
 * entries (or a negative error code). */
int sfac_profile(void *h, int enable);
int sfac_profile_read(void *h, char *names, int names_len, double *ms, int *counts, int max);

#ifdef __cplusplus
}
#endif

#endif
