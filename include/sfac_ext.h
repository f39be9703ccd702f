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
/* host-pointer form; returns E, fills vir[6], dx[3n], phi[n] */
double sfac_eval(void *h, int n, const double *w, const double *x, double vir[6],
                 double *dx, double *phi);

/* the handle's HIP stream (hipStream_t) and a blocking wait on it */
void *sfac_stream(void *h);
int sfac_sync(void *h);

/* out[0..11] = K0, K1, K2, ldim, order, flags, device bytes held, state,
 * influence-table mode (0 table in HBM, 1 closed-form Ewald from the
 * indices), graph replay enabled, y-z slab transforms in use, slab
 * transforms available for this grid */
int sfac_info(void *h, long long out[12]);

/* Transform path: 0 (default) the y-z slab transforms when the plane fits
 * in LDS and K1, K2 factor into 2, 3, 5; 1 separate z and y line passes. */
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
