/*
 * scigap.h -- C ABI of the MI355X (gfx950) GAP-TV / ADMM-TV library.
 *
 * The reference (Scientific-Research-Resource/SCI-algorithms, PnP_SCI/python)
 * is pure Python/NumPy and has no FFI; each entry point below replaces one
 * Python function or one loop of it, cited as file:line relative to the
 * reference checkout.  The Python mirror of the reference interface
 * (sci-algorithms_amd/pnp_sci) binds these with ctypes.
 *
 * Conventions
 *  - All array pointers are DEVICE pointers owned by the caller, fp32.
 *    Arithmetic is fp32; sums over frames (A x = sum_t Phi_t x_t) are
 *    accumulated in fp64 and rounded to fp32 once.
 *  - Video volumes are frame-major (T,H,W), C order: element (t,h,w) is at
 *    t*H*W + h*W + w.  The reference holds (H,W,T); scigap_hwt_to_thw /
 *    scigap_thw_to_hwt convert on the device.
 *  - Planes (measurement y, Phi_sum, ...) are (H,W), C order.
 *  - `stream` is a hipStream_t (NULL = the null stream).  Calls are
 *    asynchronous on that stream unless documented otherwise; calls on
 *    different streams with disjoint buffers may run concurrently.
 *  - Every entry point returns 0 on success or a negative SCIGAP_E* code;
 *    scigap_last_error() then returns a message (per host thread).  No C++
 *    exception crosses this boundary.  Scalar arguments are validated on the
 *    host (non-finite values are SCIGAP_EINVAL); device arrays are not
 *    scanned, and non-finite values in them propagate as in the reference's
 *    NumPy arithmetic.
 *  - Scratch memory is a caller-provided device workspace whose size the
 *    matching *_workspace_bytes() function returns; the library allocates
 *    no device memory on these paths.
 */
#ifndef SCIGAP_H
#define SCIGAP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SCIGAP_OK            0
#define SCIGAP_EINVAL       -1  /* bad shape, size or parameter */
#define SCIGAP_ENULL        -2  /* required pointer is NULL */
#define SCIGAP_EHIP         -3  /* HIP runtime error (message has details) */
#define SCIGAP_EWORKSPACE   -4  /* workspace too small */
#define SCIGAP_EALIAS       -5  /* input and output buffers overlap */
#define SCIGAP_EUNSUPPORTED -6  /* parameter combination not implemented */

/* Library version (major*10000 + minor*100 + patch). */
int scigap_version(void);

/* Message describing the last error on this host thread ("" if none). */
const char* scigap_last_error(void);

/* ---------------------------------------------------------------------- */
/* Forward model                                                          */
/* ---------------------------------------------------------------------- */

/* y[h,w] = sum_t x[t,h,w] * phi[t,h,w] (fp64 sum of exact products, one rounding).
 * Replaces A_(x, Phi), PnP_SCI/python/utils.py:10-15. */
int scigap_forward(const float* x, const float* phi, float* y,
                   int64_t H, int64_t W, int64_t T, void* stream);

/* x[t,h,w] = y[h,w] * phi[t,h,w].
 * Replaces At_(y, Phi), PnP_SCI/python/utils.py:17-26. */
int scigap_adjoint(const float* y, const float* phi, float* x,
                   int64_t H, int64_t W, int64_t T, void* stream);

/* phi_sum[h,w] = sum_t phi[t,h,w], with exact zeros replaced by 1.
 * Replaces pnp_sci_algo.py:491-492 (and pnp_sci_demo_kobe.py:77-78). */
int scigap_phi_sum(const float* phi, float* phi_sum,
                   int64_t H, int64_t W, int64_t T, void* stream);

/* Layout conversion between the reference's (H,W,T) and the library's
 * (T,H,W).  src and dst must not overlap.  No reference counterpart: the
 * reference never leaves NumPy's (H,W,T). */
int scigap_hwt_to_thw(const float* src, float* dst,
                      int64_t H, int64_t W, int64_t T, void* stream);
int scigap_thw_to_hwt(const float* src, float* dst,
                      int64_t H, int64_t W, int64_t T, void* stream);

/* sum over all elements of (a-b)^2, accumulated in fp64 into *out (device
 * pointer, overwritten).  Used for PSNR, utils.py:28-36 and
 * pnp_sci_algo.py:681-682,699-705. per_frame != 0 writes T values (one per
 * frame of a (T,H,W) volume) instead of one. */
int scigap_sq_error(const float* a, const float* b, double* out,
                    int64_t H, int64_t W, int64_t T, int per_frame,
                    void* stream);

/* Mean structural similarity of every frame of two (T,H,W) volumes, as
 * scikit-image 0.18.3 structural_similarity(X, Y, data_range) with its
 * defaults (7x7 uniform window, sample covariance, K1=0.01, K2=0.03, mean
 * over the image minus a 3-pixel border), called per frame by the reference
 * at pnp_sci_algo.py:703 and :861.  Writes T doubles to `out` (device).
 * Requires H >= 7 and W >= 7. */
int scigap_ssim(const float* a, const float* b, double* out,
                int64_t H, int64_t W, int64_t T, float data_range, void* stream);

/* ---------------------------------------------------------------------- */
/* Chambolle TV                                                           */
/* ---------------------------------------------------------------------- */

/* Per-frame 2-D Chambolle TV of a (T,H,W) volume, as scikit-image 0.18.3
 * denoise_tv_chambolle(image, weight, eps, n_iter_max, multichannel=True)
 * (skimage/restoration/_denoise.py:315-479), called by the reference at
 * pnp_sci_algo.py:650 and :812: the dual field starts at zero, n_iter_max
 * iterations perform n_iter_max-1 dual updates, and each frame stops early
 * when |E_prev - E| < eps*E_0.  `image` and `out` must not overlap.
 * `stop_iter` (optional, device, T ints) receives the iteration whose image
 * was returned per frame. */
size_t scigap_tv_workspace_bytes(int64_t H, int64_t W, int64_t T, int n_iter_max);
int scigap_tv_chambolle(const float* image, float* out,
                        int64_t H, int64_t W, int64_t T,
                        float weight, float eps, int n_iter_max,
                        int* stop_iter, void* workspace, size_t workspace_bytes,
                        void* stream);

/* The same TV with an explicit fusion depth: `depth` dual updates per launch
 * (1..6), continuing from the stored dual field between launches, with the
 * early exit applied after each launch and frames that stopped inside it
 * recomputed.  depth 0: all n_iter_max-1 updates in one launch when they are
 * at most 6, else the autotuned depth (6 when untuned).  Results do not
 * depend on the depth.  The workspace for a given depth is
 * scigap_tv_workspace_bytes_depth(); depth 1 needs the most. */
size_t scigap_tv_workspace_bytes_depth(int64_t H, int64_t W, int64_t T, int n_iter_max,
                                       int depth);
int scigap_tv_chambolle_depth(const float* image, float* out,
                              int64_t H, int64_t W, int64_t T,
                              float weight, float eps, int n_iter_max, int depth,
                              int* stop_iter, void* workspace, size_t workspace_bytes,
                              void* stream);

/* The TV of scigap_tv_chambolle for n_iter_max <= 7 without a caller
 * workspace: no counterpart in the reference; for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops).
 * Its scratch (energy slots, stop array, counter) is a device array of
 * scigap_tv_noalloc_max_bytes() bytes inside the library, one per GPU, held
 * until the process ends; shapes needing more (scigap_tv_workspace_bytes())
 * are rejected with SCIGAP_EINVAL, n_iter_max > 7 with SCIGAP_EUNSUPPORTED.
 * The call allocates nothing.  Calls on one GPU share that array: calls that
 * overlap in time on one GPU (different streams, or host threads) race. */
size_t scigap_tv_noalloc_max_bytes(void);
int scigap_tv_chambolle_noalloc(const float* image, float* out,
                                int64_t H, int64_t W, int64_t T,
                                float weight, float eps, int n_iter_max,
                                int* stop_iter, void* stream);

/* ---------------------------------------------------------------------- */
/* GAP-TV / ADMM-TV solver                                                */
/* ---------------------------------------------------------------------- */

#define SCIGAP_METHOD_GAP  0   /* pnp_sci_algo.py:536-706            */
#define SCIGAP_METHOD_ADMM 1   /* pnp_sci_algo.py:708-864            */

typedef struct scigap_params {
  int64_t H, W, T;     /* measurement H x W, T coded frames              */
  int method;          /* SCIGAP_METHOD_GAP or SCIGAP_METHOD_ADMM        */
  int accelerate;      /* GAP: accelerated update (pnp_sci_algo.py:641)  */
  int clip;            /* ADMM: clip theta to [0,1] after TV
                          (joint_pnp_sci_algo.py:633)                    */
  int iters;           /* outer iterations (sum of iter_max stages)      */
  float lambda_;       /* _lambda                                        */
  float gamma;         /* ADMM gamma (added to Phi_sum)                  */
  float tv_weight;     /* denoise_tv_chambolle weight                    */
  float tv_eps;        /* denoise_tv_chambolle eps (2e-4 by default)     */
  int tv_iter_max;     /* denoise_tv_chambolle n_iter_max (>= 1)         */
  float tv_weight_decay; /* per-iteration factor on tv_weight (1: none;
                            0.999 in ADMM_TV_rec, pnp_sci_algo.py:898)   */
  float gamma_decay;   /* per-iteration factor on gamma (1: none;
                          0.998 in ADMM_TV_rec, pnp_sci_algo.py:899)     */
  int tile_rows;       /* fused-kernel workgroup height in waves for
                          tv_iter_max = 5: 8, 12 or 16 (0 = 8); a tuning
                          knob with no effect on results                 */
  int engine;          /* fused-kernel engine for tv_iter_max = 5 (0 =
                          6): 1 two LDS exchanges per TV update,
                          2 one exchange per update, 3/4 as 1/2 with
                          interior and border tiles in separate launches,
                          5 as 3 with the next frame prefetched straight
                          into LDS, 6 as 1 with each tile's frames split
                          over up to 8 workgroups (as many as evens out
                          the last round of workgroups on the CUs); a
                          tuning knob with no
                          effect on results beyond rounding of sums      */
} scigap_params;

/* Fill *p with the reference defaults for a GAP solve of an HxWxT problem:
 * accelerate=1, lambda=1, gamma=0.01, tv_weight=0.1, tv_eps=2e-4,
 * tv_iter_max=5, iters=50, no decay (pnp_sci_algo.py:536-539,708-711). */
void scigap_default_params(scigap_params* p, int64_t H, int64_t W, int64_t T);

size_t scigap_solve_workspace_bytes(const scigap_params* p);

/* One full GAP-TV or ADMM-TV reconstruction.
 *   y        (H,W)   measurement
 *   phi      (T,H,W) masks
 *   phi_sum  (H,W)   taken as passed (the reference does not recompute it)
 *   x0       (T,H,W) warm start, or NULL for x0 = At(y) (pnp_sci_algo.py:625-627)
 *   x_out    (T,H,W) result: GAP returns x after the last TV step, ADMM
 *                    returns the last projection x (pnp_sci_algo.py:706,864)
 *   truth    (T,H,W) or NULL; when given, sq_err[k] receives
 *                    sum((truth - x_k)^2) in fp64 for every outer iteration
 *                    k (device array of `iters` doubles), the quantity behind
 *                    psnr_all (pnp_sci_algo.py:681-682)
 *   stats    optional device array of `iters` ints: per iteration, the
 *            number of frames whose TV stopped early on the energy test
 *            (ADMM: 0 in the last iteration, whose TV result never reaches
 *            the returned x and is not computed).
 * x0 may not overlap x_out.  The call is asynchronous on `stream`. */
int scigap_solve(const scigap_params* p,
                 const float* y, const float* phi, const float* phi_sum,
                 const float* x0, float* x_out,
                 const float* truth, double* sq_err, int* stats,
                 void* workspace, size_t workspace_bytes, void* stream);

/* ---------------------------------------------------------------------- */
/* Batched solves (mixed-size queue)                                      */
/* ---------------------------------------------------------------------- */

/* One independent measurement of a batch.  The reference solves coded frames
 * and measurement blocks one after another in Python loops
 * (pnp_sci_algo.py:498, pnp_sci_test_data_realexp_blkproc.py:93); a batch
 * runs all of them together, each with its own H, W, T. */
typedef struct scigap_batch_item {
  int64_t H, W, T;
  const float* y;        /* (H,W)                                  */
  const float* phi;      /* (T,H,W)                                */
  const float* phi_sum;  /* (H,W), taken as passed                 */
  const float* x0;       /* (T,H,W) warm start or NULL for At(y)   */
  float* x_out;          /* (T,H,W) result                         */
} scigap_batch_item;

typedef struct scigap_batch_options {
  int use_graph;   /* 1: capture the whole outer loop in a hipGraph on the
                      first run and replay it afterwards; 0: plain launches */
  int tile_rows;   /* waves per workgroup of the fused kernel for
                      tv_iter_max = 5 (8, 12 or 16); 0 = 8                */
  int grid;        /* persistent workgroups (0 = as many as the device
                      holds at once)                                      */
  int order;       /* 0: tiles of the most frames first; 1: submission order */
} scigap_batch_options;

typedef struct scigap_batch_plan scigap_batch_plan;  /* opaque */

/* Workspace for a batch of n solves that share the solver parameters p
 * (p->H, p->W, p->T are ignored; the items carry their own shapes). */
size_t scigap_batch_workspace_bytes(const scigap_params* p, const scigap_batch_item* items,
                                    int n, const scigap_batch_options* opt);

/* Build a plan: per-iteration launch descriptors and tile queues are written
 * into `workspace` (device memory the caller keeps alive with the plan).
 * Supports tv_iter_max in [2, 7] (the fused depth). opt may be NULL. */
int scigap_batch_plan_create(scigap_batch_plan** plan, const scigap_params* p,
                             const scigap_batch_item* items, int n, void* workspace,
                             size_t workspace_bytes, const scigap_batch_options* opt,
                             void* stream);

/* Run every solve of the plan to completion on `stream` (asynchronous).  The
 * first run with use_graph captures the launches into a hipGraph bound to
 * the plan's buffers; later runs replay it. */
int scigap_batch_run(scigap_batch_plan* plan, void* stream);

/* Free the plan (host objects and the graph; the workspace stays the caller's). */
int scigap_batch_plan_destroy(scigap_batch_plan* plan);

/* ---------------------------------------------------------------------- */
/* Differentiable unrolled GAP-TV                                         */
/* ---------------------------------------------------------------------- */

/* K outer iterations of accelerated GAP-TV (pnp_sci_algo.py:625-650, the
 * loop of gap_denoise with denoiser='tv') unrolled as a fixed computation:
 * iteration k uses its own lambda_k and tv_weight_k, and every TV call runs
 * exactly tv_iter_max-1 Chambolle dual updates (scikit-image _denoise.py:
 * 341-393 without the energy early exit, which would make the map
 * discontinuous).  A batch of `batch` samples shares one mask.  The reference
 * has no backward pass; its networks (BIRNAT, RevSCI-net) train in PyTorch. */
typedef struct scigap_unrolled {
  int64_t H, W, T;
  int batch;        /* samples sharing the mask                               */
  int iters;        /* K outer iterations, >= 1                               */
  int tv_iter_max;  /* n in [2, 7]: n-1 dual updates per TV call              */
  int logits;       /* 1: `mask` holds logits and phi = sigmoid(mask);
                       0: `mask` holds phi itself                             */
  int measure;      /* 1: `meas` holds truth videos (batch,T,H,W) and
                       y_b = sum_t phi_t*truth_{b,t} is formed on the device;
                       0: `meas` holds the measurements y (batch,H,W)         */
} scigap_unrolled;

/* Bytes of the checkpoint buffer one forward call fills and the matching
 * backward call reads: per sample, x_k (T,H,W) and the residual r_k (H,W)
 * of every outer iteration, and y. */
size_t scigap_unrolled_checkpoint_bytes(const scigap_unrolled* u);

/* Workspace of a forward or a backward call (the same size for both);
 * with batch > 1 it holds the scratch of up to 4 concurrently running
 * samples. */
size_t scigap_unrolled_workspace_bytes(const scigap_unrolled* u);

/* Forward: x_out (batch,T,H,W) = x_K of every sample.  lambdas and weights
 * are HOST arrays of K floats (weights > 0).  mask (T,H,W) per `logits`,
 * meas per `measure`.  ckpt receives the per-iteration activations. */
int scigap_unrolled_forward(const scigap_unrolled* u, const float* lambdas,
                            const float* weights, const float* mask, const float* meas,
                            float* x_out, void* ckpt, size_t ckpt_bytes,
                            void* workspace, size_t workspace_bytes, void* stream);

/* Backward of the same call, from grad_x (batch,T,H,W) = dL/dx_out:
 *   grad_mask   (T,H,W) dL/dmask (logits or phi, as passed), or NULL
 *   grad_meas   dL/dy (batch,H,W) when measure = 0, or NULL; with
 *               measure = 1 the truth videos get no gradient (pass NULL)
 *   grad_lambda, grad_weight  device arrays of K doubles, or NULL
 * Every output is overwritten.  Inner TV states are recomputed per outer
 * iteration from the checkpoint (hand-written reverse kernels). */
int scigap_unrolled_backward(const scigap_unrolled* u, const float* lambdas,
                             const float* weights, const float* mask, const float* meas,
                             const void* ckpt, size_t ckpt_bytes, const float* grad_x,
                             float* grad_mask, float* grad_meas, double* grad_lambda,
                             double* grad_weight, void* workspace, size_t workspace_bytes,
                             void* stream);

/* ---------------------------------------------------------------------- */
/* Autotuning (no reference counterpart)                                  */
/* ---------------------------------------------------------------------- */

typedef struct scigap_tune_result {
  int tile_rows;      /* chosen fused-kernel tile height (solve)            */
  int engine;         /* chosen fused-kernel engine (solve)                 */
  int tv_depth;       /* chosen TV fusion depth (TV)                        */
  int candidates;     /* configurations timed                               */
  double best_ms;     /* per outer iteration (solve) or per call (TV)       */
  double default_ms;  /* the same for the first candidate (the default)     */
} scigap_tune_result;

/* Time the fused kernel's tile heights and engines on this problem (`iters`
 * outer iterations each; x_scratch is a (T,H,W) buffer the tuner
 * overwrites), remember the fastest for the shape, and return it.
 * Synchronises `stream`.  Later scigap_solve calls with tile_rows = 0 and
 * engine = 0 use the remembered choice.  The choice never changes results. */
int scigap_autotune_solve(const scigap_params* p, const float* y, const float* phi,
                          const float* phi_sum, float* x_scratch, int iters,
                          void* workspace, size_t workspace_bytes, void* stream,
                          scigap_tune_result* result);

/* Time every TV fusion depth for this shape (workspace of at least
 * scigap_tv_workspace_bytes_depth(H,W,T,n_iter_max,1) bytes), remember the
 * fastest (used by depth 0 for deep TV and by the solvers' deep-TV path).
 * Synchronises `stream`. */
int scigap_autotune_tv(const float* image, float* out, int64_t H, int64_t W, int64_t T,
                       float weight, float eps, int n_iter_max, void* workspace,
                       size_t workspace_bytes, void* stream, scigap_tune_result* result);

/* Forget every tuned choice (in this process; the cache file is kept). */
int scigap_tune_clear(void);

/* ---------------------------------------------------------------------- */
/* Kernel timing (measurement support; no reference counterpart)          */
/* ---------------------------------------------------------------------- */

typedef struct scigap_kernel_stat {
  char name[48];     /* kernel kind, e.g. "fused_interior"                */
  int launches;      /* launches recorded since scigap_profile_begin()     */
  double total_ms;   /* sum of their durations (HIP events on the stream) */
} scigap_kernel_stat;

/* Start recording a HIP event pair around every kernel launch the library
 * makes (process-wide; adds a little launch overhead while enabled). */
int scigap_profile_begin(void);

/* Stop recording, wait for the recorded events, and write up to `max`
 * per-kind totals to `out` (host memory); *n receives the number of kinds. */
int scigap_profile_end(scigap_kernel_stat* out, int max, int* n);

#ifdef __cplusplus
}
#endif

#endif /* SCIGAP_H */
