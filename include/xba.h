/*
 * xba.h -- C ABI of the MI355X (gfx950) sliding-window bundle-adjustment solver (libxba.so).
 *
 * What each entry point replaces in the reference (paths relative to the reference checkout's src/xclean/):
 *   xba_create / xba_upload / xba_iterate / xba_download
 *       vio/window.py:889-982 `solve(pb, iters=5, lam0=1e-4)` for a BATCH of windows: the same LM iteration
 *       (build_system :291-463, prior and IMU terms :831-886, damped Cholesky :933-941, back-substitution
 *       :496-519, on-manifold update :522-540, trial cost :466-493, accept/reject :967-981), with every window's
 *       accept/reject decided on the device.  The reference binds it as a Python function over a `Problem`
 *       (window.py:768-792); xclean_amd/window.py is the Python mirror that packs Problems and calls these.
 *   xba_linearize
 *       vio/window.py:985-1008 `final_linearisation(pb)` (and one build_system at any damping): the reduced
 *       system H, b of every window, plus per-observation whitened chi2 and the landmark marginals.
 *   xba_np_solve
 *       backend/ba.py:713-869 `_solve_np(poses, pts, okf, olm, ouv, rig, free, iters, lam0, prior, rotf)` and, with
 *       lm_policy = 1, its GPU twin `_solve_gpu` (ba.py:930-1089, the same solve under that driver's LM rule, here in
 *       FP64): the local-BA drop-ins that tests/test_ba_solver.py and LocalBA.solve call.
 *   xba_np_id_solve
 *       backend/ba.py:1212-1342 `_solve_np_id(poses, rho, mbear, hidx, hpose_ext, okf, olm, ouv, rig, free, iters,
 *       lam0, prior, rotf, rho_floor)`: the inverse-depth local BA (landmarks hosted by a keyframe or by a constant
 *       out-of-window pose); with lm_policy = 1 its GPU twin `_solve_gpu_id` (ba.py:1345-1497).
 *   xba_backward
 *       no reference counterpart (the reference has no differentiable solve): the implicit-function adjoint of
 *       the last solved update with respect to observed pixels and per-observation weights.
 *
 * Conventions (vio/CONVENTIONS.md): body poses T_wb, R_wb <- R_wb Exp(dphi), p_wb <- p_wb + dp; VI blocks
 * [dphi, dp, dv, dbg, dba]; landmarks (a, b, rho) hosted in a keyframe; whitened residual = (pred - meas) * isig;
 * Huber on the whitened norm; normal equations H dx = -b.
 *
 * Memory: arrays marked (host) are read during the call only; arrays marked (device) are device pointers that
 * must stay valid for the lifetime of the context (state arrays are read by xba_upload and written by
 * xba_download).  Every call is asynchronous on `stream` (a hipStream_t, 0 = the null stream) unless noted.
 * Return codes: 0 ok; < 0 an error (no exception ever crosses this ABI); xba_strerror(code) describes it.
 */
#ifndef XBA_H
#define XBA_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define XBA_OK 0
#define XBA_E_ARG (-1)          /* invalid argument (sizes, NULL pointer, index out of range) */
#define XBA_E_HIP (-2)          /* a HIP runtime call failed */
#define XBA_E_TOO_MANY_OBS (-3) /* a landmark has more observations than XBA_MAX_OBS_PER_LM */
#define XBA_E_RANGE (-4)        /* a landmark couples slots further apart than the kernel's slot tile allows */
#define XBA_E_SIZE (-5)         /* a reduced system is larger than XBA_MAX_STATE */
#define XBA_E_STATE (-6)        /* call order: e.g. xba_iterate before xba_upload */

#define XBA_MAX_OBS_PER_LM 64   /* one wavefront lane per observation of a landmark */
#define XBA_MAX_STATE 1024      /* reduced-system dimension per window */
#define XBA_IMU_REC 301         /* packed IMU factor: dR 9, dv 3, dp 3, dt 1, J 54, bg_lin 3, ba_lin 3, W 225 */
#define XBA_WPAR 24             /* per-window parameters: fx fy cx cy B huber_k imu_k pad Rbc(9) pbc(3) g_w(3) pad */

enum xba_precision {
    XBA_FP64 = 0,  /* assembly, Schur and Cholesky in FP64 (the plain path) */
    XBA_MIXED = 1  /* per-observation Jacobian record in FP32, assembly and Schur sums in FP64 (the multi-wavefront
                      kernels; the deterministic family accumulates FP32 with Kahan compensation); FP32 Cholesky of
                      the Jacobi-scaled system with FP64 iterative refinement; FP64 SE(3) update and state */
};

/* A batch of windows, all arrays concatenated over windows in window order.  Slots, landmarks, observations,
 * IMU factors and prior blocks of window w are the ranges [*_0[w], *_0[w+1]) of the concatenated arrays; slot
 * indices stored in lm_host, ob_pose, imu_ij and prior_slot are window-local.  The observations of landmark l are
 * [ob_start[l], ob_start[l+1]) (global indices). */
typedef struct xba_batch {
    int32_t nwin;
    /* (host) structure, read by xba_create */
    const int32_t *win_slot0;   /* (nwin+1) */
    const int32_t *win_lm0;     /* (nwin+1) */
    const int32_t *win_imu0;    /* (nwin+1) */
    const int32_t *win_pblk0;   /* (nwin+1) prior blocks */
    const int32_t *win_prow0;   /* (nwin+1) prior rows (rows of S, r) */
    const int32_t *slot_dim;    /* (nslot) 6 or 15 */
    const int32_t *ob_start;    /* (nlm+1) */
    const int32_t *lm_host;     /* (nlm) */
    const int32_t *ob_pose;     /* (nobs) */
    const int32_t *imu_ij;      /* (nimu,2) */
    const int32_t *prior_slot;  /* (npblk) */
    const int32_t *prior_dim;   /* (npblk) 6 or 15 */
    /* (device) data */
    double *Rwb, *pwb;          /* (nslot,3,3) (nslot,3) state */
    double *vel, *bgs, *bas;    /* (nslot,3) state of 15-dim slots (may be NULL when every slot is 6-dim) */
    double *lmp;                /* (nlm,3) state (a, b, rho) */
    const uint8_t *lm_mask;     /* (nlm) */
    const double *rho_meas;     /* (nlm) */
    const double *rho_isig;     /* (nlm) 0 = no inverse-depth prior */
    const double *ob_z;         /* (nobs,2) measured pixel (cam 0: u, v; cam 1: u_R, unused) */
    const double *ob_isig;      /* (nobs) whitening 1/sigma */
    const uint8_t *ob_cam;      /* (nobs) 0 left 2-D, 1 right 1-D */
    const uint8_t *ob_valid;    /* (nobs) */
    const double *win_par;      /* (nwin, XBA_WPAR) */
    const double *imu_rec;      /* (nimu, XBA_IMU_REC) */
    const double *prior_x0;     /* (npblk, 21): R0 9, p0 3, v0 3, bg0 3, ba0 3 */
    const double *prior_S;      /* window w: rows [win_prow0[w], win_prow0[w+1]) x (sum of its prior dims), packed */
    const double *prior_r;      /* (nprow) */
    /* optional (NULL: 1): per-observation multiplier of the robust weight and cost of each visual row.  No
     * reference counterpart; it is the input xba_backward differentiates with respect to. */
    const double *ob_w;         /* (nobs) */
} xba_batch;

typedef struct xba_options {
    int32_t precision;          /* enum xba_precision */
    int32_t lm_per_chunk;       /* landmarks per work item of the assembly kernels; 0 = automatic */
    int32_t slot_tile;          /* slot range of one work item's LDS tile; 0 = automatic */
    int32_t grid;               /* persistent grid of the assembly kernels (workgroups); 0 = automatic */
    int32_t use_graph;          /* 1: capture one LM iteration in a hipGraph and replay it */
    int32_t want_chi2;          /* 1: xba_linearize writes per-observation chi2 */
    int32_t deterministic;      /* 1: single-wavefront assembly kernels whose every sum runs in a fixed order
                                   (bitwise reproducible; FP32 tiles use Kahan-compensated FP32 accumulation);
                                   0: multi-wavefront kernels accumulating the reduced system in FP64 LDS atomics
                                   (reproducible to rounding only) */
    int32_t big_n;              /* windows whose reduced system is larger than big_n are factored by the tiled
                                   multi-workgroup Cholesky (64 x 64 tiles, trailing updates on MFMA); smaller ones
                                   by one workgroup each.  0 = automatic (192); -1 = never tiled */
    int32_t big_fp64;           /* mixed precision: 1 factors the tiled windows in FP64 (one solve, no refinement)
                                   instead of FP32 with FP64 refinement; 2 every window (the one-workgroup solves
                                   too); ignored in FP64 */
} xba_options;

typedef struct xba_ctx xba_ctx;

/* Plan a batch: validate the structure, build the work list, allocate device workspace.  Synchronous. */
int xba_create(xba_ctx **out, const xba_batch *batch, const xba_options *opt);
int xba_destroy(xba_ctx *ctx);

/* Copy the state (and measurements, in the working precision) from the batch's device arrays into the context. */
int xba_upload(xba_ctx *ctx, double lam0, void *stream);

/* Run `iters` LM iterations of every window on device-resident data (vio/window.py:889-982).  No host sync. */
int xba_iterate(xba_ctx *ctx, int32_t iters, void *stream);

/* Write the state back into the batch's device state arrays. */
int xba_download(xba_ctx *ctx, void *stream);

/* Per-window statistics after xba_iterate, (device) out (nwin, XBA_NSTATS): cost0, cost, accepted, iters, lam,
 * last Cholesky status (1 ok, 0 failed), last accept (1/0), refinement solves of the last iteration, FP64
 * fallbacks since the upload (mixed precision: iterations whose FP32 factor failed or whose refinement did not
 * converge, re-linearised and solved again in FP64), last FP32 refinement converged (1/0: 0 when this iteration
 * fell back; always 1 in FP64). */
#define XBA_NSTATS 10
int xba_stats(xba_ctx *ctx, double *out_dev, void *stream);

/* The last iteration's full reduced update dx (sum of window state dims, FP64, device out). */
int xba_last_dx(xba_ctx *ctx, double *out_dev, void *stream);

/* One linearisation at damping `lam` at the context's current state (final_linearisation when lam == 0):
 * (device) H_out: window w's n_w x n_w block at offset sum_{v<w} n_v^2; b_out: (sum n_w); cost_out (nwin):
 * robust visual + prior + IMU cost; chi2_out (nobs, may be NULL): whitened squared residual, +inf when not
 * evaluable; Hinv_out (nlm,3,3) and ok_out (nlm) (may be NULL): the damped landmark inverses; imu_chi2_out (nimu,
 * may be NULL): the whitened chi2 of every IMU factor. */
int xba_linearize(xba_ctx *ctx, double lam, double *H_out, double *b_out, double *cost_out, double *chi2_out,
                  double *Hinv_out, uint8_t *ok_out, double *imu_chi2_out, void *stream);

/* Sizes the caller needs for the outputs above: total state dimension and total H entries. */
int xba_sizes(const xba_ctx *ctx, int64_t *n_state, int64_t *n_hentries);

/* Diagnostics (no reference counterpart): run ONE LM iteration eagerly (it advances the solve like
 * xba_iterate(ctx, 1)) with HIP events around each launch, and write the durations in ms to ms_out (host,
 * XBA_NPROF): linearise+Schur of the narrow work items, of the wide work items, reduce, prior/IMU terms, Cholesky
 * solve + update, back-substitution + trial cost (narrow, wide), trial prior/IMU cost, accept, landmark commit.
 * A class with no work items reads ~0.  Synchronous. */
#define XBA_NPROF 10
int xba_profile_iteration(xba_ctx *ctx, double *ms_out, void *stream);

/* Assembly timing over the replays of xba_iterate (no reference counterpart; a measurement aid).  on = 1: every LM
 * iteration xba_iterate runs from now on records two HIP events on the caller's stream around its assembly launch
 * (the narrow linearise+Schur kernel, with the wide one beside it on the side stream ending first); with graphs the
 * iteration is captured as two graphs, the assembly and the rest, and the events are recorded between their
 * launches.  Between two reads the first XBA_TIMING_SLOTS iterations are recorded.  on = 0: one graph, no events.
 * Synchronous. */
#define XBA_TIMING_SLOTS 4096
int xba_set_assembly_timing(xba_ctx *ctx, int32_t on);
/* The recorded assembly durations since the last call: their sum in ms and their number (host out); waits for the
 * last recorded event and starts a new record. */
int xba_assembly_times(xba_ctx *ctx, double *ms_sum, int32_t *count);

/* Plan summary (host out, 8): work items, persistent grid, max slot tile, LDS bytes of the landmark kernels, LDS
 * bytes of the dense solve, partial-tile entries, partial-b entries, max state dimension. */
int xba_plan_info(const xba_ctx *ctx, int64_t *out);

/* The Gauss-Newton step of every window at the current state, dx = -(H + 1e-9 I)^-1 b with H, b the undamped
 * (lambda = 0) reduced system (vio/window.py:291-463, 865-886, 933-941 at lambda = 0), factored in FP64; the factor
 * is kept for xba_backward.  dx_out (device, sum n) may be NULL.  No reference counterpart: it is the forward half
 * of the differentiable solve. */
int xba_gn_step(xba_ctx *ctx, double *dx_out, void *stream);

/* Implicit-function adjoint of the step xba_gn_step computed (which must be the last call on ctx that touched the
 * state), reusing its Cholesky factor: given g = dL/d(dx) (device, sum n), writes dL/dz (device, nobs x 2, the
 * observed pixel coordinates ob_z) and dL/dw (device, nobs, the per-observation weights ob_w, 1 when the batch has
 * none), in the batch's observation order.  No reference counterpart. */
int xba_backward(xba_ctx *ctx, const double *g_dx, double *grad_z, double *grad_w, void *stream);

/* ---- the Euclidean local-BA drop-in (backend/ba.py:713-869) ---------------------------------------------------
 * poses (device, nk x 4 x 4, T_wc, in/out), pts (device, nl x 3, in/out), okf/olm (device int32, N), ouv (device,
 * N x 2), free_mask (device uint8, nk), optional prior (pr_c (nl,3), pr_rho0 (nl), pr_invsig (nl), pr_valid (nl); pr_c NULL = none) and
 * rotation factors (rot_ab (n_rot,2) int32, rot_M (n_rot,3,3), rot_sig (n_rot); n_rot = 0: none).  stats_out
 * (device, 4): chi2_0, chi2_n, rms_0, rms_n.  Synchronous (host-side planning from okf/olm copies). */
typedef struct xba_np_problem {
    int32_t nk, nl, N;
    double *poses, *pts;
    const int32_t *okf, *olm;
    const double *ouv;
    const uint8_t *free_mask;
    double fx, cx, cy, huber_k;
    const double *pr_c, *pr_rho0, *pr_invsig, *pr_valid;
    int32_t n_rot;
    const int32_t *rot_ab;
    const double *rot_M, *rot_sig;
    int32_t lm_policy;  /* 0: _solve_np's LM rule; 1: the GPU driver _solve_gpu's (ba.py:930-1089): every step applied,
                           rejected by the next iteration when the Huber-weighted chi2 grew (state restored, lambda
                           x4), non-finite steps lambda x10, the rotation factors' Hessian in the reduced system,
                           SO(3) projected once at the end; stats are then its Huber-weighted chi2 / rms */
} xba_np_problem;

int xba_np_solve(const xba_np_problem *q, int32_t iters, double lam0, int32_t precision, double *stats_out,
                 double *dx_hist_out, void *stream);

/* ---- ordered observation pairs (backend/ba.py:902-927 build_pairs) -------------------------------------------
 * Replaces build_pairs(olm_sorted, nl) -> (offs, pa, pb).  Device pointers: olm_sorted (n_obs, int32) is the landmark
 * of every observation, ascending; offs (nl + 1, int32) receives the first observation of every landmark and
 * offs[nl] = n_obs; pa, pb (cap, int32) receive every ordered pair (a, b) of observations of one landmark, landmarks
 * ascending, a-major then b.  *npair (host) receives the pair count sum_l c_l^2.  pa == NULL: offs and *npair only
 * (the size query).  cap < *npair: XBA_E_SIZE and nothing written to pa, pb.  An index outside [0, nl) or a list
 * that is not ascending: XBA_E_RANGE (the reference's bincount raises on the first and returns segments that are not
 * the landmarks' on the second).  Synchronous (the count is read back). */
int xba_build_pairs(const int32_t *olm_sorted, int64_t n_obs, int32_t nl, int32_t *offs, int32_t *pa, int32_t *pb,
                    int64_t cap, int64_t *npair, void *stream);

/* ---- run-end global BA in inverse depth (backend/global_ba.py:694-1785 global_ba_id; its damped reduced camera
 * system by the deflated PCG of backend/gba_linsolve.py:447-478) -------------------------------------------------
 * Replaces global_ba_id on the path with every XTEN_GBA_* mode unset and no inertial term, rotation factors, welds
 * or nuisance states: the reprojection rows (Huber on the pixel norm, optional per-keyframe weights), the per-row
 * disparity rows, the landmark prior, the gauge keyframes, LM with the reference's accept rule and best-state
 * bookkeeping, the 3-sigma MAD rejection and the second pass (re-weighted prior, no disparity rows), the final SO(3)
 * projection.  Every damped reduced system is solved by the deflated PCG to rtol (rtol_direct at or below
 * XBA_GBA_DIRECT_MAX unknowns, where the reference solves directly); a solve that does not reach it is reported and
 * its step not used (lambda x10), as the reference does for an unconverged PCG step.
 * Device pointers; poses (nk, 16) row-major 4x4 T_wc, rho (nl), mbear (nl, 3) are overwritten with the result.
 * stats (host, XBA_GBA_NSTATS doubles, indices XBA_GBA_ST_*); hist (host, hist_cap x XBA_GBA_HIST doubles per LM
 * iteration: pass, iteration, lambda, chi, chi2 (NaN without a step), accepted, PCG iterations, relative residual,
 * PCG reason, restarts) or NULL.  Synchronous; allocates its workspace. */
#define XBA_GBA_DIRECT_MAX 30000
#define XBA_GBA_MAX_NK 65536
#define XBA_GBA_HIST 10
typedef struct {
    int32_t nk, nl, N;
    double *poses, *rho, *mbear;
    const int32_t *host, *okf, *olm;
    const double *ouv;
    const double *odisp;          /* (N) disparity [px], 0 = none; NULL: no disparity rows */
    const double *vt_w;           /* (nk) vision weight per keyframe or NULL */
    const uint8_t *free_mask;     /* (nk) 0 = gauge keyframe */
    const double *pr_m0, *pr_rho0, *pr_wbear, *pr_invsig, *pr_valid;  /* (nl,2), (nl) x 4; all NULL: no prior */
    double fx, cx, cy, fB, rho_sigma, rho_floor;
    /* gyro rotation factors (global_ba.py:1904-1967 _rot_blocks): keyframe pairs (n_rot, 2), the measured relative
       rotation M (n_rot, 9, row-major), sigma (n_rot); rot_huber > 0: the per-factor Huber threshold (rotf_huber) */
    int32_t n_rot;
    const int32_t *rot_ab;
    const double *rot_M, *rot_sig;
    double rot_huber;
    /* loop welds (global_ba.py:1969-2078 _weld_blocks, beta held): keyframe pairs (n_weld, 2), the measured relative
       pose Z (n_weld, 16, row-major 4x4, translation already divided by 1 + beta), the information diagonal
       (n_weld, 6) */
    int32_t n_weld;
    const int32_t *weld_ab;
    const double *weld_Z, *weld_w;
} xba_gba_problem;
typedef struct {
    int32_t iters, passes, maxiter, seg_len;  /* seg_len 0: the derived segment length (capped at 23 keyframes) */
    int32_t profile;                          /* 1: time the phases with HIP events (XBA_GBA_ST_*_MS) */
    double lam0, reject_sigma, time_budget_s; /* time_budget_s <= 0: none */
    double rtol, rtol_direct;
    double rtol_direct_floor;  /* at or below XBA_GBA_DIRECT_MAX unknowns a PCG that stops short of rtol_direct at
                                  its attainable accuracy is still used when its best relative residual is <= this */
    /* marginal covariances (global_ba.py:2080-2126 _marginal_cov): want_cov > 0 samples that many free keyframes
       evenly and returns the 6x6 diagonal blocks of the inverse of the last LM iteration's damped reduced system
       (6 PCG solves each, to rtol_direct), only at or below XBA_GBA_DIRECT_MAX unknowns; cov_ids (host, want_cov
       int32) and cov (host, want_cov x 36) receive XBA_GBA_ST_NCOV of them */
    int32_t want_cov;
    int32_t *cov_ids;
    double *cov;
} xba_gba_options;
enum {
    XBA_GBA_ST_RMS0 = 0, XBA_GBA_ST_RMS1, XBA_GBA_ST_OBS, XBA_GBA_ST_REJECTED, XBA_GBA_ST_THR, XBA_GBA_ST_STEREO,
    XBA_GBA_ST_CHI0V, XBA_GBA_ST_CHI0S, XBA_GBA_ST_CHIBEST, XBA_GBA_ST_ACC, XBA_GBA_ST_BEST,      /* last pass */
    XBA_GBA_ST_P1_CHI0V, XBA_GBA_ST_P1_CHI0S, XBA_GBA_ST_P1_CHIBEST, XBA_GBA_ST_P1_ACC, XBA_GBA_ST_P1_BEST,
    XBA_GBA_ST_PASSES, XBA_GBA_ST_SIGMA, XBA_GBA_ST_BIAS, XBA_GBA_ST_SIGMA1, XBA_GBA_ST_SOLVES, XBA_GBA_ST_PCG_IT,
    XBA_GBA_ST_UNCONV, XBA_GBA_ST_SECONDS, XBA_GBA_ST_NSEG, XBA_GBA_ST_SEGLEN, XBA_GBA_ST_NNZB, XBA_GBA_ST_ITERS_RUN,
    XBA_GBA_ST_ASM_MS,   /* mean device time of one reduced-system assembly (HIP events; profile on) */
    XBA_GBA_ST_SETUP_MS, /* mean device time of one PCG set-up: smoother, deflation space, coarse inverse */
    XBA_GBA_ST_PCG_MS,   /* mean device time of one PCG recursion */
    XBA_GBA_ST_ROWS,     /* observation rows of the last pass */
    XBA_GBA_ST_NCOV,     /* marginal covariance blocks returned */
    XBA_GBA_NSTATS
};
int xba_gba_solve(const xba_gba_problem *q, const xba_gba_options *o, double *stats, double *hist, int32_t hist_cap,
                  void *stream);

/* ---- the Euclidean local BA without allocation ---------------------------------------------------------------
 * No counterpart in the reference: xba_np_solve for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  Same arguments, results and return codes
 * as xba_np_solve; its workspace is one static device array per GPU (about 13 MB, held until the process ends)
 * sized for nk <= XBA_NOALLOC_MAX_NK, nl <= XBA_NOALLOC_MAX_NL, N <= XBA_NOALLOC_MAX_N (larger problems return
 * XBA_E_SIZE), and the grouping of observations by landmark runs on the device.  No hipMalloc, no hipFree, no host
 * heap.  `stream` must belong to the current device.  Every call on one GPU shares that array: calls that overlap
 * in time on one GPU (two host threads, or two streams) race and must be serialised by the caller.  Synchronous. */
#define XBA_NOALLOC_MAX_NK 32
#define XBA_NOALLOC_MAX_NL 8192
#define XBA_NOALLOC_MAX_N 65536
int xba_np_solve_noalloc(const xba_np_problem *q, int32_t iters, double lam0, int32_t precision, double *stats_out,
                         double *dx_hist_out, void *stream);

/* ---- the inverse-depth local-BA drop-in (backend/ba.py:1212-1342 `_solve_np_id`) --------------------------------
 * poses (device, nk x 4 x 4, T_wc, in/out); rho (device, nl, in/out) inverse depth and mbear (device, nl x 3, in/out)
 * bearing (x_n, y_n, 1) of every landmark in its host camera frame; hidx (device int32, nl) the host keyframe or -1,
 * in which case the host T_wc is the constant row hpose_ext[l] = [R | t] (device, nl x 12; may be NULL when no
 * hidx is -1); okf/olm/ouv/free_mask as xba_np_problem; rho_floor the lower bound of an updated inverse depth;
 * optional prior (pr_m0 (nl,3) bearing mean, pr_rho0, pr_wbear, pr_invsig, pr_valid (nl); pr_m0 NULL = none) and
 * rotation factors (their Hessian blocks enter the reduced system here, ba.py:1293-1296).  stats_out (device, 4):
 * chi2_0, chi2_n, rms_0, rms_n.  Synchronous. */
typedef struct xba_npid_problem {
    int32_t nk, nl, N;
    double *poses, *rho, *mbear;
    const int32_t *hidx;
    const double *hpose_ext;
    const int32_t *okf, *olm;
    const double *ouv;
    const uint8_t *free_mask;
    double fx, cx, cy, huber_k, rho_floor;
    const double *pr_m0, *pr_rho0, *pr_wbear, *pr_invsig, *pr_valid;
    int32_t n_rot;
    const int32_t *rot_ab;
    const double *rot_M, *rot_sig;
    int32_t lm_policy;  /* as xba_np_problem.lm_policy; 1: _solve_gpu_id (ba.py:1345-1497) */
} xba_npid_problem;

int xba_np_id_solve(const xba_npid_problem *q, int32_t iters, double lam0, int32_t precision, double *stats_out,
                    double *dx_hist_out, void *stream);

/* ---- IMU preintegration of many intervals at once --------------------------------------------------------------
 * Replaces vio/imu_factor.py:1008-1048 `preintegrate` (the numba kernel preint_core, :545-704, with its
 * sub-interval step _step, :392-542) called once per interval, and :707-746 `sqrt_info_core` for the whitening.
 * Interval f preintegrates [t0, t1] over the samples of its stream with t <= horizon (the causal rule of
 * :46-66: later samples are never read, the last released one is held past its stamp), at the linearisation biases
 * bg_lin, ba_lin, with the squared continuous densities sg2 (white gyro), sa2 (white accel), swg2, swa2 (bias walks)
 * and the nominal sample period t_nom of the sample-support rule (:67-95; the caller resolves it as the reference's
 * _t_nom does: 1 / the declared rate, else the median spacing of the released stamps, else 0).  Every pointer is a
 * device pointer; streams are given as offsets into shared sample arrays, several intervals may share a stream.
 * Outputs follow the reference's Preint: dR, dv, dp, J (rows phi v p, columns bg ba), the UNSCALED covariance C in
 * residual order [rR rv rp rbg rba] and the info vector; nothing but info is meaningful unless info[0] == 0 (the
 * reference's status codes: 1 t1 <= t0 or non-finite, 2 no sample released, 3 t0 before the first sample,
 * 4 non-finite sample used, 5 no sample period for a covariance).  Optional: the whitening W = L^-1 with
 * L L^T = kappa C + J_bg sbg2 J_bg^T + J_ba sba2 J_ba^T (w_ok 0 when not positive definite, W then zero), and the
 * window solver's IMU record (XBA_IMU_REC).  Asynchronous on `stream`. */
#define XBA_PRE_NPAR 17  /* per interval: t0, t1, horizon, bg_lin (3), ba_lin (3), sg2, sa2, swg2, swa2, t_nom, kappa, sbg2, sba2 */
#define XBA_PRE_NINFO 9  /* status, n_sub, n_inner, extrap_span, t_data_end, gap_max, n_released, t_nom, white_scale_max */
typedef struct {
    int32_t nf;            /* intervals */
    const double *t;       /* sample stamps of all streams, concatenated; non-decreasing within a stream */
    const double *gyr;     /* (samples, 3) */
    const double *acc;     /* (samples, 3) */
    const int64_t *s0;     /* (nf,) first sample of interval f's stream */
    const int32_t *ns;     /* (nf,) samples in that stream */
    const double *par;     /* (nf, XBA_PRE_NPAR) */
    double *dR, *dv, *dp;  /* (nf, 9) row-major, (nf, 3), (nf, 3) */
    double *J;             /* (nf, 54) row-major 9 x 6 */
    double *C;             /* (nf, 225) row-major */
    double *info;          /* (nf, XBA_PRE_NINFO) */
    double *W;             /* (nf, 225) or NULL */
    int32_t *w_ok;         /* (nf,) or NULL */
    double *rec;           /* (nf, XBA_IMU_REC) or NULL */
} xba_preint_batch;

int xba_preintegrate(const xba_preint_batch *b, void *stream);

/* Replaces vio/imu_factor.py:1051-1063 `sqrt_information(pre, kappa, bias_sigma_g, bias_sigma_a)` (numba
 * sqrt_info_core, :707-746) for nf intervals already preintegrated: W (nf, 225, lower triangular, row-major) =
 * L^-1 with L L^T = kappa C + J_bg sbg2 J_bg^T + J_ba sba2 J_ba^T, from C (nf, 225) and J (nf, 54) in the layout
 * xba_preintegrate writes and par (nf, 3) = kappa, sbg2, sba2 per interval.  w_ok (nf) is 0 (and W zero) when the
 * matrix is not positive definite.  Every pointer is a device pointer.  Asynchronous on `stream`. */
int xba_sqrt_information(int32_t nf, const double *C, const double *J, const double *par, double *W, int32_t *w_ok,
                         void *stream);

/* ---- Marginalisation algebra of the square-root prior, batched ----------------------------------------------
 * Replaces vio/marg.py:165-177 `schur` and :73-84 `Prior.from_info` (with :212-242 `combine`'s re-expression of b
 * at the blocks' first linearisation points), which the reference evaluates with numpy's eigh one matrix at a time.
 * Matrix w of a batch is stored in an ld x ld slab (row-major, leading dimension ld <= XBA_MAX_STATE); every pointer
 * is a device pointer; `work` holds XBA_MARG_WORK(ld) doubles per matrix.  Asynchronous on `stream`. */
#define XBA_MARG_WORK(ld) (3 * (int64_t)(ld) * (ld) + 2 * (int64_t)(ld))
#define XBA_MARG_FLOOR 1e-12  /* eigenvalues at or below this times the largest are dropped (marg.py:32 EIG_FLOOR) */

/* H2 = H_kk - H_km H_mm^+ H_mk, b2 = b_k - H_km H_mm^+ b_m over keep's order (symmetrised), H_mm^+ through the
 * eigen-decomposition of the symmetrised H_mm with the floor.  keep and marg: index lists into matrix w (nkeep[w],
 * nmarg[w] entries, nkeep + nmarg <= ld). */
typedef struct {
    int32_t nb, ld;
    const int32_t *nkeep, *nmarg;  /* (nb,) */
    const int32_t *keep, *marg;    /* (nb, ld) */
    const double *H, *b;           /* (nb, ld, ld), (nb, ld) */
    double *H2, *b2;               /* (nb, ld, ld), (nb, ld) */
    double *work;
} xba_schur_batch;

int xba_marg_schur(const xba_schur_batch *q, void *stream);

/* S = diag(sqrt l) V^T over the eigenvalues l of (H + H^T) / 2 above the floor (rows in ascending l, as numpy's eigh
 * orders them; the rows past `rank` are zero), r = diag(1 / sqrt l) V^T b'.  b' = b, or, when blk0 is given,
 * b - H d with d = x_cur [-] x0 over the matrix's blocks [blk0[w], blk0[w+1]) (block k at offset boff[k], bdim[k]
 * 6 or 15 dims; states as R 9, p 3, v 3, bg 3, ba 3), d = [Log(R0^T R), p - p0, v - v0, bg - bg0, ba - ba0]. */
typedef struct {
    int32_t nb, ld;
    const int32_t *n;              /* (nb,) */
    const double *H, *b;           /* (nb, ld, ld), (nb, ld) */
    const int32_t *blk0;           /* (nb + 1,) or NULL */
    const int32_t *boff, *bdim;    /* (blocks,) */
    const double *bcur, *bx0;      /* (blocks, 21) */
    double *S, *r;                 /* (nb, ld, ld), (nb, ld) */
    int32_t *rank;                 /* (nb,) */
    double *lam;                   /* (nb, ld) or NULL: every eigenvalue, ascending */
    double *bx;                    /* (nb, ld) or NULL: b' (b re-expressed at x0) */
    double *work;
} xba_prior_batch;

int xba_prior_from_info(const xba_prior_batch *q, void *stream);

/* Diagnostics (no reference counterpart): a plain device-to-device copy kernel of `bytes` (a multiple of 16;
 * 16-byte aligned pointers) on `stream`, so HBM counters (FETCH_SIZE, WRITE_SIZE) can be calibrated against a
 * kernel whose traffic is known exactly.  Asynchronous. */
int xba_copy_kernel(const void *src, void *dst, int64_t bytes, void *stream);

const char *xba_strerror(int code);
const char *xba_version(void);

#ifdef __cplusplus
}
#endif
#endif /* XBA_H */
