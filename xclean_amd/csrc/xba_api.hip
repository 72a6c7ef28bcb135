// libxba: the C ABI of include/xba.h -- planning (work list, offsets, workspace), upload/download of the working
// state, the LM iteration as a fixed kernel sequence (optionally replayed from one captured hipGraph), and the
// linearisation entry point.  One translation unit: the kernels are included below.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <iterator>
#include <mutex>
#include <vector>

#include "xba_internal.h"
#include "xba_lin.hip"
#include "xba_dense.hip"
#include "xba_bigchol.hip"
#include "xba_fast.hip"
#include "xba_np.hip"
#include "xba_adj.hip"
#include "xba_preint.hip"
#include "xba_marg.hip"
#include "xba_pairs.hip"
#include "xba_gba.hip"

using namespace xba;

#define HIPCK(x)                                                   \
    do {                                                           \
        hipError_t e_ = (x);                                       \
        if (e_ != hipSuccess) {                                    \
            fprintf(stderr, "xba: %s failed: %s\n", #x, hipGetErrorString(e_)); \
            return XBA_E_HIP;                                      \
        }                                                          \
    } while (0)

namespace {

template <typename T>
T *dalloc(size_t n, std::vector<void *> &owned) {
    void *p = nullptr;
    if (n == 0) n = 1;
    if (hipMalloc(&p, n * sizeof(T)) != hipSuccess) return nullptr;
    owned.push_back(p);
    return (T *)p;
}

// Working-precision conversion / packing kernels (upload and download).
template <typename T>
__global__ void k_upload(int nobs, int nlm, const int32_t *ob_pose, const uint8_t *ob_cam, const uint8_t *ob_valid,
                         const double *isig, int32_t *ob_info, T *isigT, const int32_t *lm_host,
                         const uint8_t *lm_mask, const double *rho_isig, int32_t *lm_info, T *riT,
                         const uint8_t *ob_tile, const uint8_t *lm_tile) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nobs || i < nlm; i += gridDim.x * blockDim.x) {
        if (i < nobs) {
            ob_info[i] = pack_obs(ob_pose[i], ob_cam[i] ? 1 : 0, ob_valid[i] ? 1 : 0, ob_tile ? ob_tile[i] : 0);
            isigT[i] = (T)isig[i];
        }
        if (i < nlm) {
            lm_info[i] = lm_host[i] | ((lm_mask[i] ? 1 : 0) << 12) | ((lm_tile ? lm_tile[i] : 0) << 13);
            riT[i] = (T)rho_isig[i];
        }
    }
}

__global__ void k_copy16(const uint4 *__restrict__ src, uint4 *__restrict__ dst, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

__global__ void k_init_ws(int nwin, double *ws, double lam0) {
    int w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= nwin) return;
    for (int k = 0; k < NWS; k++) ws[(size_t)NWS * w + k] = 0.0;
    ws[(size_t)NWS * w + WS_LAM] = lam0;
    ws[(size_t)NWS * w + WS_CHOLOK] = 1.0;
}

__global__ void k_stats(int nwin, const double *ws, double *out) {
    int w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= nwin) return;
    const double *s = ws + (size_t)NWS * w;
    double *o = out + XBA_NSTATS * (size_t)w;
    o[0] = s[WS_COST0]; o[1] = s[WS_CUR]; o[2] = s[WS_ACC]; o[3] = s[WS_ITERS]; o[4] = s[WS_LAM];
    o[5] = s[WS_CHOLOK]; o[6] = s[WS_ACCEPT]; o[7] = s[WS_NREF]; o[8] = s[WS_NFALLBACK]; o[9] = s[WS_CONV];
}

__global__ void k_cost_out(int nwin, const double *ws, double *out) {
    int w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w < nwin) out[w] = ws[(size_t)NWS * w + WS_CTOT];
}

}  // namespace

struct xba_ctx {
    xba_batch B;
    xba_options opt;
    int nwin = 0, nslot = 0, nlm = 0, nobs = 0, nimu = 0, npblk = 0;
    int nchunks = 0, Rmax = 1, maxn = 0, grid = 1;
    int64_t nstate = 0, nh = 0, npartH = 0, npartb = 0;
    int lin_lds = 0, solve_lds_cap = 0, fast_nw = 8, fast_nw_trial = 8, fast_lds = 0, fast_lds_trial = 0;
    int nnarrow = 0, Rnarrow = 1, Rwide = 1, wide_nw = 8, wide_lds = 0, grid_wide = 1;
    int grid_trial = 1, Rtrial = 1;  // the trial kernel runs up to XBA_TRIAL_PER_CU workgroups per CU
    int npass = 0;
    int32_t *pass_l0 = nullptr, *sched = nullptr, *sched_ctr = nullptr;
    double *pass_cost = nullptr;
    bool fast = true;
    bool uploaded = false;
    bool gn_ready = false;  // xba_gn_step ran and nothing has changed the state or the factor since
    double *adj_y = nullptr;
    std::vector<void *> owned;
    // device
    Chunk *chunks = nullptr;
    WinInfo *wins = nullptr;
    int32_t *colmap = nullptr, *slot_off = nullptr, *slot_dim = nullptr, *ob_start = nullptr, *lm_host = nullptr,
            *ob_pose = nullptr, *imu_ij = nullptr, *prior_slot = nullptr, *prior_dim = nullptr, *lm_win = nullptr;
    int32_t *ob_info = nullptr, *lm_info = nullptr;
    int32_t *chunk_slots = nullptr;                  // tile slot lists of the compacted work items
    int32_t *pair_start = nullptr, *pair_ent = nullptr, *bslot_start = nullptr, *bslot_ent = nullptr;
    uint8_t *ob_tile = nullptr, *lm_tile = nullptr;  // tile slot of every observation / landmark host (fast family)
    void *isigT = nullptr, *lmpT = nullptr, *lmp_trial = nullptr, *riT = nullptr;
    void *partH = nullptr, *partb = nullptr, *partH_lo = nullptr, *partb_lo = nullptr, *lmHi = nullptr, *lmbl = nullptr;
    double *partH64 = nullptr, *partb64 = nullptr;
    double *partc = nullptr, *partc2 = nullptr;
    uint8_t *lmok = nullptr;
    double *Rwb = nullptr, *pwb = nullptr, *vel = nullptr, *bgs = nullptr, *bas = nullptr;
    double *R2 = nullptr, *p2 = nullptr, *v2 = nullptr, *g2 = nullptr, *a2 = nullptr;
    double *H = nullptr, *b = nullptr, *dx = nullptr, *fact64 = nullptr, *wstate = nullptr;
    double *imu_out = nullptr;
    int32_t *imu_win = nullptr, *imu_color = nullptr;
    float *fact32 = nullptr;
    // windows factored by the tiled Cholesky
    int nbig = 0, big_npmax = 0, big_nbmax = 0, big_nmax = 0;
    bool big64 = false;    // mixed precision with the tiled windows factored in FP64
    bool small64 = false;  // mixed precision with the one-workgroup windows factored in FP64 too
    int small64_lds = 0;   // LDS of k_solve<double> for those windows
    BigWin *bigs = nullptr;
    float *bigF32 = nullptr;
    double *bigF64 = nullptr, *bigS = nullptr, *bigV = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    hipStream_t gstream = nullptr;
    // assembly timing (xba_set_assembly_timing): the iteration as two graphs, the assembly (gexec_a) and the rest
    // (gexec_b), with the event pair tev[2k], tev[2k + 1] recorded around the assembly of the k-th iteration since the
    // last read (tcount iterations so far)
    bool timing = false;
    hipGraph_t graph_a = nullptr, graph_b = nullptr;
    hipGraphExec_t gexec_a = nullptr, gexec_b = nullptr;
    std::vector<hipEvent_t> tev;
    int tcount = 0;
    // a second stream for independent launches of one iteration: the one-workgroup solves of the small windows
    // beside the tiled sequence of the large ones, the wide work items' assembly and trial beside the narrow ones
    hipStream_t side = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    bool solve_fork = false, wide_fork = false;
    int fb_lds_cap = 0;  // LDS of the FP64 fallback of the mixed solve
    int fb_R = 1, fb_nw = 1, fb_lds = 0, fb_grid = 1;  // its re-assembly launch: tile, waves, LDS, grid
    int maxn_small = 0;  // largest n among the windows k_solve factors
};

namespace {

template <typename T>
KArgs<T> kargs(xba_ctx *c) {
    KArgs<T> A;
    memset(&A, 0, sizeof(A));
    A.nchunks = c->nchunks;
    A.nwin = c->nwin;
    A.nlm = c->nlm;
    A.nobs = c->nobs;
    A.chunks = c->chunks;
    A.wins = c->wins;
    A.ob_start = c->ob_start;
    A.lm_info = c->lm_info;
    A.ob_info = c->ob_info;
    A.slot_off = c->slot_off;
    A.z = c->B.ob_z;
    A.isig = (const T *)c->isigT;
    A.lmp = (double *)c->lmpT;
    A.lmp_trial = (double *)c->lmp_trial;
    A.rho_meas = c->B.rho_meas;
    A.rho_isig = (const T *)c->riT;
    A.Rwb = c->Rwb;
    A.pwb = c->pwb;
    A.R2 = c->R2;
    A.p2 = c->p2;
    A.dx = c->dx;
    A.win_par = c->B.win_par;
    A.partH = (T *)c->partH;
    A.partb = (T *)c->partb;
    A.partH_lo = (T *)c->partH_lo;
    A.partb_lo = (T *)c->partb_lo;
    A.partH64 = c->partH64;
    A.partb64 = c->partb64;
    A.partc = c->partc;
    A.partc2 = c->partc2;
    A.lmHi = (T *)c->lmHi;
    A.lmbl = (T *)c->lmbl;
    A.lmok = c->lmok;
    A.wstate = c->wstate;
    A.lam_override = -1.0;
    A.pass_l0 = c->pass_l0;
    A.sched = c->sched;
    A.sched_ctr = c->sched_ctr;
    A.pass_cost = c->pass_cost;
    A.ob_w = c->B.ob_w;
    A.chunk_slots = c->chunk_slots;
    return A;
}

DArgs dargs(xba_ctx *c) {
    DArgs D;
    memset(&D, 0, sizeof(D));
    D.nwin = c->nwin;
    D.chunks = c->chunks;
    D.wins = c->wins;
    D.colmap = c->colmap;
    D.slot_off = c->slot_off;
    D.slot_dim = c->slot_dim;
    D.imu_ij = c->imu_ij;
    D.prior_slot = c->prior_slot;
    D.prior_dim = c->prior_dim;
    D.imu_rec = c->B.imu_rec;
    D.prior_x0 = c->B.prior_x0;
    D.prior_S = c->B.prior_S;
    D.prior_r = c->B.prior_r;
    D.win_par = c->B.win_par;
    D.Rwb = c->Rwb; D.pwb = c->pwb; D.vel = c->vel; D.bgs = c->bgs; D.bas = c->bas;
    D.R2 = c->R2; D.p2 = c->p2; D.v2 = c->v2; D.g2 = c->g2; D.a2 = c->a2;
    D.H = c->H;
    D.b = c->b;
    D.dx = c->dx;
    D.fact64 = c->fact64;
    D.fact32 = c->fact32;
    D.imu_out = c->imu_out;
    D.imu_win = c->imu_win;
    D.imu_color = c->imu_color;
    D.wstate = c->wstate;
    D.partc = c->partc;
    D.partc2 = c->partc2;
    D.nref_max = 6;
    D.fallback_on = c->opt.precision == XBA_MIXED ? 1 : 0;
    D.fallback_relin = (D.fallback_on && c->fast) ? 1 : 0;
    D.partH64 = c->partH64;
    D.partb64 = c->partb64;
    D.nbig = c->nbig;
    D.big64 = c->big64 ? 1 : 0;
    D.bigs = c->bigs;
    D.bigF32 = c->bigF32;
    D.bigF64 = c->bigF64;
    D.bigS = c->bigS;
    D.bigV = c->bigV;
    D.chunk_slots = c->chunk_slots;
    D.pair_start = c->pair_start;
    D.pair_ent = c->pair_ent;
    D.bslot_start = c->bslot_start;
    D.bslot_ent = c->bslot_ent;
    return D;
}

bool mixed(const xba_ctx *c) { return c->opt.precision == XBA_MIXED; }

// The distinct window slots landmark l touches (its host and the slots of its observations), ascending.
// Number of distinct slots among landmark l's host and observation rows (an upper bound of the slots its rows touch
// in the reduced system).
int distinct_slots(const xba_batch &B, int l) {
    int v[XBA_MAX_OBS_PER_LM + 1], n = 0;
    v[n++] = B.lm_host[l];
    for (int k = B.ob_start[l]; k < B.ob_start[l + 1] && n <= XBA_MAX_OBS_PER_LM; k++) {
        const int s = B.ob_pose[k];
        int i = 0;
        while (i < n && v[i] != s) i++;
        if (i == n) v[n++] = s;
    }
    return n;
}

std::vector<int> touched_slots(const xba_batch &B, int l) {
    std::vector<int> v(1, B.lm_host[l]);
    for (int k = B.ob_start[l]; k < B.ob_start[l + 1]; k++) v.push_back(B.ob_pose[k]);
    std::sort(v.begin(), v.end());
    v.erase(std::unique(v.begin(), v.end()), v.end());
    return v;
}

// The multi-wave kernels compute in FP64 whatever the precision (only their observation record is stored in the
// working type): their arguments take the FP64 whitening weights the caller uploaded and FP64 landmark blocks.
template <typename T>
KArgs<double> fast_args(const xba_ctx *c, const KArgs<T> &A) {
    KArgs<double> F;
    memset(&F, 0, sizeof(F));
    F.nchunks = A.nchunks; F.nwin = A.nwin; F.nlm = A.nlm; F.nobs = A.nobs;
    F.chunks = A.chunks; F.wins = A.wins; F.ob_start = A.ob_start; F.lm_info = A.lm_info; F.ob_info = A.ob_info;
    F.slot_off = A.slot_off; F.z = A.z;
    F.isig = c->B.ob_isig;
    F.lmp = A.lmp; F.lmp_trial = A.lmp_trial; F.rho_meas = A.rho_meas;
    F.rho_isig = c->B.rho_isig;
    F.Rwb = A.Rwb; F.pwb = A.pwb; F.R2 = A.R2; F.p2 = A.p2; F.dx = A.dx; F.win_par = A.win_par;
    F.partH64 = A.partH64; F.partb64 = A.partb64; F.partc = A.partc; F.partc2 = A.partc2;
    F.lmHi = (double *)c->lmHi; F.lmbl = (double *)c->lmbl; F.lmok = A.lmok;
    F.wstate = A.wstate; F.lam_override = A.lam_override; F.chi2_out = A.chi2_out; F.Hinv_out = A.Hinv_out;
    F.ok_out = A.ok_out; F.adj_gz = A.adj_gz; F.adj_gw = A.adj_gw; F.adj_y = A.adj_y;
    F.pass_l0 = A.pass_l0; F.sched = A.sched; F.sched_ctr = A.sched_ctr; F.pass_cost = A.pass_cost;
    F.ob_w = A.ob_w;
    F.chunk_slots = A.chunk_slots;
    F.only_fallback = A.only_fallback;
    return F;
}

// Prior and IMU terms of every window: the IMU factors in parallel (k_imu), then one workgroup per window.
void launch_terms(xba_ctx *c, const DArgs &D, hipStream_t s, int trial) {
    if (c->nimu > 0) hipLaunchKernelGGL(k_imu, dim3(c->nimu), dim3(256), 0, s, D, trial);
    hipLaunchKernelGGL(k_terms, dim3(c->nwin), dim3(1024), 0, s, D, trial);
}

// The linearisation half of one LM iteration: k_lin, k_reduce, k_terms.
// The wide work items' launch beside the narrow one (both kinds present, whole iteration): fork before the narrow
// launch, join after it.  Their outputs (per-chunk partials, per-pass costs, per-landmark records) are disjoint.
inline bool fork_wide(xba_ctx *c, hipStream_t s, int part) {
    if (part >= 0 || !c->wide_fork) return false;
    (void)hipEventRecord(c->ev_fork, s);
    (void)hipStreamWaitEvent(c->side, c->ev_fork, 0);
    return true;
}
inline void join_wide(xba_ctx *c, hipStream_t s) {
    (void)hipEventRecord(c->ev_join, c->side);
    (void)hipStreamWaitEvent(s, c->ev_join, 0);
}

// RT: the observation record type of the multi-wavefront kernels (T, or double for the mixed solve's FP64
// re-assembly of the windows it flagged, with A.only_fallback set).
template <typename T, typename RT = T>
void launch_lin_kernel(xba_ctx *c, const KArgs<T> &A, hipStream_t s, int part = -1) {
    if (c->fast) {
        const KArgs<double> F = fast_args<T>(c, A);
        const bool fork = fork_wide(c, s, part);
        if (c->nnarrow > 0 && part != 1)
            hipLaunchKernelGGL((k_lin_fast<RT, 0>), dim3(c->grid), dim3(XBA_WAVE * c->fast_nw), c->fast_lds, s, F,
                               c->Rnarrow, 0, c->nnarrow, 0);
        if (c->nchunks > c->nnarrow && part != 0)
            hipLaunchKernelGGL((k_lin_fast<RT, 1>), dim3(c->grid_wide), dim3(XBA_WAVE * c->wide_nw), c->wide_lds,
                               fork ? c->side : s, F, c->Rwide, c->nnarrow, c->nchunks, 1);
        if (fork) join_wide(c, s);
    } else if (part != 1) {
        hipLaunchKernelGGL(k_lin<T>, dim3(c->grid), dim3(XBA_WAVE), c->lin_lds, s, A, c->Rmax);
    }
}

// The trial: one launch over every work item (narrow and wide), handed out largest first.  `part` 1 (the profile's
// wide share) launches nothing.
template <typename T>
void launch_trial_kernel(xba_ctx *c, const KArgs<T> &A, hipStream_t s, int part = -1) {
    if (c->fast) {
        const KArgs<double> F = fast_args<T>(c, A);
        if (c->nchunks > 0 && part != 1)
            hipLaunchKernelGGL((k_trial_fast<0>), dim3(c->grid_trial), dim3(XBA_WAVE * c->fast_nw_trial),
                               c->fast_lds_trial, s, F, c->Rtrial, c->nchunks, 2 * c->nchunks, 2);
    } else if (part != 1) {
        hipLaunchKernelGGL(k_trial<T>, dim3(c->grid), dim3(XBA_WAVE), c->lin_lds, s, A, c->Rmax);
    }
}

template <typename T>
void launch_reduce(xba_ctx *c, const DArgs &D, hipStream_t s) {
    int gx = (int)(((int64_t)c->maxn * c->maxn + 255) / 256);
    const int gy = D.only_fallback ? 1 : c->nwin;  // the fallback's blocks visit the flagged windows themselves
    if (c->fast)
        hipLaunchKernelGGL(k_reduce<double>, dim3(std::max(gx, 1), gy), dim3(256), 0, s, D,
                           (const double *)c->partH64, (const double *)c->partb64, (const double *)nullptr,
                           (const double *)nullptr);
    else
        hipLaunchKernelGGL(k_reduce<T>, dim3(std::max(gx, 1), gy), dim3(256), 0, s, D, (const T *)c->partH,
                           (const T *)c->partb, (const T *)c->partH_lo, (const T *)c->partb_lo);
}

template <typename T>
void launch_lin(xba_ctx *c, const KArgs<T> &A, hipStream_t s, double *imu_chi2 = nullptr) {
    launch_lin_kernel<T>(c, A, s);
    DArgs D = dargs(c);
    D.imu_chi2 = imu_chi2;
    launch_reduce<T>(c, D, s);
    launch_terms(c, D, s, 0);
}

// The tiled Cholesky sequence of the large windows in precision T (FP64: one solve; FP32: refined in FP64).
template <typename T>
void launch_big(xba_ctx *c, const DArgs &D, hipStream_t s) {
    hipLaunchKernelGGL(k_big_prep<T>, dim3(c->big_npmax, c->nbig), dim3(256), 0, s, D, 0);
    for (int k = 0; k < c->big_nbmax; k++) {
        const int below = c->big_nbmax - k - 1;
        hipLaunchKernelGGL(k_big_panel<T>, dim3((below + 4) / 4, c->nbig), dim3(256), 0, s, D, k);
        if (below > 0)
            hipLaunchKernelGGL(k_big_update<T>, dim3(below * (below + 1) / 2, c->nbig), dim3(256), 0, s, D, k);
    }
    const int maxit = sizeof(T) == 8 ? 1 : D.nref_max;
    for (int it = 0; it < maxit; it++) {
        hipLaunchKernelGGL(k_big_trsv<T>, dim3(c->nbig), dim3(1024), 0, s, D, it);
        if (it + 1 < maxit)
            hipLaunchKernelGGL(k_big_resid, dim3((c->big_nmax + XBA_RES_ROWS - 1) / XBA_RES_ROWS, c->nbig),
                               dim3(256), 0, s, D);
    }
    hipLaunchKernelGGL(k_big_finish, dim3(c->nbig), dim3(256), 0, s, D);
}

// The damped solve and trial state of every window: the tiled Cholesky sequence for the large windows, one
// workgroup per window for the others.
template <typename T>
void launch_solve(xba_ctx *c, const DArgs &D, hipStream_t s) {
    const bool fork = c->solve_fork;  // both kinds present: k_solve beside the tiled sequence
    auto small = [&](hipStream_t st) {  // the one-workgroup windows (FP64 for every precision with small64)
        if (sizeof(T) == 4 && c->small64)
            hipLaunchKernelGGL(k_solve<double>, dim3(c->nwin), dim3(512), c->small64_lds, st, D, c->small64_lds, 0);
        else
            hipLaunchKernelGGL(k_solve<T>, dim3(c->nwin), dim3(512), c->solve_lds_cap, st, D, c->solve_lds_cap, 0);
    };
    if (fork) {
        (void)hipEventRecord(c->ev_fork, s);
        (void)hipStreamWaitEvent(c->side, c->ev_fork, 0);
        small(c->side);
        (void)hipEventRecord(c->ev_join, c->side);
    }
    if (c->nbig > 0) {
        if (sizeof(T) == 8 || c->big64)
            launch_big<double>(c, D, s);
        else
            launch_big<float>(c, D, s);
    }
    if (fork)
        (void)hipStreamWaitEvent(s, c->ev_join, 0);
    else if (c->nbig < c->nwin)
        small(s);
    // mixed precision: the windows flagged above are linearised again with FP64 observation records (the
    // multi-wavefront kernels; the FP32 records perturb H by ~1e-7 relative, which an ill-conditioned window
    // amplifies) and solved in FP64.  Workgroups of the other windows exit at once.
    if (sizeof(T) == 4 && D.fallback_on) {
        if (D.fallback_relin && c->nchunks > 0) {
            // every work item in one launch of the wide kernel's shape (its tile holds any work item's slots); the
            // reduction and terms of a flagged window run inside its FP64 solve below (the IMU records are current)
            KArgs<T> A = kargs<T>(c);
            A.only_fallback = 1;
            const KArgs<double> F = fast_args<T>(c, A);
            hipLaunchKernelGGL((k_lin_fast<double, 1>), dim3(c->fb_grid), dim3(XBA_WAVE * c->fb_nw), c->fb_lds, s, F,
                               c->fb_R, 0, c->nchunks, 1);
        }
        hipLaunchKernelGGL(k_solve<double>, dim3(c->nwin), dim3(512), c->fb_lds_cap, s, D, c->fb_lds_cap, 1);
    }
}

// One LM iteration (part -1), or its assembly launch (part 0: narrow and wide linearise+Schur) or the rest (part 1).
template <typename T>
void launch_iter(xba_ctx *c, hipStream_t s, int part = -1) {
    KArgs<T> A = kargs<T>(c);
    if (part != 1) launch_lin_kernel<T>(c, A, s);
    if (part == 0) return;
    DArgs D = dargs(c);
    launch_reduce<T>(c, D, s);
    launch_terms(c, D, s, 0);
    launch_solve<T>(c, D, s);
    launch_trial_kernel<T>(c, A, s);
    launch_terms(c, D, s, 1);
    hipLaunchKernelGGL(k_accept, dim3(c->nwin), dim3(256), 0, s, D);
    int gl = std::max(1, std::min((c->nlm + 255) / 256, 4096));
    hipLaunchKernelGGL(k_commit_lm<T>, dim3(gl), dim3(256), 0, s, A, (const int32_t *)c->lm_win);
}

}  // namespace

namespace {
// Working-precision copies of the measurements and the packed observation / landmark words (once per context).
int pack_measurements(xba_ctx *c, hipStream_t s) {
    const xba_batch &B = c->B;
    int n = std::max(c->nobs, c->nlm);
    int g = std::max(1, std::min((n + 255) / 256, 8192));
    if (mixed(c))
        hipLaunchKernelGGL(k_upload<float>, dim3(g), dim3(256), 0, s, c->nobs, c->nlm, c->ob_pose, B.ob_cam,
                           B.ob_valid, B.ob_isig, c->ob_info, (float *)c->isigT, c->lm_host, B.lm_mask, B.rho_isig,
                           c->lm_info, (float *)c->riT, (const uint8_t *)c->ob_tile, (const uint8_t *)c->lm_tile);
    else
        hipLaunchKernelGGL(k_upload<double>, dim3(g), dim3(256), 0, s, c->nobs, c->nlm, c->ob_pose, B.ob_cam,
                           B.ob_valid, B.ob_isig, c->ob_info, (double *)c->isigT, c->lm_host, B.lm_mask, B.rho_isig,
                           c->lm_info, (double *)c->riT, (const uint8_t *)c->ob_tile, (const uint8_t *)c->lm_tile);
    HIPCK(hipGetLastError());
    return XBA_OK;
}
}  // namespace

extern "C" {

int xba_preintegrate(const xba_preint_batch *b, void *stream) {
    if (!b || b->nf < 0) return XBA_E_ARG;
    if (b->nf == 0) return 0;
    if (!b->t || !b->gyr || !b->acc || !b->s0 || !b->ns || !b->par || !b->dR || !b->dv || !b->dp || !b->J ||
        !b->C || !b->info)
        return XBA_E_ARG;
    static_assert(PP_N == XBA_PRE_NPAR, "interval parameter record");
    const int grid = (b->nf + PRE_NG - 1) / PRE_NG;
    hipLaunchKernelGGL(k_preint, dim3(grid), dim3(PRE_NG * PRE_G), 0, (hipStream_t)stream, b->t, b->gyr, b->acc,
                       b->s0, b->ns, b->par, b->nf, b->dR, b->dv, b->dp, b->J, b->C, b->info, b->W, b->w_ok, b->rec);
    HIPCK(hipGetLastError());
    return 0;
}

int xba_sqrt_information(int32_t nf, const double *C, const double *J, const double *par, double *W, int32_t *w_ok,
                         void *stream) {
    if (nf < 0) return XBA_E_ARG;
    if (nf == 0) return 0;
    if (!C || !J || !par || !W || !w_ok) return XBA_E_ARG;
    hipLaunchKernelGGL(k_sqrt_info, dim3((nf + 63) / 64), dim3(64), 0, (hipStream_t)stream, (int)nf, C, J, par, W,
                       w_ok);
    HIPCK(hipGetLastError());
    return 0;
}

int xba_marg_schur(const xba_schur_batch *q, void *stream) {
    if (!q || q->nb < 0 || q->ld <= 0 || q->ld > XBA_MAX_STATE) return XBA_E_ARG;
    if (q->nb == 0) return 0;
    if (!q->nkeep || !q->nmarg || !q->keep || !q->marg || !q->H || !q->b || !q->H2 || !q->b2 || !q->work)
        return XBA_E_ARG;
    static_assert(MG_EIG_FLOOR == XBA_MARG_FLOOR, "floor");
    hipLaunchKernelGGL(k_marg_schur, dim3(q->nb), dim3(MG_THREADS), 0, (hipStream_t)stream, q->ld, q->nkeep,
                       q->nmarg, q->keep, q->marg, q->H, q->b, q->H2, q->b2, q->work);
    HIPCK(hipGetLastError());
    return 0;
}

int xba_prior_from_info(const xba_prior_batch *q, void *stream) {
    if (!q || q->nb < 0 || q->ld <= 0 || q->ld > XBA_MAX_STATE) return XBA_E_ARG;
    if (q->nb == 0) return 0;
    if (!q->n || !q->H || !q->b || !q->S || !q->r || !q->rank || !q->work) return XBA_E_ARG;
    if (q->blk0 && (!q->boff || !q->bdim || !q->bcur || !q->bx0)) return XBA_E_ARG;
    // LDS: room for the packed matrix and the eigenvectors of any matrix of order up to MG_PACK_N (the order rounded
    // up to even), and, when the leading dimension allows it, for an unpacked matrix of a larger order
    const int ld2 = std::min(q->ld + (q->ld & 1), MG_PACK_N);
    const size_t shm = (size_t)std::max(MG_PACK_DOUBLES(ld2), q->ld <= MG_LDS_LD ? q->ld * q->ld : 0) * sizeof(double);
    static bool attr = false;  // the LDS limit of the kernel, raised once per process
    if (!attr) {
        const int mx = std::max(MG_PACK_DOUBLES(MG_PACK_N), MG_LDS_LD * MG_LDS_LD) * (int)sizeof(double);
        HIPCK(hipFuncSetAttribute((const void *)k_prior_from_info, hipFuncAttributeMaxDynamicSharedMemorySize, mx));
        attr = true;
    }
    hipLaunchKernelGGL(k_prior_from_info, dim3(q->nb), dim3(MG_THREADS), shm, (hipStream_t)stream, q->ld, q->n, q->H,
                       q->b, q->blk0, q->boff, q->bdim, q->bcur, q->bx0, q->S, q->r, q->rank, q->lam, q->bx, q->work);
    HIPCK(hipGetLastError());
    return 0;
}

int xba_copy_kernel(const void *src, void *dst, int64_t bytes, void *stream) {
    if (!src || !dst || bytes < 0 || bytes % 16 || ((uintptr_t)src | (uintptr_t)dst) % 16) return XBA_E_ARG;
    const int64_t n = bytes / 16;
    if (n == 0) return XBA_OK;
    const int grid = (int)std::min<int64_t>((n + 255) / 256, 8192);
    hipLaunchKernelGGL(k_copy16, dim3(grid), dim3(256), 0, (hipStream_t)stream, (const uint4 *)src, (uint4 *)dst, n);
    HIPCK(hipGetLastError());
    return XBA_OK;
}

const char *xba_version(void) { return "xba 0.1 (gfx950)"; }

const char *xba_strerror(int code) {
    switch (code) {
        case XBA_OK: return "ok";
        case XBA_E_ARG: return "invalid argument";
        case XBA_E_HIP: return "HIP runtime error";
        case XBA_E_TOO_MANY_OBS: return "a landmark has more than XBA_MAX_OBS_PER_LM observations";
        case XBA_E_RANGE: return "an index is out of range or out of order, or a landmark couples slots further apart than the slot tile allows";
        case XBA_E_SIZE: return "a size exceeds a limit (XBA_MAX_STATE, a static workspace, or the output capacity given)";
        case XBA_E_STATE: return "call order (upload before iterate)";
        default: return "unknown error";
    }
}

int xba_destroy(xba_ctx *c) {
    if (!c) return XBA_OK;
    if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
    if (c->graph) (void)hipGraphDestroy(c->graph);
    if (c->gexec_a) (void)hipGraphExecDestroy(c->gexec_a);
    if (c->graph_a) (void)hipGraphDestroy(c->graph_a);
    if (c->gexec_b) (void)hipGraphExecDestroy(c->gexec_b);
    if (c->graph_b) (void)hipGraphDestroy(c->graph_b);
    for (hipEvent_t e : c->tev) (void)hipEventDestroy(e);
    if (c->ev_fork) (void)hipEventDestroy(c->ev_fork);
    if (c->ev_join) (void)hipEventDestroy(c->ev_join);
    if (c->side) (void)hipStreamDestroy(c->side);
    for (void *p : c->owned) (void)hipFree(p);
    delete c;
    return XBA_OK;
}

int xba_create(xba_ctx **out, const xba_batch *batch, const xba_options *opt) {
    if (!out || !batch || !opt || batch->nwin <= 0) return XBA_E_ARG;
    const xba_batch &B = *batch;
    if (!B.win_slot0 || !B.win_lm0 || !B.slot_dim || !B.ob_start || !B.lm_host || !B.ob_pose) return XBA_E_ARG;
    if (opt->precision != XBA_FP64 && opt->precision != XBA_MIXED) return XBA_E_ARG;
    xba_ctx *c = new xba_ctx();
    c->B = B;
    c->opt = *opt;
    const int nwin = B.nwin;
    c->nwin = nwin;
    c->nslot = B.win_slot0[nwin];
    c->nlm = B.win_lm0[nwin];
    c->nobs = B.ob_start[c->nlm];
    c->nimu = B.win_imu0 ? B.win_imu0[nwin] : 0;
    c->npblk = B.win_pblk0 ? B.win_pblk0[nwin] : 0;
    const bool mx = opt->precision == XBA_MIXED;
    const int tsz = mx ? 4 : 8;
    auto fail = [&](int code) {
        xba_destroy(c);
        return code;
    };
    // ---- per-window records
    std::vector<WinInfo> wins(nwin);
    std::vector<int32_t> slot_off(std::max(c->nslot, 1)), colmap, lm_win(std::max(c->nlm, 1)),
        imu_win(std::max(c->nimu, 1)), imu_color(std::max(c->nimu, 1));
    int64_t st = 0, h0 = 0, prS = 0;
    int maxn = 0;
    bool any_vi = false;
    for (int w = 0; w < nwin; w++) {
        WinInfo &W = wins[w];
        memset(&W, 0, sizeof(W));
        W.slot0 = B.win_slot0[w];
        W.P = B.win_slot0[w + 1] - W.slot0;
        W.lm0 = B.win_lm0[w];
        W.L = B.win_lm0[w + 1] - W.lm0;
        if (W.P <= 0 || W.P > 4095 || W.L < 0) return fail(XBA_E_ARG);
        int n = 0;
        for (int s = 0; s < W.P; s++) {
            int d = B.slot_dim[W.slot0 + s];
            if (d != 6 && d != 15) return fail(XBA_E_ARG);
            if (d == 15) W.vi = 1;
            slot_off[W.slot0 + s] = n;
            for (int i = 0; i < d; i++) colmap.push_back(i < 6 ? s * 8 + i : -1);
            n += d;
        }
        if (n > XBA_MAX_STATE) return fail(XBA_E_SIZE);
        W.n = n;
        W.st0 = st;
        W.h0 = h0;
        st += n;
        h0 += (int64_t)n * n;
        maxn = std::max(maxn, n);
        any_vi |= W.vi != 0;
        W.imu0 = B.win_imu0 ? B.win_imu0[w] : 0;
        W.imu1 = B.win_imu0 ? B.win_imu0[w + 1] : 0;
        W.pblk0 = B.win_pblk0 ? B.win_pblk0[w] : 0;
        W.pblk1 = B.win_pblk0 ? B.win_pblk0[w + 1] : 0;
        W.prow0 = B.win_prow0 ? B.win_prow0[w] : 0;
        W.prow1 = B.win_prow0 ? B.win_prow0[w + 1] : 0;
        int np_ = 0;
        for (int k = W.pblk0; k < W.pblk1; k++) {
            int s = B.prior_slot[k], d = B.prior_dim[k];
            if (s < 0 || s >= W.P || (d != 6 && d != 15) || d > B.slot_dim[W.slot0 + s]) return fail(XBA_E_ARG);
            np_ += d;
        }
        if (np_ > XBA_MAX_STATE || W.prow1 - W.prow0 > XBA_MAX_STATE) return fail(XBA_E_SIZE);
        W.np = np_;
        W.prS0 = prS;
        prS += (int64_t)(W.prow1 - W.prow0) * np_;
        for (int f = W.imu0; f < W.imu1; f++) {
            int i = B.imu_ij[2 * f], j = B.imu_ij[2 * f + 1];
            if (i < 0 || i >= W.P || j < 0 || j >= W.P || B.slot_dim[W.slot0 + i] != 15 ||
                B.slot_dim[W.slot0 + j] != 15)
                return fail(XBA_E_ARG);
        }
        for (int l = W.lm0; l < W.lm0 + W.L; l++) lm_win[l] = w;
        for (int f = W.imu0; f < W.imu1; f++) imu_win[f] = w;
        // colours: the lowest one not used by an earlier factor of the window sharing a slot
        W.imu_ncol = 0;
        for (int f = W.imu0; f < W.imu1; f++) {
            int col = 0;
            for (bool clash = true; clash; ) {
                clash = false;
                for (int g = W.imu0; g < f && !clash; g++)
                    if (imu_color[g] == col && (B.imu_ij[2 * g] == B.imu_ij[2 * f] || B.imu_ij[2 * g] == B.imu_ij[2 * f + 1] ||
                                                B.imu_ij[2 * g + 1] == B.imu_ij[2 * f] ||
                                                B.imu_ij[2 * g + 1] == B.imu_ij[2 * f + 1]))
                        clash = true;
                if (clash) col++;
            }
            imu_color[f] = col;
            W.imu_ncol = std::max(W.imu_ncol, col + 1);
        }
    }
    if (any_vi && (!B.vel || !B.bgs || !B.bas)) return fail(XBA_E_ARG);
    // windows above big_n: the tiled Cholesky (padded to whole 64 x 64 tiles)
    std::vector<BigWin> bigs;
    int64_t bigF = 0, bigSv = 0;
    {
        const int big_n = opt->big_n == 0 ? 192 : opt->big_n;
        for (int w = 0; w < nwin; w++) {
            WinInfo &W = wins[w];
            W.big = -1;
            if (big_n < 0 || W.n <= big_n) continue;
            BigWin bw;
            bw.win = w;
            bw.n = W.n;
            bw.np = (W.n + XBA_TILE - 1) / XBA_TILE * XBA_TILE;
            bw.nb = bw.np / XBA_TILE;
            bw.f0 = bigF;
            bw.s0 = bigSv;
            bw.v0 = 3 * bigSv;
            bigF += big_matrix_entries(bw.np);
            bigSv += bw.np;
            W.big = (int32_t)bigs.size();
            bigs.push_back(bw);
            c->big_npmax = std::max(c->big_npmax, bw.np);
            c->big_nbmax = std::max(c->big_nbmax, bw.nb);
            c->big_nmax = std::max(c->big_nmax, bw.n);
        }
        c->nbig = (int)bigs.size();
        for (int w = 0; w < nwin; w++)
            if (wins[w].big < 0) c->maxn_small = std::max(c->maxn_small, wins[w].n);
    }
    c->nstate = st;
    c->nh = h0;
    c->maxn = maxn;
    // ---- work list: landmark chunks whose observations stay inside a slot tile
    c->fast = opt->deterministic == 0;
    int dev = 0;
    HIPCK(hipGetDevice(&dev));
    int ncu = 0;
    HIPCK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    int tile = opt->slot_tile > 0 ? opt->slot_tile : (c->fast ? 16 : (mx ? 16 : 12));
    int tile_cap = mx ? 32 : 24;
    if (c->fast) {  // the widest slot tile whose one-wavefront layout fits the LDS
        tile_cap = 32;
        while (tile_cap > 1 && fast_layout(tile_cap, tsz, 1, false).total > 160 * 1024) tile_cap--;
    }
    int lpc = opt->lm_per_chunk;
    if (lpc <= 0) {
        if (c->fast)
            lpc = std::max(8, std::min(4096, (int)std::lround(c->nlm / (4.0 * ncu))));
        else
            lpc = std::max(4, std::min(1024, (int)std::lround(c->nlm / 2048.0)));
    }
    std::vector<int> llo(std::max(c->nlm, 1)), lhi(std::max(c->nlm, 1));
    for (int w = 0; w < nwin; w++) {
        const WinInfo &W = wins[w];
        for (int l = W.lm0; l < W.lm0 + W.L; l++) {
            int h = B.lm_host[l];
            if (h < 0 || h >= W.P) return fail(XBA_E_ARG);
            int o0 = B.ob_start[l], o1 = B.ob_start[l + 1];
            if (o1 < o0) return fail(XBA_E_ARG);
            if (o1 - o0 > XBA_MAX_OBS_PER_LM) return fail(XBA_E_TOO_MANY_OBS);
            int lo = h, hi = h;
            for (int k = o0; k < o1; k++) {
                int t = B.ob_pose[k];
                if (t < 0 || t >= W.P) return fail(XBA_E_ARG);
                lo = std::min(lo, t);
                hi = std::max(hi, t);
            }
            llo[l] = lo;
            lhi[l] = hi;
            if (c->fast && hi - lo + 1 > tile_cap) {
                // spans more slots than a tile holds: its work item's tile is the set of slots it touches
                if ((int)touched_slots(B, l).size() > tile_cap) return fail(XBA_E_RANGE);
                continue;
            }
            if (hi - lo + 1 > tile) tile = hi - lo + 1;
        }
    }
    if (tile > tile_cap) return fail(XBA_E_RANGE);
    std::vector<Chunk> chunks, wide;
    int64_t pH = 0, pb = 0;
    int Rmax = 1, Rn = 1, Rw = 1;
    // the fast family keeps landmarks that couple more slots than the narrow tile in separate (wide) work items
    const int narrow = c->fast ? std::min(tile, opt->slot_tile > 0 ? opt->slot_tile : 16) : tile;
    // wide work items run one wavefront per workgroup (their tile fills the LDS): keep them short so that they
    // spread over the CUs
    int lpc_wide = lpc;
    if (c->fast && opt->lm_per_chunk <= 0) {
        int64_t nwl = 0;
        for (int l = 0; l < c->nlm; l++) nwl += (lhi[l] - llo[l] + 1 > narrow) ? 1 : 0;
        lpc_wide = (int)std::max<int64_t>(16, std::min<int64_t>(lpc, nwl / std::max(ncu, 1)));
    }
    std::vector<int32_t> chunk_slots;
    auto emit = [&](std::vector<Chunk> &dst, int w, int cs, int ce, int lo, int hi) {
        Chunk ch;
        memset(&ch, 0, sizeof(ch));
        ch.win = w;
        ch.l0 = cs;
        ch.l1 = ce;
        ch.lo = lo;
        ch.R = hi - lo + 1;
        ch.smap = -1;
        dst.push_back(ch);
    };
    std::vector<int> wc0(nwin), wc1(nwin), ww0(nwin), ww1(nwin);
    for (int w = 0; w < nwin; w++) {
        const WinInfo &W = wins[w];
        wc0[w] = (int)chunks.size();
        ww0[w] = (int)wide.size();
        int l = W.lm0, lend = W.lm0 + W.L;
        while (l < lend) {
            if (c->fast && lhi[l] - llo[l] + 1 > tile_cap) {
                // compacted work item: consecutive such landmarks while the union of their touched slots fits
                std::vector<int> uni = touched_slots(B, l);
                const int cs = l;
                l++;
                while (l < lend && l - cs < lpc_wide && lhi[l] - llo[l] + 1 > tile_cap) {
                    std::vector<int> t = touched_slots(B, l), m;
                    std::set_union(uni.begin(), uni.end(), t.begin(), t.end(), std::back_inserter(m));
                    if ((int)m.size() > tile_cap) break;
                    uni.swap(m);
                    l++;
                }
                Chunk ch;
                memset(&ch, 0, sizeof(ch));
                ch.win = w;
                ch.l0 = cs;
                ch.l1 = l;
                ch.lo = uni.front();
                ch.R = (int)uni.size();
                ch.smap = (int32_t)chunk_slots.size();
                chunk_slots.insert(chunk_slots.end(), uni.begin(), uni.end());
                wide.push_back(ch);
                continue;
            }
            const bool isw = lhi[l] - llo[l] + 1 > narrow;
            const int lim = isw ? tile_cap : narrow;
            int cs = l, lo = llo[l], hi = lhi[l];
            l++;
            const int cap = isw ? lpc_wide : lpc;
            while (l < lend && l - cs < cap && ((lhi[l] - llo[l] + 1 > narrow) == isw)) {
                int nlo = std::min(lo, llo[l]), nhi = std::max(hi, lhi[l]);
                if (nhi - nlo + 1 > lim) break;
                lo = nlo;
                hi = nhi;
                l++;
            }
            emit(isw ? wide : chunks, w, cs, l, lo, hi);
        }
        wc1[w] = (int)chunks.size();
        ww1[w] = (int)wide.size();
    }
    const int nn = (int)chunks.size();
    for (int w = 0; w < nwin; w++) {
        wins[w].c0 = wc0[w];
        wins[w].c1 = wc1[w];
        wins[w].c2 = nn + ww0[w];
        wins[w].c3 = nn + ww1[w];
    }
    for (auto &ch : chunks) Rn = std::max(Rn, ch.R);
    for (auto &ch : wide) Rw = std::max(Rw, ch.R);
    chunks.insert(chunks.end(), wide.begin(), wide.end());
    for (auto &ch : chunks) {
        ch.partH = pH;
        ch.partb = pb;
        pH += (int64_t)(ch.R * (ch.R + 1) / 2) * 36;
        pb += 6 * ch.R;
        Rmax = std::max(Rmax, ch.R);
    }
    if (pH / 36 > INT32_MAX || pb / 6 > INT32_MAX) return fail(XBA_E_SIZE);
    // ---- reduction lists (k_reduce): per slot pair of a window, the partial blocks holding it, in the window's
    // work-item order (narrow items, then wide ones); per slot, its partial gradient segments
    std::vector<int32_t> pair_start, pair_ent, bslot_start(c->nslot + 1, 0), bslot_ent;
    {
        int64_t npairs = 0;
        for (int w = 0; w < nwin; w++) {
            if (npairs > INT32_MAX) return fail(XBA_E_SIZE);
            wins[w].pair0 = (int32_t)npairs;
            npairs += (int64_t)wins[w].P * (wins[w].P + 1) / 2;
        }
        if (npairs + 1 > INT32_MAX) return fail(XBA_E_SIZE);
        pair_start.assign(npairs + 1, 0);
        auto each_block = [&](auto &&fn) {
            for (int w = 0; w < nwin; w++) {
                for (int pass = 0; pass < 2; pass++) {
                    const int cb = pass ? wins[w].c2 : wins[w].c0, ce = pass ? wins[w].c3 : wins[w].c1;
                    for (int ci = cb; ci < ce; ci++) {
                        const Chunk &ch = chunks[ci];
                        for (int r2 = 0; r2 < ch.R; r2++) {
                            const int s2 = chunk_slot(ch, chunk_slots.data(), r2);
                            fn(w, ch, r2, s2, -1, -1);
                            for (int r1 = 0; r1 <= r2; r1++)
                                fn(w, ch, r2, s2, r1, chunk_slot(ch, chunk_slots.data(), r1));
                        }
                    }
                }
            }
        };
        each_block([&](int w, const Chunk &, int, int s2, int r1, int s1) {
            if (r1 < 0)
                bslot_start[wins[w].slot0 + s2 + 1]++;
            else
                pair_start[wins[w].pair0 + s2 * (s2 + 1) / 2 + s1 + 1]++;
        });
        for (size_t q = 1; q < pair_start.size(); q++) pair_start[q] += pair_start[q - 1];
        for (size_t q = 1; q < bslot_start.size(); q++) bslot_start[q] += bslot_start[q - 1];
        pair_ent.resize(std::max<int32_t>(pair_start.back(), 1));
        bslot_ent.resize(std::max<int32_t>(bslot_start.back(), 1));
        std::vector<int32_t> pf(pair_start.begin(), pair_start.end() - 1), bf(bslot_start.begin(), bslot_start.end() - 1);
        each_block([&](int w, const Chunk &ch, int r2, int s2, int r1, int s1) {
            if (r1 < 0)
                bslot_ent[bf[wins[w].slot0 + s2]++] = (int32_t)(ch.partb / 6 + r2);
            else
                pair_ent[pf[wins[w].pair0 + s2 * (s2 + 1) / 2 + s1]++] =
                    (int32_t)(ch.partH / 36 + r2 * (r2 + 1) / 2 + r1);
        });
    }
    // ---- pass plan and hand-out order of the multi-wave kernels: every work item is cut into passes whose
    // observations fit one lane each and whose touched slots fit the pass's U storage; work items are handed out
    // largest first within each class
    std::vector<int32_t> pass_l0, sched;
    if (c->fast) {
        std::vector<double> cost(chunks.size(), 0.0);
        for (size_t ci = 0; ci < chunks.size(); ci++) {
            Chunk &ch = chunks[ci];
            ch.p0 = (int)pass_l0.size();
            int l = ch.l0;
            while (l < ch.l1) {
                int nl = 0, S = 0, K = 0, PP = 0;
                pass_l0.push_back(l);
                while (l + nl < ch.l1 && nl < FB_LM) {
                    const int n = B.ob_start[l + nl + 1] - B.ob_start[l + nl];
                    const int k = std::min(distinct_slots(B, l + nl), ch.R);
                    const int pp = (k - 1) * (k - 2) / 2;  // the pair loop's pairs of target slots
                    if (nl > 0 && (S + n > FB_OBS || K + k > FB_HLP || PP + pp > FB_PAIRS)) break;
                    S += n;
                    K += k;
                    PP += pp;
                    nl++;
                    cost[ci] += (double)(n + 1) * (n + 1) + 4.0 * n + 8.0;
                }
                l += nl;
            }
            ch.p1 = (int)pass_l0.size();
        }
        // the assembly launches hand out each class (narrow, wide) largest first; the trial launch, which takes
        // every work item, hands them all out largest first from the second half of the list
        const size_t nc = chunks.size();
        sched.resize(2 * nc);
        for (size_t ci = 0; ci < nc; ci++) sched[ci] = sched[nc + ci] = (int32_t)ci;
        auto by_cost = [&](int32_t a, int32_t b) { return cost[a] > cost[b] || (cost[a] == cost[b] && a < b); };
        std::sort(sched.begin(), sched.begin() + nn, by_cost);
        std::sort(sched.begin() + nn, sched.begin() + nc, by_cost);
        std::sort(sched.begin() + nc, sched.end(), by_cost);
    }
    c->npass = (int)pass_l0.size();
    // tile slot of every observation and landmark host inside its work item (the multi-wave kernels read them from
    // the packed words)
    std::vector<uint8_t> ob_tile, lm_tile;
    if (c->fast) {
        ob_tile.assign(std::max(c->nobs, 1), 0);
        lm_tile.assign(std::max(c->nlm, 1), 0);
        for (const Chunk &ch : chunks)
            for (int l = ch.l0; l < ch.l1; l++) {
                lm_tile[l] = (uint8_t)chunk_tile(ch, chunk_slots.data(), B.lm_host[l]);
                for (int k = B.ob_start[l]; k < B.ob_start[l + 1]; k++)
                    ob_tile[k] = (uint8_t)chunk_tile(ch, chunk_slots.data(), B.ob_pose[k]);
            }
    }
    c->nnarrow = c->fast ? nn : (int)chunks.size();
    c->Rnarrow = Rn;
    c->Rwide = Rw;
    c->nchunks = (int)chunks.size();
    c->Rmax = Rmax;
    c->npartH = pH;
    c->npartb = pb;
    c->lin_lds = lds_layout(Rmax, tsz).total;
    const int lds_max = 160 * 1024;
    int per_cu = 1;
    if (c->fast) {
        auto pick = [&](int R, int &nw, int &lds, int &nwt, int &ldst) -> bool {
            nw = 8;
            while (nw > 1 && fast_layout(R, tsz, nw, false).total > lds_max) nw--;
            if (fast_layout(R, tsz, nw, false).total > lds_max) return false;
            lds = fast_layout(R, tsz, nw, false).total;
            nwt = 8;
            while (nwt > 1 && fast_layout(R, tsz, nwt, true).total > lds_max) nwt--;
            ldst = fast_layout(R, tsz, nwt, true).total;
            return true;
        };
        int nwt_, ldst_;  // (the trial's shape is picked below, for every work item at once)
        if (!pick(c->Rnarrow, c->fast_nw, c->fast_lds, nwt_, ldst_)) return fail(XBA_E_RANGE);
        if (!pick(c->Rwide, c->wide_nw, c->wide_lds, nwt_, ldst_)) return fail(XBA_E_RANGE);
        {  // the mixed fallback's re-assembly: one launch over every work item, tile of the widest
            int nwt, ldst;
            c->fb_R = std::max(c->Rnarrow, c->Rwide);
            if (!pick(c->fb_R, c->fb_nw, c->fb_lds, nwt, ldst)) return fail(XBA_E_RANGE);
            c->fb_grid = std::max(1, std::min(c->nchunks, ncu));
        }
        per_cu = 1;
        c->grid_wide = std::max(1, std::min(c->nchunks - c->nnarrow, ncu));
        // the trial: one launch over every work item with the widest tile (its LDS is small: the slot poses, steps
        // and per-wavefront row arrays); k_trial_fast is compiled for 4 waves per SIMD, two 8-wave workgroups per CU
        // where the LDS holds both
        {
            int nw, lds;
            c->Rtrial = std::max(c->Rnarrow, c->Rwide);
            if (!pick(c->Rtrial, nw, lds, c->fast_nw_trial, c->fast_lds_trial)) return fail(XBA_E_RANGE);
        }
        auto tpc = [&](int lds) { return std::max(1, std::min(XBA_TRIAL_PER_CU, lds_max / std::max(lds, 1))); };
        c->grid_trial = std::max(1, std::min(c->nchunks, ncu * tpc(c->fast_lds_trial)));
    } else {
        per_cu = std::max(1, std::min(8, lds_max / std::max(c->lin_lds, 1)));
    }
    c->grid = opt->grid > 0 ? opt->grid : std::max(1, std::min(c->fast ? c->nnarrow : c->nchunks, ncu * per_cu));
    if (c->nchunks == 0) c->grid = 1;
    if (opt->grid > 0) c->grid_trial = opt->grid;
    // ---- device allocations
    auto &O = c->owned;
    c->chunks = dalloc<Chunk>(chunks.size(), O);
    c->pass_l0 = dalloc<int32_t>(pass_l0.size(), O);
    c->sched = dalloc<int32_t>(sched.size(), O);
    c->sched_ctr = dalloc<int32_t>(8, O);
    c->pass_cost = dalloc<double>(pass_l0.size(), O);
    c->wins = dalloc<WinInfo>(nwin, O);
    c->colmap = dalloc<int32_t>(colmap.size(), O);
    c->slot_off = dalloc<int32_t>(c->nslot, O);
    c->slot_dim = dalloc<int32_t>(c->nslot, O);
    c->ob_start = dalloc<int32_t>(c->nlm + 1, O);
    c->lm_host = dalloc<int32_t>(c->nlm, O);
    c->ob_pose = dalloc<int32_t>(c->nobs, O);
    c->lm_win = dalloc<int32_t>(c->nlm, O);
    c->imu_ij = dalloc<int32_t>(2 * (size_t)c->nimu, O);
    c->prior_slot = dalloc<int32_t>(c->npblk, O);
    c->prior_dim = dalloc<int32_t>(c->npblk, O);
    c->ob_info = dalloc<int32_t>(c->nobs, O);
    c->chunk_slots = dalloc<int32_t>(chunk_slots.size(), O);
    c->pair_start = dalloc<int32_t>(pair_start.size(), O);
    c->pair_ent = dalloc<int32_t>(pair_ent.size(), O);
    c->bslot_start = dalloc<int32_t>(bslot_start.size(), O);
    c->bslot_ent = dalloc<int32_t>(bslot_ent.size(), O);
    if (c->fast) {
        c->ob_tile = dalloc<uint8_t>(c->nobs, O);
        c->lm_tile = dalloc<uint8_t>(c->nlm, O);
    }
    c->lm_info = dalloc<int32_t>(c->nlm, O);
    c->isigT = mx ? (void *)dalloc<float>(c->nobs, O) : (void *)dalloc<double>(c->nobs, O);
    c->lmpT = (void *)dalloc<double>(3 * (size_t)c->nlm, O);
    c->lmp_trial = (void *)dalloc<double>(3 * (size_t)c->nlm, O);
    c->riT = mx ? (void *)dalloc<float>(c->nlm, O) : (void *)dalloc<double>(c->nlm, O);
    const int64_t pHd = c->fast ? 1 : pH, pbd = c->fast ? 1 : pb;
    c->partH = mx ? (void *)dalloc<float>(pHd, O) : (void *)dalloc<double>(pHd, O);
    c->partb = mx ? (void *)dalloc<float>(pbd, O) : (void *)dalloc<double>(pbd, O);
    c->partH_lo = (mx && !c->fast) ? (void *)dalloc<float>(pH, O) : (void *)dalloc<double>(1, O);
    c->partb_lo = (mx && !c->fast) ? (void *)dalloc<float>(pb, O) : (void *)dalloc<double>(1, O);
    c->partH64 = c->fast ? dalloc<double>(pH, O) : dalloc<double>(1, O);
    c->partb64 = c->fast ? dalloc<double>(pb, O) : dalloc<double>(1, O);
    const bool lm32 = mx && !c->fast;  // the multi-wave kernels keep the landmark blocks in FP64
    c->lmHi = lm32 ? (void *)dalloc<float>(6 * (size_t)c->nlm, O) : (void *)dalloc<double>(6 * (size_t)c->nlm, O);
    c->lmbl = lm32 ? (void *)dalloc<float>(3 * (size_t)c->nlm, O) : (void *)dalloc<double>(3 * (size_t)c->nlm, O);
    c->partc = dalloc<double>(c->nchunks, O);
    c->partc2 = dalloc<double>(c->nchunks, O);
    c->lmok = dalloc<uint8_t>(c->nlm, O);
    c->Rwb = dalloc<double>(9 * (size_t)c->nslot, O);
    c->pwb = dalloc<double>(3 * (size_t)c->nslot, O);
    c->R2 = dalloc<double>(9 * (size_t)c->nslot, O);
    c->p2 = dalloc<double>(3 * (size_t)c->nslot, O);
    c->vel = dalloc<double>(3 * (size_t)c->nslot, O);
    c->bgs = dalloc<double>(3 * (size_t)c->nslot, O);
    c->bas = dalloc<double>(3 * (size_t)c->nslot, O);
    c->v2 = dalloc<double>(3 * (size_t)c->nslot, O);
    c->g2 = dalloc<double>(3 * (size_t)c->nslot, O);
    c->a2 = dalloc<double>(3 * (size_t)c->nslot, O);
    c->H = dalloc<double>(h0, O);
    c->b = dalloc<double>(st, O);
    c->dx = dalloc<double>(st, O);
    c->fact64 = dalloc<double>(h0, O);  // FP64 factor: the FP64 solve and the Gauss-Newton step of the adjoint
    c->adj_y = dalloc<double>(c->nstate, O);
    c->fact32 = mx ? dalloc<float>(h0, O) : dalloc<float>(1, O);
    c->imu_out = dalloc<double>(IMU_OUT * (size_t)std::max(c->nimu, 1), O);
    c->imu_win = dalloc<int32_t>(std::max(c->nimu, 1), O);
    c->imu_color = dalloc<int32_t>(std::max(c->nimu, 1), O);
    c->wstate = dalloc<double>((size_t)NWS * nwin, O);
    c->bigs = dalloc<BigWin>(bigs.size(), O);
    c->big64 = mx && opt->big_fp64 != 0;
    c->small64 = mx && opt->big_fp64 >= 2;
    c->bigF32 = (mx && !c->big64 && c->nbig) ? dalloc<float>(bigF, O) : dalloc<float>(1, O);
    c->bigF64 = ((!mx || c->big64) && c->nbig) ? dalloc<double>(bigF, O) : dalloc<double>(1, O);
    c->bigS = dalloc<double>(bigSv, O);
    c->bigV = dalloc<double>(3 * bigSv, O);
    for (void *p : O)
        if (!p) return fail(XBA_E_HIP);
    auto up = [&](void *dst, const void *src, size_t bytes) -> bool {
        if (bytes == 0 || !src) return true;
        return hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice) == hipSuccess;
    };
    bool ok = up(c->chunks, chunks.data(), chunks.size() * sizeof(Chunk)) &&
              (pass_l0.empty() || up(c->pass_l0, pass_l0.data(), pass_l0.size() * 4)) &&
              (sched.empty() || up(c->sched, sched.data(), sched.size() * 4)) &&
              up(c->wins, wins.data(), nwin * sizeof(WinInfo)) &&
              (bigs.empty() || up(c->bigs, bigs.data(), bigs.size() * sizeof(BigWin))) &&
              (chunk_slots.empty() || up(c->chunk_slots, chunk_slots.data(), chunk_slots.size() * 4)) &&
              up(c->pair_start, pair_start.data(), pair_start.size() * 4) &&
              up(c->pair_ent, pair_ent.data(), pair_ent.size() * 4) &&
              up(c->bslot_start, bslot_start.data(), bslot_start.size() * 4) &&
              up(c->bslot_ent, bslot_ent.data(), bslot_ent.size() * 4) &&
              (!c->fast || (up(c->ob_tile, ob_tile.data(), (size_t)c->nobs) && up(c->lm_tile, lm_tile.data(), (size_t)c->nlm))) &&
              up(c->colmap, colmap.data(), colmap.size() * 4) && up(c->slot_off, slot_off.data(), c->nslot * 4) &&
              up(c->slot_dim, B.slot_dim, c->nslot * 4) && up(c->ob_start, B.ob_start, (c->nlm + 1) * 4) &&
              up(c->lm_host, B.lm_host, c->nlm * 4) && up(c->ob_pose, B.ob_pose, (size_t)c->nobs * 4) &&
              up(c->lm_win, lm_win.data(), c->nlm * 4) && up(c->imu_win, imu_win.data(), (size_t)c->nimu * 4) &&
              up(c->imu_color, imu_color.data(), (size_t)c->nimu * 4) && up(c->imu_ij, B.imu_ij, 2 * (size_t)c->nimu * 4) &&
              up(c->prior_slot, B.prior_slot, c->npblk * 4) && up(c->prior_dim, B.prior_dim, c->npblk * 4);
    if (!ok) return fail(XBA_E_HIP);
    if (hipMemset(c->sched_ctr, 0, 8 * sizeof(int32_t)) != hipSuccess) return fail(XBA_E_HIP);
    // LDS budgets
    if (c->fast) {
        if (mx) {
            HIPCK(hipFuncSetAttribute((const void *)k_lin_fast<float, 0>, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
            HIPCK(hipFuncSetAttribute((const void *)k_lin_fast<float, 1>, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
        }
        // FP64 records: the FP64 path, and the mixed solve's FP64 re-assembly of flagged windows (same layout)
        HIPCK(hipFuncSetAttribute((const void *)k_lin_fast<double, 0>, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
        HIPCK(hipFuncSetAttribute((const void *)k_lin_fast<double, 1>, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
        HIPCK(hipFuncSetAttribute((const void *)k_trial_fast<0>, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
    }
    if (!c->fast && c->lin_lds > 64 * 1024) {  // the single-wavefront kernels' tile (their launch only)
        if (c->lin_lds > 160 * 1024) return fail(XBA_E_RANGE);
        if (mx) {
            HIPCK(hipFuncSetAttribute((const void *)k_lin<float>, hipFuncAttributeMaxDynamicSharedMemorySize, c->lin_lds));
            HIPCK(hipFuncSetAttribute((const void *)k_trial<float>, hipFuncAttributeMaxDynamicSharedMemorySize, c->lin_lds));
        } else {
            HIPCK(hipFuncSetAttribute((const void *)k_lin<double>, hipFuncAttributeMaxDynamicSharedMemorySize, c->lin_lds));
            HIPCK(hipFuncSetAttribute((const void *)k_trial<double>, hipFuncAttributeMaxDynamicSharedMemorySize, c->lin_lds));
        }
    }
    hipFuncAttributes fa;
    HIPCK(hipFuncGetAttributes(&fa, mx ? (const void *)k_solve<float> : (const void *)k_solve<double>));
    int cap = 160 * 1024 - (int)fa.sharedSizeBytes - 256;
    // k_solve keeps a window's matrix in LDS, rows padded to a multiple of 4 entries, when it fits
    int need = (int)std::min<int64_t>((int64_t)c->maxn_small * ((c->maxn_small + 3) & ~3) * tsz, (int64_t)cap);
    c->solve_lds_cap = std::max(0, need);
    if (c->solve_lds_cap > 64 * 1024)
        HIPCK(hipFuncSetAttribute(mx ? (const void *)k_solve<float> : (const void *)k_solve<double>,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, c->solve_lds_cap));
    if (mx) {  // the FP64 fallback of the mixed solve: the matrix in LDS when it fits, else in global memory
        HIPCK(hipFuncGetAttributes(&fa, (const void *)k_solve<double>));
        const int cap64 = 160 * 1024 - (int)fa.sharedSizeBytes - 256;
        const int64_t n64 = std::max(c->maxn, 1), ns64 = std::max(c->maxn_small, 1);
        c->fb_lds_cap = (int)std::max<int64_t>(0, std::min<int64_t>(n64 * ((n64 + 3) & ~3) * 8, cap64));
        // (the FP64 solve of the one-workgroup windows with small64: sized to those windows only)
        c->small64_lds = (int)std::max<int64_t>(0, std::min<int64_t>(ns64 * ((ns64 + 3) & ~3) * 8, cap64));
        if (c->fb_lds_cap > 64 * 1024 || c->small64_lds > 64 * 1024)
            HIPCK(hipFuncSetAttribute((const void *)k_solve<double>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                      std::max(c->fb_lds_cap, c->small64_lds)));
    }
    c->solve_fork = c->nbig > 0 && c->nbig < c->nwin;
    c->wide_fork = c->fast && c->nnarrow > 0 && c->nchunks > c->nnarrow;
    if (c->solve_fork || c->wide_fork) {
        HIPCK(hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking));
        HIPCK(hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming));
        HIPCK(hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming));
    }
    {
        int rc = pack_measurements(c, nullptr);
        if (rc != XBA_OK) return fail(rc);
        if (hipDeviceSynchronize() != hipSuccess) return fail(XBA_E_HIP);
    }
    *out = c;
    return XBA_OK;
}

int xba_sizes(const xba_ctx *c, int64_t *n_state, int64_t *n_h) {
    if (!c) return XBA_E_ARG;
    if (n_state) *n_state = c->nstate;
    if (n_h) *n_h = c->nh;
    return XBA_OK;
}

int xba_upload(xba_ctx *c, double lam0, void *stream) {
    if (!c) return XBA_E_ARG;
    hipStream_t s = (hipStream_t)stream;
    const xba_batch &B = c->B;
    HIPCK(hipMemcpyAsync(c->lmpT, B.lmp, 3 * (size_t)c->nlm * 8, hipMemcpyDeviceToDevice, s));
    size_t ns = (size_t)c->nslot;
    HIPCK(hipMemcpyAsync(c->Rwb, B.Rwb, 9 * ns * 8, hipMemcpyDeviceToDevice, s));
    HIPCK(hipMemcpyAsync(c->pwb, B.pwb, 3 * ns * 8, hipMemcpyDeviceToDevice, s));
    if (B.vel) HIPCK(hipMemcpyAsync(c->vel, B.vel, 3 * ns * 8, hipMemcpyDeviceToDevice, s));
    if (B.bgs) HIPCK(hipMemcpyAsync(c->bgs, B.bgs, 3 * ns * 8, hipMemcpyDeviceToDevice, s));
    if (B.bas) HIPCK(hipMemcpyAsync(c->bas, B.bas, 3 * ns * 8, hipMemcpyDeviceToDevice, s));
    hipLaunchKernelGGL(k_init_ws, dim3((c->nwin + 255) / 256), dim3(256), 0, s, c->nwin, c->wstate, lam0);
    HIPCK(hipGetLastError());
    c->uploaded = true;
    c->gn_ready = false;
    return XBA_OK;
}

int xba_iterate(xba_ctx *c, int32_t iters, void *stream) {
    if (!c || iters < 0) return XBA_E_ARG;
    if (!c->uploaded) return XBA_E_STATE;
    c->gn_ready = false;
    hipStream_t s = (hipStream_t)stream;
    // with assembly timing, the event pair of the next iteration (none once the record is full)
    auto ev_next = [&]() -> hipEvent_t * {
        return (c->timing && c->tcount < XBA_TIMING_SLOTS) ? &c->tev[2 * c->tcount] : nullptr;
    };
    if (c->opt.use_graph && c->timing) {
        for (int it = 0; it < iters; it++) {
            hipEvent_t *ev = ev_next();
            if (ev) HIPCK(hipEventRecord(ev[0], s));
            HIPCK(hipGraphLaunch(c->gexec_a, s));
            if (ev) {
                HIPCK(hipEventRecord(ev[1], s));
                c->tcount++;
            }
            HIPCK(hipGraphLaunch(c->gexec_b, s));
        }
        return XBA_OK;
    }
    if (c->opt.use_graph) {
        if (!c->gexec || c->gstream != s) {
            if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
            if (c->graph) (void)hipGraphDestroy(c->graph);
            c->gexec = nullptr;
            c->graph = nullptr;
            hipStream_t cs = s;
            bool own = false;
            if (cs == nullptr) {  // the null stream cannot be captured: capture on a private stream
                HIPCK(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
                own = true;
            }
            HIPCK(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
            if (mixed(c))
                launch_iter<float>(c, cs);
            else
                launch_iter<double>(c, cs);
            HIPCK(hipStreamEndCapture(cs, &c->graph));
            HIPCK(hipGraphInstantiate(&c->gexec, c->graph, nullptr, nullptr, 0));
            if (own) HIPCK(hipStreamDestroy(cs));
            c->gstream = s;
        }
        for (int it = 0; it < iters; it++) HIPCK(hipGraphLaunch(c->gexec, s));
        return XBA_OK;
    }
    for (int it = 0; it < iters; it++) {
        hipEvent_t *ev = ev_next();
        if (ev) HIPCK(hipEventRecord(ev[0], s));
        if (mixed(c))
            launch_iter<float>(c, s, 0);
        else
            launch_iter<double>(c, s, 0);
        if (ev) {
            HIPCK(hipEventRecord(ev[1], s));
            c->tcount++;
        }
        if (mixed(c))
            launch_iter<float>(c, s, 1);
        else
            launch_iter<double>(c, s, 1);
        HIPCK(hipGetLastError());
    }
    return XBA_OK;
}

int xba_set_assembly_timing(xba_ctx *c, int32_t on) {
    if (!c) return XBA_E_ARG;
    const bool t = on != 0;
    if (t && c->tev.empty()) {
        c->tev.resize(2 * XBA_TIMING_SLOTS, nullptr);
        for (auto &e : c->tev) HIPCK(hipEventCreate(&e));
    }
    if (t && c->opt.use_graph && !c->gexec_a) {
        // the assembly and the rest captured as two graphs (on a private stream: a captured graph can be launched
        // on any stream), uploaded so that their first launches cost what the later ones do
        hipStream_t cs;
        HIPCK(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
        for (int part = 0; part < 2; part++) {
            hipGraph_t *g = part == 0 ? &c->graph_a : &c->graph_b;
            hipGraphExec_t *e = part == 0 ? &c->gexec_a : &c->gexec_b;
            HIPCK(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
            if (mixed(c))
                launch_iter<float>(c, cs, part);
            else
                launch_iter<double>(c, cs, part);
            HIPCK(hipStreamEndCapture(cs, g));
            HIPCK(hipGraphInstantiate(e, *g, nullptr, nullptr, 0));
            HIPCK(hipGraphUpload(*e, cs));
        }
        HIPCK(hipStreamSynchronize(cs));
        HIPCK(hipStreamDestroy(cs));
    }
    c->timing = t;
    c->tcount = 0;
    return XBA_OK;
}

int xba_assembly_times(xba_ctx *c, double *ms_sum, int32_t *count) {
    if (!c || !ms_sum || !count) return XBA_E_ARG;
    double sum = 0.0;
    const int n = c->tcount;
    if (n > 0) HIPCK(hipEventSynchronize(c->tev[2 * n - 1]));
    for (int k = 0; k < n; k++) {
        float ms = 0.f;
        HIPCK(hipEventElapsedTime(&ms, c->tev[2 * k], c->tev[2 * k + 1]));
        sum += ms;
    }
    *ms_sum = sum;
    *count = n;
    c->tcount = 0;
    return XBA_OK;
}

int xba_download(xba_ctx *c, void *stream) {
    if (!c) return XBA_E_ARG;
    if (!c->uploaded) return XBA_E_STATE;
    hipStream_t s = (hipStream_t)stream;
    const xba_batch &B = c->B;
    size_t ns = (size_t)c->nslot;
    HIPCK(hipMemcpyAsync(B.Rwb, c->Rwb, 9 * ns * 8, hipMemcpyDeviceToDevice, s));
    HIPCK(hipMemcpyAsync(B.pwb, c->pwb, 3 * ns * 8, hipMemcpyDeviceToDevice, s));
    if (B.vel) HIPCK(hipMemcpyAsync(B.vel, c->vel, 3 * ns * 8, hipMemcpyDeviceToDevice, s));
    if (B.bgs) HIPCK(hipMemcpyAsync(B.bgs, c->bgs, 3 * ns * 8, hipMemcpyDeviceToDevice, s));
    if (B.bas) HIPCK(hipMemcpyAsync(B.bas, c->bas, 3 * ns * 8, hipMemcpyDeviceToDevice, s));
    HIPCK(hipMemcpyAsync(B.lmp, c->lmpT, 3 * (size_t)c->nlm * 8, hipMemcpyDeviceToDevice, s));
    return XBA_OK;
}

int xba_stats(xba_ctx *c, double *out, void *stream) {
    if (!c || !out) return XBA_E_ARG;
    hipLaunchKernelGGL(k_stats, dim3((c->nwin + 255) / 256), dim3(256), 0, (hipStream_t)stream, c->nwin, c->wstate, out);
    HIPCK(hipGetLastError());
    return XBA_OK;
}

int xba_last_dx(xba_ctx *c, double *out, void *stream) {
    if (!c || !out) return XBA_E_ARG;
    HIPCK(hipMemcpyAsync(out, c->dx, (size_t)c->nstate * 8, hipMemcpyDeviceToDevice, (hipStream_t)stream));
    return XBA_OK;
}

int xba_linearize(xba_ctx *c, double lam, double *H_out, double *b_out, double *cost_out, double *chi2_out,
                  double *Hinv_out, uint8_t *ok_out, double *imu_chi2_out, void *stream) {
    if (!c || lam < 0.0) return XBA_E_ARG;
    if (!c->uploaded) return XBA_E_STATE;
    c->gn_ready = false;
    hipStream_t s = (hipStream_t)stream;
    if (mixed(c)) {
        KArgs<float> A = kargs<float>(c);
        A.lam_override = lam;
        A.chi2_out = chi2_out;
        A.Hinv_out = Hinv_out;
        A.ok_out = ok_out;
        launch_lin<float>(c, A, s, imu_chi2_out);
    } else {
        KArgs<double> A = kargs<double>(c);
        A.lam_override = lam;
        A.chi2_out = chi2_out;
        A.Hinv_out = Hinv_out;
        A.ok_out = ok_out;
        launch_lin<double>(c, A, s, imu_chi2_out);
    }
    HIPCK(hipGetLastError());
    if (H_out) HIPCK(hipMemcpyAsync(H_out, c->H, (size_t)c->nh * 8, hipMemcpyDeviceToDevice, s));
    if (b_out) HIPCK(hipMemcpyAsync(b_out, c->b, (size_t)c->nstate * 8, hipMemcpyDeviceToDevice, s));
    if (cost_out) {
        hipLaunchKernelGGL(k_cost_out, dim3((c->nwin + 255) / 256), dim3(256), 0, s, c->nwin, c->wstate, cost_out);
        HIPCK(hipGetLastError());
    }
    return XBA_OK;
}

int xba_profile_iteration(xba_ctx *c, double *ms_out, void *stream) {
    if (!c || !ms_out) return XBA_E_ARG;
    if (!c->uploaded) return XBA_E_STATE;
    c->gn_ready = false;
    hipStream_t s = (hipStream_t)stream;
    hipEvent_t ev[XBA_NPROF + 1];
    for (int i = 0; i <= XBA_NPROF; i++) HIPCK(hipEventCreate(&ev[i]));
    DArgs D = dargs(c);
    int gl = std::max(1, std::min((c->nlm + 255) / 256, 4096));
    auto run = [&](auto tag) {
        using T = decltype(tag);
        KArgs<T> A = kargs<T>(c);
        (void)hipEventRecord(ev[0], s);
        launch_lin_kernel<T>(c, A, s, 0);
        (void)hipEventRecord(ev[1], s);
        launch_lin_kernel<T>(c, A, s, 1);
        (void)hipEventRecord(ev[2], s);
        launch_reduce<T>(c, D, s);
        (void)hipEventRecord(ev[3], s);
        launch_terms(c, D, s, 0);
        (void)hipEventRecord(ev[4], s);
        launch_solve<T>(c, D, s);
        (void)hipEventRecord(ev[5], s);
        launch_trial_kernel<T>(c, A, s, 0);
        (void)hipEventRecord(ev[6], s);
        launch_trial_kernel<T>(c, A, s, 1);
        (void)hipEventRecord(ev[7], s);
        launch_terms(c, D, s, 1);
        (void)hipEventRecord(ev[8], s);
        hipLaunchKernelGGL(k_accept, dim3(c->nwin), dim3(256), 0, s, D);
        (void)hipEventRecord(ev[9], s);
        hipLaunchKernelGGL(k_commit_lm<T>, dim3(gl), dim3(256), 0, s, A, (const int32_t *)c->lm_win);
        (void)hipEventRecord(ev[10], s);
    };
    if (mixed(c))
        run(0.0f);
    else
        run(0.0);
    HIPCK(hipGetLastError());
    HIPCK(hipEventSynchronize(ev[XBA_NPROF]));
    for (int i = 0; i < XBA_NPROF; i++) {
        float ms = 0.f;
        HIPCK(hipEventElapsedTime(&ms, ev[i], ev[i + 1]));
        ms_out[i] = ms;
    }
    for (int i = 0; i <= XBA_NPROF; i++) (void)hipEventDestroy(ev[i]);
    return XBA_OK;
}

int xba_plan_info(const xba_ctx *c, int64_t *out) {
    if (!c || !out) return XBA_E_ARG;
    out[0] = c->nchunks; out[1] = c->grid; out[2] = c->Rmax; out[3] = c->fast ? c->fast_lds : c->lin_lds;
    out[4] = c->solve_lds_cap;
    if (c->fast) out[1] = c->grid * 100000 + c->fast_nw * 1000 + (c->nchunks - c->nnarrow);
    out[5] = c->npartH; out[6] = c->npartb; out[7] = c->maxn;
    return XBA_OK;
}

int xba_gn_step(xba_ctx *c, double *dx_out, void *stream) {
    if (!c) return XBA_E_ARG;
    if (!c->uploaded) return XBA_E_STATE;
    if (mixed(c) && !c->fast) return XBA_E_STATE;  // the deterministic mixed family keeps FP32 landmark blocks
    hipStream_t s = (hipStream_t)stream;
    c->gn_ready = false;
    if (mixed(c)) {
        KArgs<float> A = kargs<float>(c);
        A.lam_override = 0.0;
        launch_lin<float>(c, A, s);
    } else {
        KArgs<double> A = kargs<double>(c);
        A.lam_override = 0.0;
        launch_lin<double>(c, A, s);
    }
    hipLaunchKernelGGL(k_gn_solve, dim3(c->nwin), dim3(256), 0, s, dargs(c));
    HIPCK(hipGetLastError());
    if (dx_out) HIPCK(hipMemcpyAsync(dx_out, c->dx, (size_t)c->nstate * 8, hipMemcpyDeviceToDevice, s));
    c->gn_ready = true;
    return XBA_OK;
}

int xba_backward(xba_ctx *c, const double *g_dx, double *grad_z, double *grad_w, void *stream) {
    if (!c || !g_dx || !grad_z || !grad_w) return XBA_E_ARG;
    if (!c->gn_ready) return XBA_E_STATE;
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(k_adj_y, dim3(c->nwin), dim3(256), 0, s, dargs(c), g_dx, c->adj_y);
    const KArgs<double> F = mixed(c) ? fast_args<float>(c, kargs<float>(c)) : kargs<double>(c);
    if (c->nlm > 0)
        hipLaunchKernelGGL(k_adj_obs, dim3((c->nlm + 127) / 128), dim3(128), 0, s, F, (const int32_t *)c->lm_win,
                           (const double *)c->adj_y, grad_z, grad_w);
    HIPCK(hipGetLastError());
    return XBA_OK;
}

}  // extern "C"

namespace {
// Argument checks shared by the Euclidean local-BA entry points (backend/ba.py:713-869).
int np_check(const xba_np_problem *q, int32_t iters, int32_t precision, const double *stats_out) {
    if (!q || iters < 0 || q->nk < 1 || q->nl < 0 || q->N < 0 || !q->poses || !stats_out) return XBA_E_ARG;
    if (precision != XBA_FP64) return XBA_E_ARG;  // the reference solve is FP64 throughout
    if ((q->nl > 0 && !q->pts) || (q->N > 0 && (!q->okf || !q->olm || !q->ouv)) || !q->free_mask) return XBA_E_ARG;
    if (q->pr_c && (!q->pr_rho0 || !q->pr_invsig || !q->pr_valid)) return XBA_E_ARG;
    if (q->n_rot < 0 || (q->n_rot > 0 && (!q->rot_ab || !q->rot_M || !q->rot_sig))) return XBA_E_ARG;
    if (6 * q->nk > XBA_MAX_STATE) return XBA_E_SIZE;
    return XBA_OK;
}

// The problem's fields into the kernel arguments (the workspace pointers are set by the caller).
NpArgs np_args(const xba_np_problem *q, double *dx_hist_out) {
    NpArgs A{};
    A.nk = q->nk; A.nl = q->nl; A.N = q->N; A.n_rot = q->n_rot; A.has_prior = q->pr_c ? 1 : 0;
    A.fx = q->fx; A.cx = q->cx; A.cy = q->cy; A.K = q->huber_k;
    A.pts = q->pts; A.okf = q->okf; A.olm = q->olm; A.ouv = q->ouv; A.free_mask = q->free_mask;
    A.pr_c = q->pr_c; A.pr_rho0 = q->pr_rho0; A.pr_invsig = q->pr_invsig; A.pr_valid = q->pr_valid;
    A.rot_ab = q->rot_ab; A.rot_M = q->rot_M; A.rot_sig = q->rot_sig;
    A.dx_hist = dx_hist_out;
    A.policy = q->lm_policy ? 1 : 0;
    A.rot_h = A.policy;
    A.jstride = 20;
    A.rwoff = 18;
    return A;
}

// Workspace of the Euclidean solve, carved from `base` in 256-byte-aligned regions (base NULL: sizes only).
// Returns the doubles used.  The int32 regions (order N, lstart nl + 1, fill nl, err 1) follow the doubles.
struct NpWs {
    int32_t *order, *lstart, *fill, *err;
};
constexpr int64_t np_round(int64_t n) { return (n + 31) & ~(int64_t)31; }
constexpr int64_t np_ws_doubles(int64_t nk, int64_t nl, int64_t N) {
    return 3 * np_round(16 * nk) + 2 * np_round(3 * nl) + np_round(20 * N) + np_round(N) + np_round(9 * nl) +
           np_round(3 * nl) + np_round(36 * nk * nk) + 2 * np_round(6 * nk) + np_round(NST) +
           np_round((N + 2 * nl + 2 + 1) / 2 + 1);
}
int64_t np_carve(double *base, int nk, int nl, int N, NpArgs &A, NpWs &W) {
    int64_t o = 0;
    auto take = [&](int64_t n) {
        double *p = base ? base + o : nullptr;
        o += np_round(n);
        return p;
    };
    const int64_t n = 6 * (int64_t)nk;
    A.Tcw = take(16 * (int64_t)nk);
    A.Tnew = take(16 * (int64_t)nk);
    A.Tbest = take(16 * (int64_t)nk);
    A.pnew = take(3 * (int64_t)nl);
    A.pbest = take(3 * (int64_t)nl);
    A.J = take(20 * (int64_t)N);
    A.r2 = take(N);
    A.Hlli = take(9 * (int64_t)nl);
    A.bl = take(3 * (int64_t)nl);
    A.H = take(n * n);
    A.bp = take(n);
    A.dx = take(n);
    A.st = take(NST);
    int32_t *ip = (int32_t *)take(((int64_t)N + 2 * (int64_t)nl + 2 + 1) / 2 + 1);
    W.order = ip;
    W.lstart = ip ? ip + N : nullptr;
    W.fill = ip ? ip + N + nl + 1 : nullptr;
    W.err = ip ? ip + N + 2 * (int64_t)nl + 1 : nullptr;
    return o;
}

#ifndef XBA_NP_SCHUR_LDS
#define XBA_NP_SCHUR_LDS 1
#endif
// The reduced system of a local-BA solve (k_np_schur_pairs): in each workgroup's LDS when the upper triangle and
// the gradient fit (a few workgroups, each flushed once into H), else with global atomics.
template <bool ID>
int np_launch_schur(const NpArgs &A, int n, hipStream_t s) {
    const int nblk = (A.nl + NP_LB - 1) / NP_LB;
    if (nblk == 0) return XBA_OK;
    const size_t lds = ((size_t)n * (n + 1) / 2 + n) * sizeof(double);
    constexpr int cap = 150 * 1024;
    if (XBA_NP_SCHUR_LDS && lds <= (size_t)cap) {
        HIPCK(hipFuncSetAttribute((const void *)k_np_schur_pairs<ID, true>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  cap));
        hipLaunchKernelGGL((k_np_schur_pairs<ID, true>), dim3(std::min(nblk, 2 * 256)), dim3(256), lds, s, A, n);
    } else {
        hipLaunchKernelGGL((k_np_schur_pairs<ID, false>), dim3(std::min(nblk, 1024)), dim3(256), 0, s, A, n);
    }
    return XBA_OK;
}

// Dynamic LDS of k_np_solve: the damped n x n matrix with rows padded to a multiple of 4 entries when it fits.
int np_solve_lds(int n) {
    hipFuncAttributes fa;
    if (hipFuncGetAttributes(&fa, (const void *)k_np_solve) != hipSuccess) return -1;
    const int cap = 160 * 1024 - (int)fa.sharedSizeBytes - 256;  // beside its static LDS
    const size_t need = (size_t)n * ((n + 3) & ~3) * sizeof(double);
    if (cap <= 0 || need > (size_t)cap) return 0;
    if (hipFuncSetAttribute((const void *)k_np_solve, hipFuncAttributeMaxDynamicSharedMemorySize, (int)need) !=
        hipSuccess)
        return -1;
    return (int)need;
}

// The LM iterations of the Euclidean solve on a planned problem (A.order / A.lstart ready on the device):
// shared by xba_np_solve and xba_np_solve_noalloc.  Synchronous.
int np_solve_launch(const xba_np_problem *q, NpArgs &A, int32_t iters, double lam0, double *stats_out,
                    double *dx_hist_out, hipStream_t s) {
    const int nk = q->nk, nl = q->nl, N = q->N, n = 6 * nk;
    const double st0[NST] = {lam0, 0, 0, 0, 1, 0, 0, 0, 0, 0, -1, -1, -1, 0, 0, 0};
    HIPCK(hipMemcpyAsync(A.st, st0, sizeof(st0), hipMemcpyHostToDevice, s));
    if (dx_hist_out && iters) {
        const int64_t m = (int64_t)iters * n;
        k_np_fill_f64<<<(int)std::min<int64_t>((m + 255) / 256, 1024), 256, 0, s>>>(dx_hist_out, m, 0.0);
    }
    const int gk = (nk + 255) / 256, gN = (N + 255) / 256, gl = (nl + 255) / 256;
    const int gz = std::min((n * n + 255) / 256, 1024), gp = (3 * nl + 255) / 256;
    const int slds = np_solve_lds(n);
    if (slds < 0) return XBA_E_HIP;
    k_np_inv_poses<<<gk, 256, 0, s>>>(nk, q->poses, A.Tcw);
    if (gN) k_np_r2_only<<<gN, 256, 0, s>>>(A);
    k_np_chi<<<1, 256, 0, s>>>(A, 5);
    for (int it = 0; it < iters; it++) {
        if (gN) k_np_obs<<<gN, 256, 0, s>>>(A);
        if (A.policy) k_np_gate<<<1, 256, 0, s>>>(A, 0);
        if (gl) k_np_lm<<<gl, 256, 0, s>>>(A);
        k_np_zero<<<gz, 256, 0, s>>>(A, n);
        if (gl && np_launch_schur<false>(A, n, s) != XBA_OK) return XBA_E_HIP;
        k_np_solve<<<1, 512, slds, s>>>(A, n, it, slds);
        if (gl) k_np_backsub<<<gl, 256, 0, s>>>(A);
        if (A.policy) {
            k_np_apply<<<1, 256, 0, s>>>(A, n, 0);
        } else {
            if (gN) k_np_trial_r2<<<gN, 256, 0, s>>>(A);
            k_np_accept<<<1, 256, 0, s>>>(A);
            if (gp) k_np_commit_pts<<<gp, 256, 0, s>>>(A);
        }
    }
    if (A.policy) {
        k_np_project<<<gk, 256, 0, s>>>(nk, A.Tcw);
        if (gN) k_np_obs<<<gN, 256, 0, s>>>(A);
        k_np_wstats<<<1, 256, 0, s>>>(A, stats_out);
    } else {
        if (gN) k_np_r2_only<<<gN, 256, 0, s>>>(A);
        k_np_chi<<<1, 256, 0, s>>>(A, 6);
        k_np_stats<<<1, 1, 0, s>>>(A.st, N, stats_out);
    }
    k_np_inv_poses<<<gk, 256, 0, s>>>(nk, A.Tcw, q->poses);
    HIPCK(hipGetLastError());
    HIPCK(hipStreamSynchronize(s));
    return XBA_OK;
}

// The allocation-free entry point's workspace: one array per GPU, for the problem sizes XBA_NOALLOC_MAX_*.  Host
// code never passes the array's name to a kernel: its per-device address comes from hipGetSymbolAddress, taken once
// per device (with that device current) on the entry point's first call there.
constexpr int64_t NP_SCRATCH_DOUBLES = np_ws_doubles(XBA_NOALLOC_MAX_NK, XBA_NOALLOC_MAX_NL, XBA_NOALLOC_MAX_N);
static __device__ __align__(256) double xba_np_scratch[NP_SCRATCH_DOUBLES];
#define XBA_MAX_DEVICES 64
std::mutex g_np_scratch_mu;
double *g_np_scratch[XBA_MAX_DEVICES];

int np_scratch_ptr(double **out) {
    int dev = -1;
    HIPCK(hipGetDevice(&dev));
    if (dev < 0 || dev >= XBA_MAX_DEVICES) return XBA_E_ARG;
    std::lock_guard<std::mutex> lock(g_np_scratch_mu);
    if (!g_np_scratch[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(xba_np_scratch));
        if (e != hipSuccess || !p) return XBA_E_HIP;
        g_np_scratch[dev] = (double *)p;
    }
    *out = g_np_scratch[dev];
    return XBA_OK;
}
}  // namespace

extern "C" {

int xba_np_solve(const xba_np_problem *q, int32_t iters, double lam0, int32_t precision, double *stats_out,
                 double *dx_hist_out, void *stream) {
    const int rc0 = np_check(q, iters, precision, stats_out);
    if (rc0 != XBA_OK) return rc0;
    const int nk = q->nk, nl = q->nl, N = q->N;
    hipStream_t s = (hipStream_t)stream;

    // host planning: observations grouped by landmark (stable), index validation
    std::vector<int32_t> okf(N), olm(N), order(N), lstart(nl + 1, 0), rab(2 * (size_t)q->n_rot);
    if (N) {
        HIPCK(hipMemcpyAsync(okf.data(), q->okf, sizeof(int32_t) * N, hipMemcpyDeviceToHost, s));
        HIPCK(hipMemcpyAsync(olm.data(), q->olm, sizeof(int32_t) * N, hipMemcpyDeviceToHost, s));
    }
    if (q->n_rot)
        HIPCK(hipMemcpyAsync(rab.data(), q->rot_ab, sizeof(int32_t) * rab.size(), hipMemcpyDeviceToHost, s));
    HIPCK(hipStreamSynchronize(s));
    for (int k = 0; k < N; k++) {
        if (okf[k] < 0 || okf[k] >= nk || olm[k] < 0 || olm[k] >= nl) return XBA_E_RANGE;
        lstart[olm[k] + 1]++;
    }
    for (int f = 0; f < 2 * q->n_rot; f++)
        if (rab[f] < 0 || rab[f] >= nk) return XBA_E_RANGE;
    for (int j = 0; j < nl; j++) lstart[j + 1] += lstart[j];
    {
        std::vector<int32_t> fill(lstart.begin(), lstart.end() - 1);
        for (int k = 0; k < N; k++) order[fill[olm[k]]++] = k;
    }

    NpArgs A = np_args(q, dx_hist_out);
    NpWs W{};
    const int64_t need = np_carve(nullptr, nk, nl, N, A, W);
    double *ws = nullptr;
    if (hipMalloc(&ws, sizeof(double) * need) != hipSuccess) return XBA_E_HIP;
    np_carve(ws, nk, nl, N, A, W);
    A.order = W.order;
    A.lstart = W.lstart;
    int rc = XBA_OK;
    if (N && hipMemcpyAsync(W.order, order.data(), sizeof(int32_t) * N, hipMemcpyHostToDevice, s) != hipSuccess)
        rc = XBA_E_HIP;
    if (rc == XBA_OK &&
        hipMemcpyAsync(W.lstart, lstart.data(), sizeof(int32_t) * (nl + 1), hipMemcpyHostToDevice, s) != hipSuccess)
        rc = XBA_E_HIP;
    if (rc == XBA_OK) rc = np_solve_launch(q, A, iters, lam0, stats_out, dx_hist_out, s);
    (void)hipStreamSynchronize(s);
    (void)hipFree(ws);
    return rc;
}

int xba_np_solve_noalloc(const xba_np_problem *q, int32_t iters, double lam0, int32_t precision, double *stats_out,
                         double *dx_hist_out, void *stream) {
    const int rc0 = np_check(q, iters, precision, stats_out);
    if (rc0 != XBA_OK) return rc0;
    const int nk = q->nk, nl = q->nl, N = q->N;
    if (nk > XBA_NOALLOC_MAX_NK || nl > XBA_NOALLOC_MAX_NL || N > XBA_NOALLOC_MAX_N) return XBA_E_SIZE;
    hipStream_t s = (hipStream_t)stream;
    double *base = nullptr;
    const int rcs = np_scratch_ptr(&base);
    if (rcs != XBA_OK) return rcs;
    NpArgs A = np_args(q, dx_hist_out);
    NpWs W{};
    if (np_carve(nullptr, nk, nl, N, A, W) > NP_SCRATCH_DOUBLES) return XBA_E_SIZE;
    np_carve(base, nk, nl, N, A, W);
    A.order = W.order;
    A.lstart = W.lstart;
    // device planning: counts (the array keeps the previous call's contents, so it is cleared first), offsets,
    // scatter, ordered segments; the index checks read back once
    const int nrot2 = 2 * q->n_rot, gc = (std::max(N, nrot2) + 255) / 256;
    k_np_fill_i32<<<(nl + 1 + 255) / 256, 256, 0, s>>>(W.lstart, nl + 1, 0);
    k_np_fill_i32<<<1, 64, 0, s>>>(W.err, 1, 0);
    if (gc) k_np_plan_count<<<gc, 256, 0, s>>>(N, nk, nl, q->okf, q->olm, nrot2, q->rot_ab, W.lstart, W.err);
    k_np_plan_scan<<<1, 1024, 0, s>>>(nl, W.lstart, W.fill, W.err);
    if (N) k_np_plan_scatter<<<(N + 255) / 256, 256, 0, s>>>(N, nl, q->olm, W.fill, W.order, W.err);
    if (nl) k_np_plan_sort<<<(nl + 255) / 256, 256, 0, s>>>(nl, W.lstart, W.order, W.err);
    HIPCK(hipGetLastError());
    int32_t herr = 0;
    HIPCK(hipMemcpyAsync(&herr, W.err, sizeof(herr), hipMemcpyDeviceToHost, s));
    HIPCK(hipStreamSynchronize(s));
    if (herr) return XBA_E_RANGE;
    return np_solve_launch(q, A, iters, lam0, stats_out, dx_hist_out, s);
}

int xba_np_id_solve(const xba_npid_problem *q, int32_t iters, double lam0, int32_t precision, double *stats_out,
                    double *dx_hist_out, void *stream) {
    if (!q || iters < 0 || q->nk < 1 || q->nl < 0 || q->N < 0 || !q->poses || !stats_out) return XBA_E_ARG;
    if (precision != XBA_FP64) return XBA_E_ARG;  // the reference solve is FP64 throughout
    if ((q->nl > 0 && (!q->rho || !q->mbear || !q->hidx)) || (q->N > 0 && (!q->okf || !q->olm || !q->ouv)) ||
        !q->free_mask)
        return XBA_E_ARG;
    if (q->pr_m0 && (!q->pr_rho0 || !q->pr_wbear || !q->pr_invsig || !q->pr_valid)) return XBA_E_ARG;
    if (q->n_rot < 0 || (q->n_rot > 0 && (!q->rot_ab || !q->rot_M || !q->rot_sig))) return XBA_E_ARG;
    const int nk = q->nk, nl = q->nl, N = q->N, n = 6 * nk;
    if (n > XBA_MAX_STATE) return XBA_E_SIZE;
    hipStream_t s = (hipStream_t)stream;

    // host planning: observations grouped by landmark (stable), index validation
    std::vector<int32_t> okf(N), olm(N), order(N), lstart(nl + 1, 0), rab(2 * (size_t)q->n_rot), hidx(nl);
    if (N) {
        HIPCK(hipMemcpyAsync(okf.data(), q->okf, sizeof(int32_t) * N, hipMemcpyDeviceToHost, s));
        HIPCK(hipMemcpyAsync(olm.data(), q->olm, sizeof(int32_t) * N, hipMemcpyDeviceToHost, s));
    }
    if (nl) HIPCK(hipMemcpyAsync(hidx.data(), q->hidx, sizeof(int32_t) * nl, hipMemcpyDeviceToHost, s));
    if (q->n_rot)
        HIPCK(hipMemcpyAsync(rab.data(), q->rot_ab, sizeof(int32_t) * rab.size(), hipMemcpyDeviceToHost, s));
    HIPCK(hipStreamSynchronize(s));
    for (int k = 0; k < N; k++) {
        if (okf[k] < 0 || okf[k] >= nk || olm[k] < 0 || olm[k] >= nl) return XBA_E_RANGE;
        lstart[olm[k] + 1]++;
    }
    bool any_ext = false;
    for (int j = 0; j < nl; j++) {
        if (hidx[j] >= nk || hidx[j] < -1) return XBA_E_RANGE;
        any_ext |= hidx[j] < 0;
    }
    if (any_ext && !q->hpose_ext) return XBA_E_ARG;
    for (int f = 0; f < 2 * q->n_rot; f++)
        if (rab[f] < 0 || rab[f] >= nk) return XBA_E_RANGE;
    for (int j = 0; j < nl; j++) lstart[j + 1] += lstart[j];
    {
        std::vector<int32_t> fill(lstart.begin(), lstart.end() - 1);
        for (int k = 0; k < N; k++) order[fill[olm[k]]++] = k;
    }

    std::vector<void *> owned;
    auto release = [&]() {
        for (void *p : owned) (void)hipFree(p);
    };
    NpArgs A{};
    A.nk = nk; A.nl = nl; A.N = N; A.n_rot = q->n_rot; A.has_prior = q->pr_m0 ? 1 : 0; A.rot_h = 1;
    A.fx = q->fx; A.cx = q->cx; A.cy = q->cy; A.K = q->huber_k; A.rho_floor = q->rho_floor;
    A.okf = q->okf; A.olm = q->olm; A.ouv = q->ouv; A.free_mask = q->free_mask;
    A.hidx = q->hidx; A.hext = q->hpose_ext; A.rho = q->rho; A.mb = q->mbear;
    A.pr_m0 = q->pr_m0; A.pr_rho0 = q->pr_rho0; A.pr_wbear = q->pr_wbear; A.pr_invsig = q->pr_invsig;
    A.pr_valid = q->pr_valid;
    A.rot_ab = q->rot_ab; A.rot_M = q->rot_M; A.rot_sig = q->rot_sig;
    A.dx_hist = dx_hist_out;
    A.policy = q->lm_policy ? 1 : 0;
    A.jstride = 32;
    A.rwoff = 30;
    A.Tcw = dalloc<double>(16 * (size_t)nk, owned);
    A.Tnew = dalloc<double>(16 * (size_t)nk, owned);
    A.Twc = dalloc<double>(16 * (size_t)nk, owned);
    A.Twn = dalloc<double>(16 * (size_t)nk, owned);
    A.rhon = dalloc<double>(nl, owned);
    A.mbn = dalloc<double>(3 * (size_t)nl, owned);
    A.J = dalloc<double>(32 * (size_t)N, owned);
    A.r2 = dalloc<double>(N, owned);
    A.Hlli = dalloc<double>(9 * (size_t)nl, owned);
    A.bl = dalloc<double>(3 * (size_t)nl, owned);
    A.H = dalloc<double>((size_t)n * n, owned);
    A.bp = dalloc<double>(n, owned);
    A.dx = dalloc<double>(n, owned);
    A.st = dalloc<double>(NST, owned);
    A.Tbest = dalloc<double>(16 * (size_t)nk, owned);
    A.Wbest = dalloc<double>(16 * (size_t)nk, owned);
    A.rbest = dalloc<double>(nl, owned);
    A.mbest = dalloc<double>(3 * (size_t)nl, owned);
    int32_t *d_order = dalloc<int32_t>(N, owned), *d_lstart = dalloc<int32_t>(nl + 1, owned);
    if (owned.size() != 20) {
        release();
        return XBA_E_HIP;
    }
    A.order = d_order;
    A.lstart = d_lstart;
    const double st0[NST] = {lam0, 0, 0, 0, 1, 0, 0, 0, 0, 0, -1, -1, -1, 0, 0, 0};
    const int slds = np_solve_lds(n);
    if (slds < 0) {
        release();
        return XBA_E_HIP;
    }
    auto run = [&]() -> int {
        if (N) HIPCK(hipMemcpyAsync(d_order, order.data(), sizeof(int32_t) * N, hipMemcpyHostToDevice, s));
        HIPCK(hipMemcpyAsync(d_lstart, lstart.data(), sizeof(int32_t) * (nl + 1), hipMemcpyHostToDevice, s));
        HIPCK(hipMemcpyAsync(A.st, st0, sizeof(st0), hipMemcpyHostToDevice, s));
        HIPCK(hipMemcpyAsync(A.Twc, q->poses, 16 * sizeof(double) * nk, hipMemcpyDeviceToDevice, s));
        if (dx_hist_out && iters)
            HIPCK(hipMemsetAsync(dx_hist_out, 0, sizeof(double) * (size_t)iters * n, s));
        const int gk = (nk + 255) / 256, gN = (N + 255) / 256, gl = (nl + 255) / 256;
        const int gz = std::min((n * n + 255) / 256, 1024);
        const int gc = (int)((std::max<int64_t>(16 * (int64_t)nk, 3 * (int64_t)nl) + 255) / 256);
        k_np_inv_poses<<<gk, 256, 0, s>>>(nk, A.Twc, A.Tcw);
        if (gN) k_id_r2<<<gN, 256, 0, s>>>(A, 0);
        k_np_chi<<<1, 256, 0, s>>>(A, 5);
        for (int it = 0; it < iters; it++) {
            if (gN) k_id_obs<<<gN, 256, 0, s>>>(A);
            if (A.policy) k_np_gate<<<1, 256, 0, s>>>(A, 1);
            if (gl) k_id_lm<<<gl, 256, 0, s>>>(A);
            k_np_zero<<<gz, 256, 0, s>>>(A, n);
            if (gl && np_launch_schur<true>(A, n, s) != XBA_OK) return XBA_E_HIP;
            k_np_solve<<<1, 512, slds, s>>>(A, n, it, slds);
            k_np_inv_poses<<<gk, 256, 0, s>>>(nk, A.Tnew, A.Twn);
            if (gl) k_id_backsub<<<gl, 256, 0, s>>>(A);
            if (A.policy) {
                k_np_apply<<<1, 256, 0, s>>>(A, n, 1);
            } else {
                if (gN) k_id_r2<<<gN, 256, 0, s>>>(A, 1);
                k_np_accept<<<1, 256, 0, s>>>(A);
                k_id_commit<<<gc, 256, 0, s>>>(A);
            }
        }
        if (A.policy) {
            k_np_project<<<gk, 256, 0, s>>>(nk, A.Tcw);
            k_np_inv_poses<<<gk, 256, 0, s>>>(nk, A.Tcw, A.Twc);
            if (gN) k_id_obs<<<gN, 256, 0, s>>>(A);
            k_np_wstats<<<1, 256, 0, s>>>(A, stats_out);
        } else {
            if (gN) k_id_r2<<<gN, 256, 0, s>>>(A, 0);
            k_np_chi<<<1, 256, 0, s>>>(A, 6);
            k_np_stats<<<1, 1, 0, s>>>(A.st, N, stats_out);
        }
        HIPCK(hipMemcpyAsync(q->poses, A.Twc, 16 * sizeof(double) * nk, hipMemcpyDeviceToDevice, s));
        HIPCK(hipGetLastError());
        HIPCK(hipStreamSynchronize(s));
        return XBA_OK;
    };
    const int rc = run();
    release();
    return rc;
}

int xba_build_pairs(const int32_t *olm_sorted, int64_t n_obs, int32_t nl, int32_t *offs, int32_t *pa, int32_t *pb,
                    int64_t cap, int64_t *npair, void *stream) {
    if (n_obs < 0 || nl < 0 || !offs || !npair || (n_obs && !olm_sorted) || (pa && (!pb || cap < 0)))
        return XBA_E_ARG;
    if (n_obs > INT32_MAX) return XBA_E_ARG;   // observation indices are int32
    hipStream_t s = (hipStream_t)stream;
    const int32_t nb = (int32_t)(((int64_t)nl + 1 + PAIRS_BS - 1) / PAIRS_BS);   // blocks over l in [0, nl]
    // workspace: err (padded to 8 bytes), bsum (nb + 1), base (nl + 1)
    char *ws = nullptr;
    const size_t bytes = 8 + sizeof(int64_t) * ((size_t)nb + 1 + (size_t)nl + 1);
    if (hipMalloc(&ws, bytes) != hipSuccess) return XBA_E_HIP;
    int32_t *err = (int32_t *)ws;
    int64_t *bsum = (int64_t *)(ws + 8), *base = bsum + nb + 1;
    int rc = XBA_OK;
    auto run = [&]() -> int {
        HIPCK(hipMemsetAsync(err, 0, sizeof(int32_t), s));
        const int64_t span = std::max<int64_t>(n_obs, (int64_t)nl + 1);
        const int g = (int)std::min<int64_t>((span + 255) / 256, 4096);
        k_pairs_offs<<<g, 256, 0, s>>>(olm_sorted, n_obs, nl, offs, err);
        k_pairs_bsum<<<nb, PAIRS_BS, 0, s>>>(offs, nl, bsum, err);
        k_pairs_bscan<<<1, PAIRS_BS, 0, s>>>(bsum, nb, err);
        k_pairs_base<<<nb, PAIRS_BS, 0, s>>>(offs, nl, bsum, base, err);
        HIPCK(hipGetLastError());
        int32_t herr = 0;
        int64_t tot = 0;
        HIPCK(hipMemcpyAsync(&herr, err, sizeof(int32_t), hipMemcpyDeviceToHost, s));
        HIPCK(hipMemcpyAsync(&tot, base + nl, sizeof(int64_t), hipMemcpyDeviceToHost, s));
        HIPCK(hipStreamSynchronize(s));
        if (herr) return XBA_E_RANGE;
        *npair = tot;
        if (!pa) return XBA_OK;
        if (cap < tot) return XBA_E_SIZE;
        if (tot) {
            const int ge = (int)std::min<int64_t>((tot + 255) / 256, 65536);
            k_pairs_expand<<<ge, 256, 0, s>>>(offs, base, nl, tot, pa, pb);
            HIPCK(hipGetLastError());
        }
        HIPCK(hipStreamSynchronize(s));
        return XBA_OK;
    };
    rc = run();
    (void)hipFree(ws);
    return rc;
}

}  // extern "C"
