// Internal declarations of libxba: work-list records, per-window records and the kernel argument block.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/xba.h"

#define XBA_WAVE 64

// One work item of the landmark kernels: a contiguous landmark range of one window whose observations touch only
// R slots, its tile slots: [lo, lo + R) when smap < 0, else the ascending window slots chunk_slots[smap .. smap + R)
// (work items of landmarks spanning more slots than a tile holds, which touch few of them).  Its reduced-system
// partial lives at partH (R(R+1)/2 lower-triangular 6x6 blocks over the tile slots) and partb (6R entries) of the
// partial buffers.  The multi-wave kernels walk it in the passes [p0, p1) of the pass list (each pass: the
// landmarks [pass_l0[p], pass_l0[p + 1])).
struct Chunk {
    int32_t win, l0, l1, lo, R, p0, p1, smap;
    int64_t partH, partb;
};

// Window slot of tile slot t of a work item.
__host__ __device__ inline int chunk_slot(const Chunk &ch, const int32_t *chunk_slots, int t) {
    return ch.smap < 0 ? ch.lo + t : chunk_slots[ch.smap + t];
}

// Tile slot of window slot s in a work item, or -1 when the work item does not touch it.
__host__ __device__ inline int chunk_tile(const Chunk &ch, const int32_t *chunk_slots, int s) {
    if (ch.smap < 0) return (s >= ch.lo && s < ch.lo + ch.R) ? s - ch.lo : -1;
    for (int t = 0; t < ch.R; t++) {
        const int v = chunk_slots[ch.smap + t];
        if (v == s) return t;
        if (v > s) break;
    }
    return -1;
}

// Per-window record.
struct WinInfo {
    int32_t slot0, P, lm0, L, n, c0, c1, imu0, imu1, pblk0, pblk1, prow0, prow1, np, vi;
    int32_t big;     // index into the tiled-solve window list, -1 when the one-workgroup solve handles the window
    int32_t c2, c3;  // second work-item range of the window (landmarks coupling more slots than the narrow tile)
    int32_t pair0;   // first slot pair of the window in the reduction lists (pairs s2 >= s1, s2 (s2 + 1) / 2 + s1)
    int32_t imu_ncol;  // colours of the window's IMU factors
    int64_t st0;   // offset of the window's state vector (dx, b)
    int64_t h0;    // offset of the window's n x n matrices
    int64_t prS0;  // offset of the window's prior S
};

// Per-window scalar state (double[nwin * NWS]).
enum {
    WS_LAM = 0,
    WS_COSTVIS,
    WS_COSTREST,
    WS_CTOT,
    WS_CNEW,
    WS_COST0,
    WS_ACC,
    WS_CHOLOK,
    WS_ACCEPT,
    WS_NREF,
    WS_HAVE0,
    WS_CUR,
    WS_ITERS,
    WS_COSTTRIAL_REST,
    WS_BIGOK,       // tiled Cholesky of the window: 1 while every pivot so far was positive and finite
    WS_BIGCONV,     // tiled solve: 1 once the refinement has converged in this LM iteration
    WS_FALLBACK,    // mixed precision: 1 when this iteration's FP32 factor failed or its refinement did not converge,
                    // so the FP64 solve (k_solve<double> in fallback mode) produced the update
    WS_NFALLBACK,   // FP64 fallbacks since the last upload
    WS_CONV,        // 1 when this iteration's solve met the refinement criterion (always 1 in FP64)
    NWS = 20
};

// A window whose reduced system is factored by the tiled multi-workgroup Cholesky: its matrix is padded to np, a
// multiple of XBA_TILE, with an identity block on the padding, and stored row-major at f0 of the tiled-factor
// buffer, followed by the inverses of the factor's diagonal tiles (np x XBA_TILE, row-major); its Jacobi scaling
// (or ones) at s0 of the scale buffer; its refinement vectors (solution x, residual r, per-workgroup residual and
// right-hand-side maxima) at v0 of the refinement buffer (3 np entries).
#define XBA_TILE 64
// Mixed-precision refinement stops once the last FP64-accumulated correction is at most this fraction of the
// solution (max norms), or once the FP64 residual is at most 1e-15 of the right-hand side.  The error left after
// that solve is the correction times the refinement's contraction factor (below 1 whenever it converges), so at
// most 1e-8 of the solution, two orders below the mixed path's 1e-6 error target.
#define XBA_REFINE_TOL 1e-8
#define XBA_RES_ROWS 4  // rows per workgroup of the refinement residual
struct BigWin {
    int32_t win, n, np, nb;  // window, state dimension, padded dimension, tiles per side (np / XBA_TILE)
    int64_t f0, s0, v0;
};

// Packed observation word: slot (bits 0-11) | cam << 12 | valid << 13 | tile slot << 14 (bits 14-19, the slot's
// index in its landmark's work-item tile).  Packed landmark word: host slot (bits 0-11) | mask << 12 | host tile
// slot << 13 (bits 13-18).
__host__ __device__ inline int32_t pack_obs(int slot, int cam, int valid, int tile) {
    return slot | (cam << 12) | (valid << 13) | (tile << 14);
}
__device__ inline int obs_tile(int info) { return (info >> 14) & 63; }
__device__ inline int lm_tile(int li) { return (li >> 13) & 63; }

template <typename T>
struct KArgs {
    int32_t nchunks, nwin, nlm, nobs;
    const Chunk *chunks;
    const WinInfo *wins;
    const int32_t *ob_start;   // (nlm+1)
    const int32_t *lm_info;    // (nlm) packed landmark word
    const int32_t *ob_info;    // (nobs) packed observation word
    const int32_t *slot_off;   // (nslot) offset of the slot's block inside its window's state vector
    const double *z;           // (nobs,2) measurements stay FP64: residuals are formed in FP64
    const T *isig;             // (nobs)
    double *lmp;               // (nlm,3) landmark state (FP64)
    double *lmp_trial;         // (nlm,3)
    const double *rho_meas;    // (nlm)
    const T *rho_isig;         // (nlm)
    const double *Rwb, *pwb, *R2, *p2;  // (nslot,9), (nslot,3)
    const double *dx;          // (sum n)
    const double *win_par;     // (nwin, XBA_WPAR)
    T *partH, *partb;
    T *partH_lo, *partb_lo;    // FP32 partials: the low parts of the compensated tile sums
    double *partH64, *partb64; // partials of the multi-wave kernels (FP64 tiles)
    double *partc, *partc2;
    T *lmHi, *lmbl;            // (nlm,6) symmetric damped inverse, (nlm,3)
    uint8_t *lmok;
    double *wstate;            // (nwin, NWS)
    double lam_override;       // >= 0: damping for every window (linearize); < 0: per-window lambda
    double *chi2_out;          // (nobs) or null
    double *Hinv_out;          // (nlm,9) or null
    uint8_t *ok_out;           // (nlm) or null
    double *adj_gz, *adj_gw;   // adjoint outputs (nobs,2), (nobs) or null
    const double *adj_y;       // adjoint solve H^-1 g (sum n) or null
    const int32_t *pass_l0;    // multi-wave kernels: first landmark of every pass (npass + 1)
    const int32_t *sched;      // multi-wave kernels: work items in the order they are handed out
    int32_t *sched_ctr;        // multi-wave kernels: per launch kind, a hand-out counter and a finished counter
    double *pass_cost;         // multi-wave kernels: robust cost of every pass (summed per work item in pass order)
    const double *ob_w;        // (nobs) per-observation weight multiplier, or null (1)
    const int32_t *chunk_slots; // tile slot lists of the work items with smap >= 0
    int32_t only_fallback;     // 1: only the work items of windows flagged WS_FALLBACK (the FP64 re-assembly)
};
