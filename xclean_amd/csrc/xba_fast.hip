// Throughput kernels of the window solve (the default for every precision; xba_options.deterministic selects the
// single-wave kernels of xba_lin.hip instead).
//
// A workgroup of NW wavefronts owns one work item (a landmark chunk of one window) at a time and shares one FP64
// LDS tile of the chunk's reduced-system partial; each wavefront takes a contiguous share of the chunk's landmarks
// and walks it in passes of up to 64 observations (several landmarks per pass, one lane per observation), then
// switches to one lane per landmark for the 3x3 algebra, then spreads each landmark's Schur blocks over the lanes.
// Tile updates are LDS atomics (ds_add_f64); per-landmark sums over observations run in the working precision.
//
// Reference arithmetic (paths relative to the reference's src/xclean/): as xba_lin.hip --
//   residual / Jacobians vio/window.py:125-146, 164-263;  Huber :284-288;  build_system :324-462;
//   back_substitute :496-519;  landmark update rule :533-540;  visual_cost :466-493.
#include "xba_internal.h"

namespace xba {

#define FB_OBS 64   // observations per pass (one lane each)
#define FB_LM 32    // landmarks per pass
#define FB_HLP 64   // sum over a pass's landmarks of their touched slots
#define TBLK 37     // tile block stride (doubles): 36 entries + 1 pad, spreads a wave's atomics over the banks
#define FB_LREC 12  // per landmark (FP64): ok, L^-1 (00 10 11 20 21 22), g = L^-1 bl, unique-target flag
#define FB_PAIRS 512  // sum over a pass's landmarks of (k - 1)(k - 2) / 2 (their pairs of target slots)
// per-landmark accumulator: Hll (6), bl (3), active rows, active rows off the host slot
enum { LA_H = 0, LA_B = 6, LA_NACT = 9, LA_NX = 10, LA_N = 11 };
// LR_UNIQ: 1 when each of the landmark's active cross rows has a target slot of its own (then each row alone forms
// its slot's U entry)
enum { LR_OK = 0, LR_L = 1, LR_G = 7, LR_UNIQ = 10 };
// trial kernel, per landmark of a pass (FP64): state at the current point, trial state, symmetric damped inverse
// (00 01 02 11 12 22), bl, inverse-depth prior (measurement, inverse sigma)
enum { TL_CUR = 0, TL_NEW = 3, TL_HI = 6, TL_BL = 12, TL_RM = 15, TL_RI = 16, TL_N = 18 };
#define TL_LMOK (1 << 20)  // in the trial's per-landmark word (packed landmark word | this): the landmark is solvable

// Orders the wavefront's LDS traffic: LDS requests of one wavefront complete in issue order, so a compiler fence
// plus a wait on the LDS counter makes every lane's earlier LDS writes visible to every lane's later reads.
#define WAVE_SYNC()                                          \
    do {                                                     \
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront"); \
        __builtin_amdgcn_s_waitcnt(0xc07f);                  \
        __builtin_amdgcn_wave_barrier();                     \
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront"); \
    } while (0)

struct FastLayout {
    int sc, sc64, wc, wd, tile, btile, ctl, wave, wave_stride, total;
    // offsets inside one wavefront's region
    int o_oblm, o_rcq, o_u, o_lmask, o_lobs0, o_lnobs, o_loff, o_poff, o_slots, o_hlp, o_lrec, o_ldbl,
        o_lrow, o_ptab;
};

// rts: bytes per record entry (the observation record's storage type); everything else is FP64.
__host__ __device__ inline FastLayout fast_layout(int R, int rts, int nw, bool trial) {
    FastLayout L;
    int o = 0;
    const int nc = trial ? 2 : 1, ts = 8;
    L.sc = o;    o = align16(o + nc * R * SC_STRIDE * ts);
    L.sc64 = o;  o = align16(o + (trial ? 0 : R * SC64_STRIDE * 8));  // the trial reads FP64 poses from sc
    L.wc = o;    o = align16(o + 32 * ts);
    L.wd = o;    o = align16(o + 8 * 8);
    L.tile = o;  o = align16(o + (trial ? 6 * R * ts : (R * (R + 1) / 2) * TBLK * 8));  // trial: dp per slot
    L.btile = o; o = align16(o + (trial ? 0 : 6 * R * 8));
    L.ctl = o;   o = align16(o + 4 * 4);
    L.wave = o;
    int w = 0;
    L.o_oblm = w;  w = align16(w + FB_OBS * 4);
    L.o_rcq = w;   // (the assembly kernels place it in U's storage below; the trial keeps its costs in registers)
    L.o_u = w;     // (unused)
    L.o_lmask = w; w = align16(w + FB_LM * 4);
    L.o_lobs0 = w; w = align16(w + FB_LM * 4);
    L.o_lnobs = w; w = align16(w + FB_LM * 4);
    L.o_loff = w;  w = align16(w + FB_LM * 4);
    L.o_poff = w;  w = align16(w + FB_LM * 4);
    L.o_slots = w; w = align16(w + (trial ? 0 : FB_HLP * 4));
    L.o_hlp = w;   w = align16(w + (trial ? 0 : FB_HLP * 18 * ts));
    L.o_lrec = w;  w = align16(w + (trial ? 0 : FB_LM * FB_LREC * ts));
    L.o_ldbl = w;  w = align16(w + (trial ? FB_LM * TL_N * 8 : 0));
    L.o_lrow = w;  // (trial: unused)
    if (!trial) {
        // the observation costs and the per-landmark sums (Hll, bl, active count) live from the observation phase
        // to the landmark phase, U from after it: they share U's storage (one wavefront's phases are ordered by its
        // wave syncs), so that eight wavefronts of a 16-slot tile fit the LDS
        static_assert(FB_OBS * 8 + FB_LM * LA_N * 8 <= FB_HLP * 18 * 8, "U storage holds the phase-1 arrays");
        L.o_rcq = L.o_hlp;
        L.o_lrow = L.o_hlp + FB_OBS * 8;
    }
    L.o_ptab = w;  w = align16(w + (trial ? 0 : FB_PAIRS * 2));            // per slot pair: first entry, si, sj
    L.wave_stride = w;
    L.total = o + nw * w;
    return L;
}

// Wavefront scans and sums with DPP lane moves (VALU only: no LDS round trip and no per-lane shuffle addresses held in
// registers): lane i adds lane i - d for d = 1, 2, 4, 8 inside its row of 16 lanes, then the row-15 broadcast adds
// row 0's (row 2's) total into row 1 (row 3) and the lane-31 broadcast adds the first half's total into rows 2 and 3.
template <int CTRL, int ROWS>
__device__ inline int dpp_i(int v) {
    return __builtin_amdgcn_update_dpp(0, v, CTRL, ROWS, 0xf, true);
}

// Inclusive prefix sum over the wavefront.
__device__ inline int wave_incl_scan(int v, int lane) {
    const int r = lane & 15, row = lane >> 4;
    int t;
    t = dpp_i<0x111, 0xf>(v); if (r >= 1) v += t;   // row_shr:1
    t = dpp_i<0x112, 0xf>(v); if (r >= 2) v += t;   // row_shr:2
    t = dpp_i<0x114, 0xf>(v); if (r >= 4) v += t;   // row_shr:4
    t = dpp_i<0x118, 0xf>(v); if (r >= 8) v += t;   // row_shr:8
    t = dpp_i<0x142, 0xa>(v); if (row & 1) v += t;  // row_bcast:15 into rows 1, 3
    t = dpp_i<0x143, 0xc>(v); if (row >= 2) v += t; // row_bcast:31 into rows 2, 3
    return v;
}

template <int CTRL, int ROWS>
__device__ inline double dpp_d(double v);

// Sum over the wavefront in a fixed order (the scan above, lane 63's total), the result in every lane.
__device__ inline double wave_sum(double v, int lane) {
    const int r = lane & 15, row = lane >> 4;
    double t;
    t = dpp_d<0x111, 0xf>(v); if (r >= 1) v += t;
    t = dpp_d<0x112, 0xf>(v); if (r >= 2) v += t;
    t = dpp_d<0x114, 0xf>(v); if (r >= 4) v += t;
    t = dpp_d<0x118, 0xf>(v); if (r >= 8) v += t;
    t = dpp_d<0x142, 0xa>(v); if (row & 1) v += t;
    t = dpp_d<0x143, 0xc>(v); if (row >= 2) v += t;
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)b, 63), hi = __builtin_amdgcn_readlane((int)(b >> 32), 63);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

template <typename T>
__device__ inline void lds_add(T *p, T v) {
    __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// A pass's plan (the B landmarks from lb, planned on the host so that their observations fit one lane each and
// their touched slots fit the pass's U storage) and its landmarks' row offsets (lane < B: o0, o1 = ob_start of
// landmark lb + lane and the next), loaded one pass ahead: each wavefront takes its next pass and issues these loads
// while it works on the current one.
struct PassAhead {
    int p, lb, B, o0, o1;
};

__device__ inline int next_pass(int *ctl, int lane);

// Takes the wavefront's next pass of the work item and issues the loads of its plan entries (the scalar part; the
// row offsets follow in pass_rows once the plan has arrived).
__device__ inline void pass_take(const KArgs<double> &A, const Chunk &ch, int *ctl, int lane, PassAhead &n) {
    n.p = next_pass(ctl, lane);
    n.lb = 0;
    n.B = 0;
    if (n.p < ch.p1) {
        n.lb = A.pass_l0[n.p];
        n.B = (n.p + 1 < ch.p1 ? A.pass_l0[n.p + 1] : ch.l1) - n.lb;
    }
}

__device__ inline void pass_rows(const int32_t *ob_start, int lane, PassAhead &n) {
    const bool in = lane < n.B;
    n.o0 = in ? ob_start[n.lb + lane] : 0;
    n.o1 = in ? ob_start[n.lb + lane + 1] : 0;
}

// One pass from its plan: per landmark its first record row and observation count, per record row its landmark.
// Returns the pass's observation count.
__device__ inline int load_pass(const PassAhead &pa, int lane, int *lobs0, int *lnobs, int *ob_lm, int *base) {
    const int B = pa.B;
    const bool in = lane < B;
    const int o0 = pa.o0;
    const int n = in ? (pa.o1 - o0) : 0;
    const int S = wave_incl_scan(n, lane);
    if (in) {
        lobs0[lane] = S - n;
        lnobs[lane] = n;
        for (int q = 0; q < n; q++) ob_lm[S - n + q] = lane;
    }
    *base = __builtin_amdgcn_readlane(o0, 0);  // the pass's observation rows are contiguous from here
    return __builtin_amdgcn_readlane(S, B - 1);
}

// Segmented sums over consecutive lanes (a segment: one landmark's observation rows), accumulated into the
// segment's last lane with DPP lane moves (VALU only, no LDS traffic): lane i adds lane i - d for d = 1, 2, 4, 8
// inside its row of 16 lanes, then the row-15 and lane-31 broadcasts carry row sums across rows, each move taken
// only when its source lane lies in lane i's segment [ss, i].  Only the last lane's result is the segment sum; a
// lane with ss > lane (ss = 64) belongs to no segment.  A step no lane takes is skipped.  Every lane's values must be
// finite: a move not taken is a multiply by 0.
template <int CTRL, int ROWS>
__device__ inline double dpp_d(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)b, CTRL, ROWS, 0xf, true);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, ROWS, 0xf, true);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

template <int CTRL, int ROWS, int N>
__device__ inline void seg_step(double (&v)[N], bool take) {
    if (__ballot(take) == 0) return;
    const double f = take ? 1.0 : 0.0;  // exact: v + 1 * o rounds once, v + 0 * o is v (o finite)
    for (int i = 0; i < N; i++) v[i] = fma(dpp_d<CTRL, ROWS>(v[i]), f, v[i]);
}

template <int N>
__device__ inline void seg_sum(double (&v)[N], int lane, int ss) {
    const int r = lane & 15, row = lane >> 4;
    seg_step<0x111, 0xf>(v, r >= 1 && lane - 1 >= ss);  // row_shr:1
    seg_step<0x112, 0xf>(v, r >= 2 && lane - 2 >= ss);  // row_shr:2
    seg_step<0x114, 0xf>(v, r >= 4 && lane - 4 >= ss);  // row_shr:4
    seg_step<0x118, 0xf>(v, r >= 8 && lane - 8 >= ss);  // row_shr:8
    seg_step<0x142, 0xa>(v, (row & 1) && lane - r - 1 >= ss);  // row_bcast:15 into rows 1, 3
    seg_step<0x143, 0xc>(v, row >= 2 && 31 >= ss);              // row_bcast:31 into rows 2, 3
}

template <int CTRL, int ROWS>
__device__ inline unsigned seg_or_step(unsigned v, bool take) {
    const unsigned o = (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, ROWS, 0xf, true);
    return take ? (v | o) : v;
}

__device__ inline unsigned seg_or(unsigned v, int lane, int ss) {
    const int r = lane & 15, row = lane >> 4;
    v = seg_or_step<0x111, 0xf>(v, r >= 1 && lane - 1 >= ss);
    v = seg_or_step<0x112, 0xf>(v, r >= 2 && lane - 2 >= ss);
    v = seg_or_step<0x114, 0xf>(v, r >= 4 && lane - 4 >= ss);
    v = seg_or_step<0x118, 0xf>(v, r >= 8 && lane - 8 >= ss);
    v = seg_or_step<0x142, 0xa>(v, (row & 1) && lane - r - 1 >= ss);
    v = seg_or_step<0x143, 0xc>(v, row >= 2 && 31 >= ss);
    return v;
}

// The work item's robust cost from its passes' costs (written by the workgroup's wavefronts before the last
// barrier), summed by one wavefront in a fixed order: lane i takes the passes p0 + i, p0 + i + 64, ... (its loads in
// flight together across the lanes), then the lanes are summed.
__device__ inline void chunk_cost(const double *pass_cost, const Chunk &ch, int lane, double *out) {
    double cs = 0.0;
    for (int q = ch.p0 + lane; q < ch.p1; q += XBA_WAVE) cs += pass_cost[q];
    cs = wave_sum(cs, lane);
    if (lane == 0) *out = cs;
}

// Work items are handed out dynamically, largest first (the host orders `sched`): thread 0 takes the next index of
// sched[cbeg, cend) from the launch kind's counter, and the last workgroup to find the list exhausted resets both
// counters for the next launch of that kind.  ctl[1] is the chunk's next pass, taken by the wavefronts in turn.
// Returns the work item, or -1 when there is none left.
__device__ inline int next_chunk(const KArgs<double> &A, int kind, int cbeg, int cend, int *ctl) {
    __syncthreads();
    if (threadIdx.x == 0) {
        int32_t *ctr = A.sched_ctr + 2 * kind;
        const int i = __hip_atomic_fetch_add(ctr, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        int c = -1;
        if (cbeg + i < cend) {
            c = A.sched[cbeg + i];
        } else if (__hip_atomic_fetch_add(ctr + 1, 1, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT) ==
                   (int)gridDim.x - 1) {
            __hip_atomic_store(ctr, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(ctr + 1, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        ctl[0] = c;
        ctl[1] = c >= 0 ? A.chunks[c].p0 : 0;
    }
    __syncthreads();
    return ctl[0];
}

// The mixed solve's FP64 re-assembly (A.only_fallback): this workgroup's share of the launch's work items,
// sched[i] with i = cbeg + blockIdx.x + k gridDim.x, k = 0, 1, ..., keeping those whose window was flagged
// WS_FALLBACK; one per call, -1 when none is left.  The threads read a round of blockDim.x items' flags together,
// so a launch with nothing flagged costs one round per workgroup (no hand-out counter).  ctl[2] holds the next k
// (0 before the first call), ctl[3] the round's first flagged thread.
__device__ inline int next_fallback_chunk(const KArgs<double> &A, int cbeg, int cend, int *ctl) {
    for (;;) {
        __syncthreads();
        const int k0 = ctl[2];
        if (cbeg + (int)blockIdx.x + k0 * (int)gridDim.x >= cend) return -1;  // uniform
        if (threadIdx.x == 0) ctl[3] = INT_MAX;
        __syncthreads();
        const int i = cbeg + (int)blockIdx.x + (k0 + (int)threadIdx.x) * (int)gridDim.x;
        if (i < cend) {
            const int c = A.sched[i];
            if (A.wstate[(size_t)NWS * A.chunks[c].win + WS_FALLBACK] != 0.0) atomicMin(ctl + 3, (int)threadIdx.x);
        }
        __syncthreads();
        const int t = ctl[3];
        __syncthreads();
        if (threadIdx.x == 0) ctl[2] = t == INT_MAX ? k0 + (int)blockDim.x : k0 + t + 1;
        if (t == INT_MAX) continue;
        const int c = A.sched[cbeg + (int)blockIdx.x + (k0 + t) * (int)gridDim.x];
        if (threadIdx.x == 0) {
            ctl[0] = c;
            ctl[1] = A.chunks[c].p0;
        }
        __syncthreads();
        return c;
    }
}

// The next pass of the current work item for this wavefront (uniform across the wavefront).
__device__ inline int next_pass(int *ctl, int lane) {
    int p = 0;
    if (lane == 0) p = __hip_atomic_fetch_add(ctl + 1, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    return __builtin_amdgcn_readlane(p, 0);
}

// ------------------------------------------------------------------------------------------ k_lin_fast --
// One observation row held in its lane's registers in the record type RT: Huber weight (times the row weight),
// whitened residual, whitened Jacobians wrt the landmark, the target pose block and the host block's rotation
// columns (its translation columns are the negated target ones, see obs_eval), and the target slot in the tile.
template <typename RT>
struct ObsRow {
    RT w, e[2], jl[6], jt[12], jhr[6];
    int t;
    __device__ inline double jh(int r, int j) const { return j < 3 ? (double)jhr[3 * r + j] : -(double)jt[6 * r + j]; }
    __device__ inline void zero() {
        w = (RT)0;
        for (int i = 0; i < 2; i++) e[i] = (RT)0;
        for (int i = 0; i < 6; i++) jl[i] = jhr[i] = (RT)0;
        for (int i = 0; i < 12; i++) jt[i] = (RT)0;
    }
    // false when any entry is Inf or NaN (their sum is then not finite)
    __device__ inline bool finite() const {
        RT s = w + e[0] + e[1];
        for (int i = 0; i < 6; i++) s += jl[i] + jhr[i];
        for (int i = 0; i < 12; i++) s += jt[i];
        return isfinite(s);
    }
};

// Residual, Huber weight and cost (FP64) and whitened Jacobians (FP64, stored in RT) of observation row k of a
// landmark hosted at tile slot h (vio/window.py:125-146, 164-263, 284-288).  Returns the robust cost (0 inactive).
template <typename RT>
__device__ inline double obs_row(const KArgs<double> &A, int k, int h, double a, double b, double rho,
                                 const double *sc64, const double *sc, const double *wc, const double *wd,
                                 ObsRow<RT> &o) {
    const int info = A.ob_info[k];
    const int t = obs_tile(info), cam = (info >> 12) & 1, valid = (info >> 13) & 1;
    const int nd = cam == 0 ? 2 : 1;
    o.t = t;
    double ev[2] = {0.0, 0.0}, Jl[2][3], Jt[2][6], Jh[2][6];
    bool act = false;
    if (valid)
        act = obs_eval<double>(sc64 + h * SC64_STRIDE, sc64 + t * SC64_STRIDE, sc + h * SC_STRIDE,
                               sc + t * SC_STRIDE, wc, wd, a, b, rho, cam, A.z[2 * (size_t)k], A.z[2 * (size_t)k + 1],
                               t == h, true, ev, Jl, Jt, Jh);
    double e[2] = {ev[0], ev[1]};  // a fresh copy, only ever indexed by constants
    double w = 0.0, c = 0.0, en2 = 0.0;
    if (act) {
        const double isg = A.isig[k];
        // fixed trip counts (rows past nd hold zeros): the arrays stay in registers
        for (int i = 0; i < 2; i++) {
            e[i] *= isg;
            if (i < nd) en2 += e[i] * e[i];
        }
        huber(sqrt(en2), wd[5], &w, &c);
        if (A.ob_w) {
            w *= A.ob_w[k];
            c *= A.ob_w[k];
        }
        for (int r = 0; r < 2; r++) {
            for (int j = 0; j < 3; j++) Jl[r][j] *= isg;
            for (int j = 0; j < 6; j++) {
                Jt[r][j] *= isg;
                Jh[r][j] *= isg;
            }
        }
    }
    if (A.chi2_out) A.chi2_out[k] = act ? en2 : (double)INFINITY;
    o.w = act ? (RT)w : (RT)0;
    for (int r = 0; r < 2; r++) {
        const bool on = act && r < nd;
        o.e[r] = on ? (RT)e[r] : (RT)0;
        for (int j = 0; j < 3; j++) {
            o.jl[3 * r + j] = on ? (RT)Jl[r][j] : (RT)0;
            o.jhr[3 * r + j] = on ? (RT)Jh[r][j] : (RT)0;
        }
        for (int j = 0; j < 6; j++) o.jt[6 * r + j] = on ? (RT)Jt[r][j] : (RT)0;
    }
    return act ? c : 0.0;
}

// RT is the type the per-observation record (whitened residual, weight and Jacobians) is stored in; every sum and
// product is formed in FP64.  CLS names the work-item class the launch covers (0: narrow chunks, 1: wide chunks);
// the code is the same, the two classes get separate symbols so that profiles report them apart.
template <typename RT, int CLS>
__global__ __launch_bounds__(512) void k_lin_fast(KArgs<double> A, int Rmax, int cbeg, int cend, int kind) {
    using T = double;
    extern __shared__ __align__(16) unsigned char smem[];
    const int NW = blockDim.x / XBA_WAVE;
    const FastLayout FL = fast_layout(Rmax, (int)sizeof(RT), NW, false);
    T *sc = (T *)(smem + FL.sc), *wc = (T *)(smem + FL.wc);
    double *sc64 = (double *)(smem + FL.sc64), *wd = (double *)(smem + FL.wd);
    double *tile = (double *)(smem + FL.tile), *btile = (double *)(smem + FL.btile);
    const int tid = threadIdx.x, lane = tid % XBA_WAVE, wv = tid / XBA_WAVE;
    unsigned char *wb = smem + FL.wave + wv * FL.wave_stride;
    int *ob_lm = (int *)(wb + FL.o_oblm);
    double *rcq = (double *)(wb + FL.o_rcq);
    unsigned *lmask = (unsigned *)(wb + FL.o_lmask);
    int *lobs0 = (int *)(wb + FL.o_lobs0), *lnobs = (int *)(wb + FL.o_lnobs), *loff = (int *)(wb + FL.o_loff);
    int *poff = (int *)(wb + FL.o_poff);
    int *slots = (int *)(wb + FL.o_slots);
    T *hlp = (T *)(wb + FL.o_hlp), *lrec = (T *)(wb + FL.o_lrec);
    double *lacc = (double *)(wb + FL.o_lrow);
    uint16_t *ptab = (uint16_t *)(wb + FL.o_ptab);
    int *ctl = (int *)(smem + FL.ctl);

    if (threadIdx.x == 0) ctl[2] = 0;
    for (;;) {
        // the mixed solve's FP64 re-assembly visits only the windows it flagged
        const int c = A.only_fallback ? next_fallback_chunk(A, cbeg, cend, ctl) : next_chunk(A, kind, cbeg, cend, ctl);
        if (c < 0) break;
        const Chunk ch = A.chunks[c];
        const int32_t *smap = ch.smap >= 0 ? A.chunk_slots + ch.smap : nullptr;
        const WinInfo W = A.wins[ch.win];
        const double *wp = A.win_par + (size_t)XBA_WPAR * ch.win;
        const int R = ch.R, nblk = R * (R + 1) / 2;
        const double lamd = A.lam_override >= 0.0 ? A.lam_override : A.wstate[(size_t)NWS * ch.win + WS_LAM];
        if (wv == 0) {
            load_slots<T>(sc, sc64, A.Rwb, A.pwb, wp, W.slot0, ch.lo, R, lane, smap);
            load_wconst<T>(wc, wp, lane);
            load_wd(wd, wp, lane);
        }
        for (int i = tid; i < nblk * TBLK; i += blockDim.x) tile[i] = 0.0;
        for (int i = tid; i < 6 * R; i += blockDim.x) btile[i] = 0.0;
        __syncthreads();
        // the chunk's passes, taken by the wavefronts in turn (each one pass ahead)
        PassAhead pa;
        pass_take(A, ch, ctl, lane, pa);
        pass_rows(A.ob_start, lane, pa);
        for (;;) {
            const int p = pa.p;
            if (p >= ch.p1) break;
            const int lb = pa.lb, B = pa.B;
            int obase;
            const int nob = load_pass(pa, lane, lobs0, lnobs, ob_lm, &obase);
            pass_take(A, ch, ctl, lane, pa);  // the next pass; its row offsets are loaded after the next wave sync
            const int pl_l = lb + lane;  // this lane's landmark (lane < B)
            // the landmark lanes' own loads, issued now so that they arrive during the observation phase, and their
            // accumulators cleared
            int pl_li = 0;
            double pl_rho = 0.0, pl_rm = 0.0, pl_ri = 0.0;
            if (lane < B) {
                pl_li = A.lm_info[pl_l];
                pl_rho = A.lmp[3 * (size_t)pl_l + 2];
                pl_rm = A.rho_meas[pl_l];
                pl_ri = (double)A.rho_isig[pl_l];
                for (int i = 0; i < LA_N; i++) lacc[LA_N * lane + i] = 0.0;
                lmask[lane] = 0u;
            }
            WAVE_SYNC();
            pass_rows(A.ob_start, lane, pa);
            // ---- observations (one lane each, kept in registers): whitened residual, Huber weight and cost,
            //      whitened Jacobians.  The per-landmark sums (Hll, bl, active rows, touched slots) and the host
            //      blocks (h,h) and host gradient (one host slot per landmark) are summed over the landmark's lanes
            //      and written once per landmark by its last lane.  The row's own target blocks (t,t), (t,h) and
            //      target gradient are added later, together with their Schur parts.
            ObsRow<RT> ob;
            ob.zero();
            ob.t = 0;
            int ob_bb = 0, ob_h = 0, seg_ss = XBA_WAVE, seg_last = 0;
            if (lane < nob) {
                ob_bb = ob_lm[lane];
                seg_ss = lobs0[ob_bb];
                seg_last = lane == seg_ss + lnobs[ob_bb] - 1;
                const int l = lb + ob_bb, k = obase + lane;
                const int li = A.lm_info[l];
                ob_h = lm_tile(li);
                double cq = 0.0;
                if ((li >> 12) & 1)
                    cq = obs_row<RT>(A, k, ob_h, A.lmp[3 * (size_t)l], A.lmp[3 * (size_t)l + 1],
                                     A.lmp[3 * (size_t)l + 2], sc64, sc, wc, wd, ob);
                rcq[lane] = cq;
            }
            // the first active cross row of each (landmark, target slot) adds that slot's Schur parts
            bool ob_first = false;
            // the host block (h,h) and host gradient of the landmark's rows, summed over its lanes (complete in its
            // last lane, which adds them to the tile together with their Schur parts)
            double hg[6], hh[21];
            {
                // The lane sums below need finite values in every lane (lanes with nothing to add hold zeros).  A
                // row with non-finite entries (from non-finite inputs) is left out of them and marks its landmark's
                // Hll non-finite instead, so that the landmark fails the determinant guard as it would with the
                // row summed in; its own target terms still carry the non-finite values into the tile.
                const bool bad = ob.w != (RT)0 && !ob.finite();
                const double w = bad ? 0.0 : (double)ob.w;
                const bool cross = w != 0.0 && ob.t != ob_h;
                const bool head = seg_last;
                double s[LA_N + 1];
                {
                    const double l00 = bad ? 0.0 : (double)ob.jl[0], l01 = bad ? 0.0 : (double)ob.jl[1];
                    const double l02 = bad ? 0.0 : (double)ob.jl[2], l10 = bad ? 0.0 : (double)ob.jl[3];
                    const double l11 = bad ? 0.0 : (double)ob.jl[4], l12 = bad ? 0.0 : (double)ob.jl[5];
                    const double e0 = bad ? 0.0 : (double)ob.e[0], e1 = bad ? 0.0 : (double)ob.e[1];
                    s[LA_H + 0] = w * (l00 * l00 + l10 * l10);
                    s[LA_H + 1] = w * (l00 * l01 + l10 * l11);
                    s[LA_H + 2] = w * (l00 * l02 + l10 * l12);
                    s[LA_H + 3] = w * (l01 * l01 + l11 * l11);
                    s[LA_H + 4] = w * (l01 * l02 + l11 * l12);
                    s[LA_H + 5] = w * (l02 * l02 + l12 * l12);
                    s[LA_B + 0] = w * (l00 * e0 + l10 * e1);
                    s[LA_B + 1] = w * (l01 * e0 + l11 * e1);
                    s[LA_B + 2] = w * (l02 * e0 + l12 * e1);
                    s[LA_NACT] = w != 0.0 ? 1.0 : 0.0;
                    s[LA_NX] = cross ? 1.0 : 0.0;
                    s[LA_N] = bad ? 1.0 : 0.0;
                }
                seg_sum(s, lane, seg_ss);
                if (s[LA_N] != 0.0) s[LA_H] = (double)NAN;
                // an inclusive scan over the landmark's lanes: its last lane holds the touched-slot mask, and a
                // cross row is the first of its target slot when the scan up to the lane before it lacks that slot
                const unsigned m = seg_or(cross ? (1u << ob.t) | (1u << ob_h) : 0u, lane, seg_ss);
                const unsigned mprev = (unsigned)__builtin_amdgcn_ds_bpermute(((lane + XBA_WAVE - 1) % XBA_WAVE) << 2,
                                                                              (int)m);
                ob_first = cross && !(lane > seg_ss && ((mprev >> ob.t) & 1u));
                if (head) {
                    double *la = lacc + LA_N * ob_bb;
                    for (int i = 0; i < LA_N; i++) la[i] = s[i];
                    lmask[ob_bb] = m;
                }
                double xh[2][6];
                for (int r = 0; r < 2; r++)
                    for (int j = 0; j < 6; j++) xh[r][j] = cross ? ob.jh(r, j) : 0.0;
                const double e0 = cross ? w * (double)ob.e[0] : 0.0, e1 = cross ? w * (double)ob.e[1] : 0.0;
                for (int j = 0; j < 6; j++) hg[j] = xh[0][j] * e0 + xh[1][j] * e1;
                seg_sum(hg, lane, seg_ss);
                for (int i = 0, o = 0; i < 6; i++)
                    for (int j = 0; j <= i; j++, o++) hh[o] = w * (xh[0][i] * xh[0][j] + xh[1][i] * xh[1][j]);
                seg_sum(hh, lane, seg_ss);
            }
            WAVE_SYNC();
            // ---- landmarks (one lane each): touched slots, Hll, bl, prior, damping, det guard, FP64 inverse,
            //      Hll = L L^T
            double lcost = 0.0;
            int kb = 0;
            bool lok = false;
            unsigned msk = 0u;
            if (lane < B) {
                const int l = pl_l, li = pl_li;
                T *lr = lrec + lane * FB_LREC;
                lr[LR_OK] = 0.0;
                lr[LR_UNIQ] = 0.0;
                if ((li >> 12) & 1) {
                    const int q0 = lobs0[lane], n = lnobs[lane];
                    for (int q = q0; q < q0 + n; q++) lcost += rcq[q];
                    const double *la = lacc + LA_N * lane;
                    const int nact = (int)la[LA_NACT];
                    msk = lmask[lane];
                    kb = __popc(msk);
                    double H6[6], bl[3];
                    for (int i = 0; i < 6; i++) H6[i] = la[LA_H + i];
                    for (int i = 0; i < 3; i++) bl[i] = la[LA_B + i];
                    const double rho_l = pl_rho;
                    const T ris = pl_ri;
                    if (ris > (T)0) {
                        double risd = (double)ris, r0 = (rho_l - pl_rm) * risd, w, cc;
                        huber(fabs(r0), wd[5], &w, &cc);
                        lcost += cc;
                        H6[5] += w * risd * risd;
                        bl[2] += w * risd * r0;
                    }
                    double Hll[9] = {H6[0], H6[1], H6[2], H6[1], H6[3], H6[4], H6[2], H6[4], H6[5]};
                    for (int i = 0; i < 3; i++) Hll[4 * i] += lamd * Hll[4 * i] + 1e-12;
                    double det = (Hll[0] * (Hll[4] * Hll[8] - Hll[5] * Hll[7]) -
                                  Hll[1] * (Hll[3] * Hll[8] - Hll[5] * Hll[6]) +
                                  Hll[2] * (Hll[3] * Hll[7] - Hll[4] * Hll[6]));
                    double tr = Hll[0] + Hll[4] + Hll[8];
                    bool ok = nact > 0 && (det > 1e-18 * tr * tr * tr) && isfinite(det);
                    double Hi[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
                    if (ok) {
                        double id = 1.0 / det;
                        Hi[0] = (Hll[4] * Hll[8] - Hll[5] * Hll[7]) * id;
                        Hi[1] = (Hll[2] * Hll[7] - Hll[1] * Hll[8]) * id;
                        Hi[2] = (Hll[1] * Hll[5] - Hll[2] * Hll[4]) * id;
                        Hi[3] = (Hll[5] * Hll[6] - Hll[3] * Hll[8]) * id;
                        Hi[4] = (Hll[0] * Hll[8] - Hll[2] * Hll[6]) * id;
                        Hi[5] = (Hll[2] * Hll[3] - Hll[0] * Hll[5]) * id;
                        Hi[6] = (Hll[3] * Hll[7] - Hll[4] * Hll[6]) * id;
                        Hi[7] = (Hll[1] * Hll[6] - Hll[0] * Hll[7]) * id;
                        Hi[8] = (Hll[0] * Hll[4] - Hll[1] * Hll[3]) * id;
                        // Hll = L L^T: the Schur blocks become (L^-1 A2)^T (L^-1 A1)
                        const double c00 = sqrt(fmax(Hll[0], 1e-300));
                        const double c10 = Hll[3] / c00, c20 = Hll[6] / c00;
                        const double c11 = sqrt(fmax(Hll[4] - c10 * c10, 1e-300));
                        const double c21 = (Hll[7] - c20 * c10) / c11;
                        const double c22 = sqrt(fmax(Hll[8] - c20 * c20 - c21 * c21, 1e-300));
                        const double i00 = 1.0 / c00, i11 = 1.0 / c11, i22 = 1.0 / c22;
                        const double i10 = -c10 * i00 * i11;
                        const double i21 = -c21 * i11 * i22;
                        const double i20 = -(c20 * i00 + c21 * i10) * i22;
                        lr[LR_L + 0] = i00; lr[LR_L + 1] = i10; lr[LR_L + 2] = i11;
                        lr[LR_L + 3] = i20; lr[LR_L + 4] = i21; lr[LR_L + 5] = i22;
                        lr[LR_G + 0] = i00 * bl[0];
                        lr[LR_G + 1] = i10 * bl[0] + i11 * bl[1];
                        lr[LR_G + 2] = i20 * bl[0] + i21 * bl[1] + i22 * bl[2];
                    }
                    lok = ok;
                    lr[LR_OK] = ok ? 1.0 : 0.0;
                    lr[LR_UNIQ] = (int)la[LA_NX] == kb - 1 ? 1.0 : 0.0;
                    A.lmok[l] = ok ? 1 : 0;
                    if (A.ok_out) A.ok_out[l] = ok ? 1 : 0;
                    T *ho = A.lmHi + 6 * (size_t)l;
                    ho[0] = Hi[0]; ho[1] = Hi[1]; ho[2] = Hi[2]; ho[3] = Hi[4]; ho[4] = Hi[5];
                    ho[5] = Hi[8];
                    T *bo = A.lmbl + 3 * (size_t)l;
                    bo[0] = bl[0]; bo[1] = bl[1]; bo[2] = bl[2];
                    if (A.Hinv_out)
                        for (int i = 0; i < 9; i++) A.Hinv_out[9 * (size_t)l + i] = ok ? Hi[i] : 0.0;
                } else {
                    A.lmok[l] = 0;
                }
            }
            const int koff = wave_incl_scan(kb, lane) - kb;
            const int ktot = __builtin_amdgcn_readlane(koff + kb, B - 1);
            // the pair loop's share of the landmark's Schur blocks: the blocks (s2, s1) of two different target slots
            // s2 > s1; a landmark that failed the det guard has no Schur part, one without cross rows touches no slot
            const int npb = (lok && kb > 0) ? (kb - 1) * (kb - 2) / 2 : 0;
            const int pof = wave_incl_scan(npb, lane) - npb;
            const int ptot = __builtin_amdgcn_readlane(pof + npb, B - 1);
            if (lane < B) {
                poff[lane] = pof;
                loff[lane] = koff;
                unsigned m = msk;
                int i = 0;
                while (m) {
                    slots[koff + i] = __ffs(m) - 1;
                    m &= m - 1u;
                    i++;
                }
                if (npb > 0) {
                    const int he = __popc(msk & ((1u << lm_tile(pl_li)) - 1u));  // the host's touched-slot entry
                    int o = pof;
                    for (int si = 0; si < kb; si++)
                        for (int sj = 0; sj < si; sj++)
                            if (si != he && sj != he) ptab[o++] = (uint16_t)((koff << 10) | (si << 5) | sj);
                }
            }
            const double pcost = wave_sum(lcost, lane);
            if (lane == 0) A.pass_cost[p] = pcost;
            WAVE_SYNC();
            // ---- U = L^-1 Hlp, one lane per observation: its two 3x6 terms G Jt and G Jh (G = w L^-1 Jl^T); the
            //      target term is added into U of its target slot, the host terms are summed over the landmark's
            //      lanes and stored once into U of the host slot by the landmark's last lane (the slot's entry is
            //      its rank in the landmark's touched-slot mask)
            // the entries are summed with atomics from cleared storage only for landmarks whose rows share a target
            // slot; otherwise every target entry is stored once by its row and the host entry by the last lane
            if (__ballot(lane < B && lok && kb > 0 && lrec[lane * FB_LREC + LR_UNIQ] == 0.0) != 0)
                for (int e = lane; e < ktot * 18; e += XBA_WAVE) hlp[e] = 0.0;
            WAVE_SYNC();
            double ut[18];  // the row's term G Jt of its target slot's U entry (the whole entry when LR_UNIQ)
            for (int i = 0; i < 18; i++) ut[i] = 0.0;
            {
                double uh[18];
                for (int i = 0; i < 18; i++) uh[i] = 0.0;
                const int bb = ob_bb, t = ob.t, h = ob_h;
                const T *lr = lrec + bb * FB_LREC;
                const unsigned m = lane < nob ? lmask[bb] : 0u;
                if (lane < nob) {
                    const T w = (T)ob.w;
                    if (lr[LR_OK] != 0.0 && w != 0.0 && t != h) {
                        T *Ut = hlp + (loff[bb] + __popc(m & ((1u << t) - 1u))) * 18;
                        T G[3][2];
                        for (int r = 0; r < 2; r++) {
                            const T j0 = ob.jl[3 * r], j1 = ob.jl[3 * r + 1], j2 = ob.jl[3 * r + 2];
                            G[0][r] = w * (lr[LR_L] * j0);
                            G[1][r] = w * (lr[LR_L + 1] * j0 + lr[LR_L + 2] * j1);
                            G[2][r] = w * (lr[LR_L + 3] * j0 + lr[LR_L + 4] * j1 + lr[LR_L + 5] * j2);
                        }
                        for (int j = 0; j < 6; j++) {
                            const T t0 = ob.jt[j], t1 = ob.jt[6 + j];
                            const T h0 = ob.jh(0, j), h1 = ob.jh(1, j);
                            for (int c = 0; c < 3; c++) {
                                ut[6 * c + j] = G[c][0] * t0 + G[c][1] * t1;
                                uh[6 * c + j] = G[c][0] * h0 + G[c][1] * h1;
                            }
                        }
                        if (lr[LR_UNIQ] != 0.0)
                            for (int i = 0; i < 18; i++) Ut[i] = ut[i];
                        else
                            for (int i = 0; i < 18; i++) lds_add<double>(Ut + i, ut[i]);
                    }
                }
                seg_sum(uh, lane, seg_ss);
                if (seg_last && m != 0u) {
                    // the landmark's host block and host gradient: the rows' sums, and for a landmark that passed the
                    // det guard their Schur parts -U_h^T U_h and -U_h^T g (U_h is zero otherwise)
                    const bool ok = lr[LR_OK] != 0.0;
                    if (ok) {
                        T *Uh = hlp + (loff[bb] + __popc(m & ((1u << h) - 1u))) * 18;
                        for (int i = 0; i < 18; i++) Uh[i] = uh[i];
                    }
                    const double g0 = ok ? lr[LR_G] : 0.0, g1 = ok ? lr[LR_G + 1] : 0.0, g2 = ok ? lr[LR_G + 2] : 0.0;
                    for (int j = 0; j < 6; j++)
                        lds_add<double>(btile + 6 * h + j, hg[j] - (uh[j] * g0 + uh[6 + j] * g1 + uh[12 + j] * g2));
                    double *dhh = tile + (h * (h + 1) / 2 + h) * TBLK;
                    for (int i = 0, o = 0; i < 6; i++)
                        for (int j = 0; j <= i; j++, o++)
                            lds_add<double>(dhh + 6 * i + j,
                                            hh[o] - (uh[i] * uh[j] + uh[6 + i] * uh[6 + j] + uh[12 + i] * uh[12 + j]));
                }
            }
            WAVE_SYNC();
            // ---- the rows' own blocks: every active cross row adds w Jt^T Jt into (t,t), its (t,h) block and
            //      w Jt^T e into the gradient of t; the first row of each (landmark, target slot) of a landmark that
            //      passed the det guard adds the slot's Schur parts with them: -U_t^T U_t, -U_t^T U_h (stored under
            //      the larger slot, so -U_h^T U_t when t < h) and -U_t^T g
            if (lane < nob && ob.w != (RT)0 && ob.t != ob_h) {
                const int bb = ob_bb, t = ob.t, h = ob_h;
                const T *lr = lrec + bb * FB_LREC;
                double uu[18], gl[3] = {0.0, 0.0, 0.0};
                for (int i = 0; i < 18; i++) uu[i] = 0.0;
                if (ob_first && lr[LR_OK] != 0.0) {
                    const unsigned m = lmask[bb];
                    const T *Uh = hlp + (loff[bb] + __popc(m & ((1u << h) - 1u))) * 18;
                    for (int i = 0; i < 18; i++) uu[i] = Uh[i];
                    if (lr[LR_UNIQ] == 0.0) {  // the entry summed over the rows sharing the slot
                        const T *Ut = hlp + (loff[bb] + __popc(m & ((1u << t) - 1u))) * 18;
                        for (int i = 0; i < 18; i++) ut[i] = Ut[i];
                    }
                    for (int c = 0; c < 3; c++) gl[c] = lr[LR_G + c];
                } else {
                    for (int i = 0; i < 18; i++) ut[i] = 0.0;
                }
                const double w = (double)ob.w;
                double xt[2][6], xh[2][6];
                for (int r = 0; r < 2; r++)
                    for (int j = 0; j < 6; j++) {
                        xt[r][j] = (double)ob.jt[6 * r + j];
                        xh[r][j] = ob.jh(r, j);
                    }
                const double e0 = w * (double)ob.e[0], e1 = w * (double)ob.e[1];
                for (int j = 0; j < 6; j++)
                    lds_add<double>(btile + 6 * t + j, (xt[0][j] * e0 + xt[1][j] * e1) -
                                                           (ut[j] * gl[0] + ut[6 + j] * gl[1] + ut[12 + j] * gl[2]));
                double *dtt = tile + (t * (t + 1) / 2 + t) * TBLK;
                for (int i = 0; i < 6; i++)
                    for (int j = 0; j <= i; j++)
                        lds_add<double>(dtt + 6 * i + j, w * (xt[0][i] * xt[0][j] + xt[1][i] * xt[1][j]) -
                                                             (ut[i] * ut[j] + ut[6 + i] * ut[6 + j] +
                                                              ut[12 + i] * ut[12 + j]));
                const bool tlo = t < h;  // the (t,h) block is stored under the larger slot
                double *dth = tile + (tlo ? h * (h + 1) / 2 + t : t * (t + 1) / 2 + h) * TBLK;
                for (int i = 0; i < 6; i++)
                    for (int j = 0; j < 6; j++) {
                        const double v = tlo ? w * (xh[0][i] * xt[0][j] + xh[1][i] * xt[1][j]) -
                                                   (uu[i] * ut[j] + uu[6 + i] * ut[6 + j] + uu[12 + i] * ut[12 + j])
                                             : w * (xt[0][i] * xh[0][j] + xt[1][i] * xh[1][j]) -
                                                   (ut[i] * uu[j] + ut[6 + i] * uu[6 + j] + ut[12 + i] * uu[12 + j]);
                        lds_add<double>(dth + 6 * i + j, v);
                    }
            }
            // ---- reduced system, the remaining blocks: one lane per (landmark, pair of target slots s2 > s1) of the
            //      landmarks that passed the det guard (the table entry holds the landmark's first touched-slot entry
            //      and the pair); both U entries held in registers, -U2^T U1 formed one row at a time, each row's
            //      atomics issued after its products
            for (int o = lane; o < ptot; o += XBA_WAVE) {
                const int pk = ptab[o];
                const int base = pk >> 10, si = (pk >> 5) & 31, sj = pk & 31;
                const int s2 = slots[base + si], s1 = slots[base + sj];
                const T *U2 = hlp + (base + si) * 18, *U1 = hlp + (base + sj) * 18;
                T u2[18], a1[18];
                for (int e = 0; e < 18; e++) {
                    u2[e] = U2[e];
                    a1[e] = U1[e];
                }
                double *dst = tile + (s2 * (s2 + 1) / 2 + s1) * TBLK;
                for (int i = 0; i < 6; i++) {
                    const T c0 = u2[i], c1 = u2[6 + i], c2 = u2[12 + i];
                    for (int jj = 0; jj < 6; jj++)
                        lds_add<double>(dst + 6 * i + jj, -(c0 * a1[jj] + c1 * a1[6 + jj] + c2 * a1[12 + jj]));
                }
            }
            WAVE_SYNC();
        }
        __syncthreads();
        double *pH = A.partH64 + ch.partH, *pb = A.partb64 + ch.partb;
        for (int i = tid; i < nblk * 36; i += blockDim.x) pH[i] = tile[(i / 36) * TBLK + i % 36];
        for (int i = tid; i < 6 * R; i += blockDim.x) pb[i] = btile[i];
        if (wv == 0) chunk_cost(A.pass_cost, ch, lane, A.partc + c);
    }
}

// ---------------------------------------------------------------------------------------- k_trial_fast --
// Latency-bound (LDS reads, wave-local syncs, a few FP64 divides per observation): compiled for 4 waves per SIMD
// (at most 128 VGPRs), so that two 8-wave workgroups share a CU.
#define XBA_TRIAL_WAVES_PER_SIMD 4
#define XBA_TRIAL_PER_CU 2
template <int CLS>
__global__ __launch_bounds__(512, XBA_TRIAL_WAVES_PER_SIMD) void k_trial_fast(KArgs<double> A, int Rmax, int cbeg, int cend, int kind) {
    using T = double;
    extern __shared__ __align__(16) unsigned char smem[];
    const int NW = blockDim.x / XBA_WAVE;
    const FastLayout FL = fast_layout(Rmax, 8, NW, true);
    T *sc = (T *)(smem + FL.sc), *wc = (T *)(smem + FL.wc), *dps = (T *)(smem + FL.tile);
    // In FP64 the slot records sc hold the camera poses (their first 12 entries) as the linearisation's FP64 copies
    // do, and wc + 18 holds fx fy cx cy B huber_k: the observations read those, one copy of each value.
    const double *wdc = wc + 18;
    const int tid = threadIdx.x, lane = tid % XBA_WAVE, wv = tid / XBA_WAVE;
    unsigned char *wb = smem + FL.wave + wv * FL.wave_stride;
    int *ob_lm = (int *)(wb + FL.o_oblm);
    double *ldbl = (double *)(wb + FL.o_ldbl);
    int *lobs0 = (int *)(wb + FL.o_lobs0), *lnobs = (int *)(wb + FL.o_lnobs), *lmask = (int *)(wb + FL.o_lmask);
    int *ctl = (int *)(smem + FL.ctl);

    for (;;) {
        const int c = next_chunk(A, kind, cbeg, cend, ctl);
        if (c < 0) break;
        const Chunk ch = A.chunks[c];
        const int32_t *smap = ch.smap >= 0 ? A.chunk_slots + ch.smap : nullptr;
        const WinInfo W = A.wins[ch.win];
        if (A.wstate[(size_t)NWS * ch.win + WS_CHOLOK] == 0.0) {
            if (tid == 0) A.partc2[c] = 0.0;
            continue;
        }
        const double *wp = A.win_par + (size_t)XBA_WPAR * ch.win;
        const int R = ch.R;
        T *sc2 = sc + R * SC_STRIDE;
        if (wv == 0) {
            load_slots<T>(sc, nullptr, A.Rwb, A.pwb, wp, W.slot0, ch.lo, R, lane, smap);
            load_wconst<T>(wc, wp, lane);
        } else if (wv == 1 % NW) {
            load_slots<T>(sc2, nullptr, A.R2, A.p2, wp, W.slot0, ch.lo, R, lane, smap);
        }
        if (NW == 1) load_slots<T>(sc2, nullptr, A.R2, A.p2, wp, W.slot0, ch.lo, R, lane, smap);
        for (int i = tid; i < 6 * R; i += blockDim.x) {
            int s = i / 6, j = i % 6;
            dps[i] = (T)A.dx[W.st0 + A.slot_off[W.slot0 + chunk_slot(ch, A.chunk_slots, s)] + j];
        }
        __syncthreads();
        for (int s = tid; s < R; s += blockDim.x) dir_slot<T>(dps + 6 * s, sc + s * SC_STRIDE, wc);
        __syncthreads();
        for (;;) {
            PassAhead pa;
            pass_take(A, ch, ctl, lane, pa);
            const int p = pa.p;
            if (p >= ch.p1) break;
            const int lb = pa.lb, B = pa.B;
            // The landmark lanes' loads depend only on the plan: issued first, together with the row offsets, and
            // parked in LDS (landmark record, word) or registers (prior) for the two phases that read them.
            const bool lm_in = lane < B;
            int lw = 0;
            pass_rows(A.ob_start, lane, pa);
            if (lm_in) {
                const int l = lb + lane;
                double *d = ldbl + TL_N * lane;
                for (int i = 0; i < 3; i++) d[TL_CUR + i] = A.lmp[3 * (size_t)l + i];
                for (int i = 0; i < 6; i++) d[TL_HI + i] = (double)A.lmHi[6 * (size_t)l + i];
                for (int i = 0; i < 3; i++) d[TL_BL + i] = (double)A.lmbl[3 * (size_t)l + i];
                lw = A.lm_info[l] | (A.lmok[l] ? TL_LMOK : 0);
                lmask[lane] = lw;
                d[TL_RM] = A.rho_meas[l];
                d[TL_RI] = (double)A.rho_isig[l];
            }
            int obase;
            const int nob = load_pass(pa, lane, lobs0, lnobs, ob_lm, &obase);
            const bool q_in = lane < nob;
            const int q_k = obase + lane;
            const int info = q_in ? A.ob_info[q_k] : 0;
            WAVE_SYNC();
            // ---- observations at the current state: u_q = w Jl^T (Jt dp_t + Jh dp_h), summed per landmark over its
            //      lanes with DPP moves (complete in the landmark's last lane)
            int q_t = 0, q_h = 0, q_cam = 0, q_valid = 0, qb = 0, seg_ss = XBA_WAVE;
            bool q_mask = false;
            double uv[3] = {0.0, 0.0, 0.0};
            if (q_in) {
                qb = ob_lm[lane];
                seg_ss = lobs0[qb];
                const int li = lmask[qb];
                q_mask = (li >> 12) & 1;
                q_h = lm_tile(li);
                q_t = obs_tile(info);
                q_cam = (info >> 12) & 1;
                q_valid = (info >> 13) & 1;
                if (q_mask && q_valid && (li & TL_LMOK) && q_t != q_h) {
                    const double *d = ldbl + TL_N * qb;
                    const double a = d[TL_CUR], b = d[TL_CUR + 1], rho = d[TL_CUR + 2];
                    double e[2];
                    T Jl[2][3], jdv[2];
                    if (obs_dir(sc + q_h * SC_STRIDE, sc + q_t * SC_STRIDE, wdc, a, b, rho, q_cam, A.z[2 * (size_t)q_k],
                                A.z[2 * (size_t)q_k + 1], dps + 6 * q_t, dps + 6 * q_h, e, Jl, jdv)) {
                        const int nd = q_cam == 0 ? 2 : 1;
                        const double isg = (double)A.isig[q_k];
                        double en2 = 0.0, w, cst;
                        for (int i = 0; i < 2; i++)
                            if (i < nd) en2 += (e[i] * isg) * (e[i] * isg);
                        huber(sqrt(en2), wdc[5], &w, &cst);
                        if (A.ob_w) w *= A.ob_w[q_k];
                        const T isgT = A.isig[q_k];
                        T s[3] = {0, 0, 0};
                        for (int r = 0; r < 2; r++) {
                            const T jd = r < nd ? jdv[r] * (isgT * isgT) : (T)0;
                            for (int i = 0; i < 3; i++) s[i] += Jl[r][i] * jd;
                        }
                        for (int i = 0; i < 3; i++) uv[i] = (T)w * s[i];
                    }
                }
            }
            // The lane sums need finite values in every lane: a non-finite row is left out and fails its landmark's
            // update instead (as it would summed in: the step's rho would not be finite).
            const bool ubad = !(isfinite(uv[0]) && isfinite(uv[1]) && isfinite(uv[2]));
            if (ubad) uv[0] = uv[1] = uv[2] = 0.0;
            const uint64_t badm = __ballot(ubad);
            seg_sum(uv, lane, seg_ss);
            // ---- landmarks: back-substitution and the update rule; prior cost at the trial state
            const int l_n = lm_in ? lnobs[lane] : 0, l_s0 = lm_in ? lobs0[lane] : 0;
            const int src = (l_s0 + (l_n > 0 ? l_n - 1 : 0)) << 2;
            double v[3];
            for (int i = 0; i < 3; i++) {
                const long long bits = __double_as_longlong(uv[i]);
                const int lo = __builtin_amdgcn_ds_bpermute(src, (int)bits);
                const int hi = __builtin_amdgcn_ds_bpermute(src, (int)(bits >> 32));
                v[i] = l_n > 0 ? __longlong_as_double(((long long)hi << 32) | (unsigned)lo) : 0.0;
            }
            double lcost = 0.0;
            if (lm_in) {
                const int l = lb + lane;
                const double *d = ldbl + TL_N * lane;
                const double a = d[TL_CUR], b = d[TL_CUR + 1], rho = d[TL_CUR + 2];
                double na = a, nb = b, nr = rho;
                const uint64_t segm = l_n >= 64 ? ~0ull : ((1ull << l_n) - 1ull);
                if (((lw >> 12) & 1) && (lw & TL_LMOK) && ((badm >> l_s0) & segm) == 0ull) {
                    const double *Hi = d + TL_HI, *bl = d + TL_BL;
                    const double v0 = v[0] + bl[0], v1 = v[1] + bl[1], v2 = v[2] + bl[2];
                    const T d0 = -(Hi[0] * v0 + Hi[1] * v1 + Hi[2] * v2);
                    const T d1 = -(Hi[1] * v0 + Hi[3] * v1 + Hi[4] * v2);
                    const T d2 = -(Hi[2] * v0 + Hi[4] * v1 + Hi[5] * v2);
                    const double r = rho + (double)d2;
                    if (r > 1e-9 && isfinite(r)) {
                        na = a + (double)d0;
                        nb = b + (double)d1;
                        nr = r;
                    }
                }
                A.lmp_trial[3 * (size_t)l] = na;
                A.lmp_trial[3 * (size_t)l + 1] = nb;
                A.lmp_trial[3 * (size_t)l + 2] = nr;
                double *dn = ldbl + TL_N * lane + TL_NEW;
                dn[0] = na;
                dn[1] = nb;
                dn[2] = nr;
                if (((lw >> 12) & 1) && d[TL_RI] > 0.0) {
                    double r0 = (nr - d[TL_RM]) * d[TL_RI], w, cc;
                    huber(fabs(r0), wdc[5], &w, &cc);
                    lcost += cc;
                }
            }
            WAVE_SYNC();
            // ---- observations at the trial state: robust cost (+1e6 behind the camera)
            if (q_in) {
                double cq = 0.0;
                if (q_mask && q_valid) {
                    double e[2];
                    T Jl[2][3], Jt[2][6], Jh[2][6];
                    bool act = obs_eval<T>(sc2 + q_h * SC_STRIDE, sc2 + q_t * SC_STRIDE, sc2 + q_h * SC_STRIDE,
                                           sc2 + q_t * SC_STRIDE, wc, wdc, ldbl[TL_N * qb + TL_NEW],
                                           ldbl[TL_N * qb + TL_NEW + 1], ldbl[TL_N * qb + TL_NEW + 2], q_cam,
                                           A.z[2 * (size_t)q_k], A.z[2 * (size_t)q_k + 1], q_t == q_h, false, e, Jl,
                                           Jt, Jh);
                    if (!act) {
                        cq = 1e6;
                    } else {
                        double isg = (double)A.isig[q_k], en2 = e[0] * isg * (e[0] * isg), w;
                        if (q_cam == 0) en2 += e[1] * isg * (e[1] * isg);
                        huber(sqrt(en2), wdc[5], &w, &cq);
                        if (A.ob_w) cq *= A.ob_w[q_k];
                    }
                }
                lcost += cq;
            }
            const double pcost = wave_sum(lcost, lane);
            if (lane == 0) A.pass_cost[p] = pcost;
            WAVE_SYNC();
        }
        __syncthreads();
        if (tid == 0) {
            double cs = 0.0;
            for (int q = ch.p0; q < ch.p1; q++) cs += A.pass_cost[q];
            A.partc2[c] = cs;
        }
    }
}

template __global__ void k_lin_fast<float, 0>(KArgs<double>, int, int, int, int);
template __global__ void k_lin_fast<float, 1>(KArgs<double>, int, int, int, int);
template __global__ void k_lin_fast<double, 0>(KArgs<double>, int, int, int, int);
template __global__ void k_lin_fast<double, 1>(KArgs<double>, int, int, int, int);
template __global__ void k_trial_fast<0>(KArgs<double>, int, int, int, int);

}  // namespace xba
