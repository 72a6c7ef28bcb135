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
#define FB_HLP 96   // sum over a pass's landmarks of their touched slots
#define FB_LREC 20  // per landmark: Hll 6 (00 01 02 11 12 22), bl 3, Hi 6 (same order), HiB 3, ok, pad

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
    int sc, sc64, wc, wd, tile, btile, wcost, wave, wave_stride, total;
    // offsets inside one wavefront's region
    int o_oblm, o_obt, o_obact, o_rcq, o_u, o_lmask, o_lobs0, o_lnobs, o_loff, o_slots, o_hlp, o_lrec, o_ldbl, o_lrow;
};

__host__ __device__ inline FastLayout fast_layout(int R, int ts, int nw, bool trial) {
    FastLayout L;
    int o = 0;
    const int nc = trial ? 2 : 1;
    L.sc = o;    o = align16(o + nc * R * SC_STRIDE * ts);
    L.sc64 = o;  o = align16(o + nc * R * SC64_STRIDE * 8);
    L.wc = o;    o = align16(o + 32 * ts);
    L.wd = o;    o = align16(o + 8 * 8);
    L.tile = o;  o = align16(o + (trial ? 6 * R * ts : (R * (R + 1) / 2) * 36 * 8));  // trial: dp per slot
    L.btile = o; o = align16(o + (trial ? 0 : 6 * R * 8));
    L.wcost = o; o = align16(o + 16 * 8);
    L.wave = o;
    int w = 0;
    L.o_oblm = w;  w = align16(w + FB_OBS * 4);
    L.o_obt = w;   w = align16(w + FB_OBS * 4);
    L.o_obact = w; w = align16(w + FB_OBS * 4);
    L.o_rcq = w;   w = align16(w + FB_OBS * 8);
    L.o_u = w;     w = align16(w + (trial ? FB_OBS * 3 * ts : 0));
    L.o_lmask = w; w = align16(w + FB_LM * 4);
    L.o_lobs0 = w; w = align16(w + FB_LM * 4);
    L.o_lnobs = w; w = align16(w + FB_LM * 4);
    L.o_loff = w;  w = align16(w + FB_LM * 4);
    L.o_slots = w; w = align16(w + (trial ? 0 : FB_HLP * 4));
    L.o_hlp = w;   w = align16(w + (trial ? 0 : FB_HLP * 18 * ts));
    L.o_lrec = w;  w = align16(w + FB_LM * FB_LREC * ts);
    L.o_ldbl = w;  w = align16(w + FB_LM * 4 * 8);
    L.o_lrow = w;  w = align16(w + (trial ? 0 : FB_OBS * 9 * ts));  // per observation: Hll 6 + bl 3 terms
    L.wave_stride = w;
    L.total = o + nw * w;
    return L;
}

__device__ inline int wave_incl_scan(int v, int lane) {
    for (int d = 1; d < XBA_WAVE; d <<= 1) {
        int t = __shfl_up(v, d, XBA_WAVE);
        if (lane >= d) v += t;
    }
    return v;
}

// Sum over the wavefront in a fixed order (a butterfly), the result in every lane.
__device__ inline double wave_sum(double v) {
    for (int d = XBA_WAVE / 2; d >= 1; d >>= 1) v += __shfl_xor(v, d, XBA_WAVE);
    return v;
}

template <typename T>
__device__ inline void lds_add(T *p, T v) {
    __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// One pass's landmark set: as many landmarks from `lb` as fit the pass limits.  Writes lobs0/lnobs/ob_lm and
// returns the count B (>= 1) and the pass's observation count.
__device__ inline int select_pass(const int32_t *ob_start, int lb, int lend, int lane, int *lobs0, int *lnobs,
                                  int *ob_lm, int *nob_out) {
    const int l = lb + lane;
    const bool in = lane < FB_LM && l < lend;
    const int n = in ? (ob_start[l + 1] - ob_start[l]) : 1000;
    const int S = wave_incl_scan(n, lane);
    const int K = wave_incl_scan(n + 1, lane);
    const bool fit = in && S <= FB_OBS && K <= FB_HLP;
    const uint64_t bal = __ballot(fit);
    const int B = __popcll(bal);
    *nob_out = __shfl(S, B - 1, XBA_WAVE);
    if (lane < B) {
        lobs0[lane] = S - n;
        lnobs[lane] = n;
        for (int q = 0; q < n; q++) ob_lm[S - n + q] = lane;
    }
    return B;
}

// ------------------------------------------------------------------------------------------ k_lin_fast --
template <typename T>
__global__ __launch_bounds__(512) void k_lin_fast(KArgs<T> A, int Rmax) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int NW = blockDim.x / XBA_WAVE;
    const FastLayout FL = fast_layout(Rmax, (int)sizeof(T), NW, false);
    T *sc = (T *)(smem + FL.sc), *wc = (T *)(smem + FL.wc);
    double *sc64 = (double *)(smem + FL.sc64), *wd = (double *)(smem + FL.wd);
    double *tile = (double *)(smem + FL.tile), *btile = (double *)(smem + FL.btile);
    double *wcost = (double *)(smem + FL.wcost);
    const int tid = threadIdx.x, lane = tid % XBA_WAVE, wv = tid / XBA_WAVE;
    unsigned char *wb = smem + FL.wave + wv * FL.wave_stride;
    int *ob_lm = (int *)(wb + FL.o_oblm), *ob_t = (int *)(wb + FL.o_obt), *ob_act = (int *)(wb + FL.o_obact);
    double *rcq = (double *)(wb + FL.o_rcq);
    unsigned *lmask = (unsigned *)(wb + FL.o_lmask);
    int *lobs0 = (int *)(wb + FL.o_lobs0), *lnobs = (int *)(wb + FL.o_lnobs), *loff = (int *)(wb + FL.o_loff);
    int *slots = (int *)(wb + FL.o_slots);
    T *hlp = (T *)(wb + FL.o_hlp), *lrec = (T *)(wb + FL.o_lrec), *lrow = (T *)(wb + FL.o_lrow);

    for (int c = blockIdx.x; c < A.nchunks; c += gridDim.x) {
        const Chunk ch = A.chunks[c];
        const WinInfo W = A.wins[ch.win];
        const double *wp = A.win_par + (size_t)XBA_WPAR * ch.win;
        const int R = ch.R, nblk = R * (R + 1) / 2;
        const double lamd = A.lam_override >= 0.0 ? A.lam_override : A.wstate[(size_t)NWS * ch.win + WS_LAM];
        if (wv == 0) {
            load_slots<T>(sc, sc64, A.Rwb, A.pwb, wp, W.slot0, ch.lo, R, lane);
            load_wconst<T>(wc, wp, lane);
            load_wd(wd, wp, lane);
        }
        for (int i = tid; i < nblk * 36; i += blockDim.x) tile[i] = 0.0;
        for (int i = tid; i < 6 * R; i += blockDim.x) btile[i] = 0.0;
        __syncthreads();
        // this wavefront's share of the chunk
        const int nl = ch.l1 - ch.l0;
        const int wbeg = ch.l0 + (int)(((int64_t)nl * wv) / NW), wend = ch.l0 + (int)(((int64_t)nl * (wv + 1)) / NW);
        double wcst = 0.0;
        for (int lb = wbeg; lb < wend;) {
            int nob;
            const int B = select_pass(A.ob_start, lb, wend, lane, lobs0, lnobs, ob_lm, &nob);
            if (lane < B) lmask[lane] = 0u;
            WAVE_SYNC();
            // ---- observations: projection, activity, touched slots
            int q_b = 0, q_l = 0, q_k = 0, q_t = 0, q_h = 0, q_cam = 0;
            bool q_in = lane < nob;
            double a = 0, b = 0, rho = 0;
            bool q_mask = false;
            if (q_in) {
                q_b = ob_lm[lane];
                q_l = lb + q_b;
                q_k = A.ob_start[q_l] + (lane - lobs0[q_b]);
                const int li = A.lm_info[q_l];
                q_mask = (li >> 12) & 1;
                q_h = (li & 0xfff) - ch.lo;
                const int info = A.ob_info[q_k];
                q_t = (info & 0xfff) - ch.lo;
                q_cam = (info >> 12) & 1;
                const int valid = (info >> 13) & 1;
                a = A.lmp[3 * (size_t)q_l];
                b = A.lmp[3 * (size_t)q_l + 1];
                rho = A.lmp[3 * (size_t)q_l + 2];
                bool act = false;
                if (q_mask && valid) {
                    const double *Rh = sc64 + q_h * SC64_STRIDE, *Rt = sc64 + q_t * SC64_STRIDE;
                    double d0 = Rh[9] - Rt[9], d1 = Rh[10] - Rt[10], d2 = Rh[11] - Rt[11];
                    double Y0 = Rh[0] * a + Rh[1] * b + Rh[2] + rho * d0;
                    double Y1 = Rh[3] * a + Rh[4] * b + Rh[5] + rho * d1;
                    double Y2 = Rh[6] * a + Rh[7] * b + Rh[8] + rho * d2;
                    double q2 = Rt[2] * Y0 + Rt[5] * Y1 + Rt[8] * Y2;
                    act = q2 > 1e-9;
                }
                ob_act[lane] = act ? 1 : 0;
                ob_t[lane] = q_t;
                if (act && q_t != q_h) atomicOr(&lmask[q_b], (1u << q_t) | (1u << q_h));
            }
            WAVE_SYNC();
            // ---- landmarks: touched-slot lists (compact, ascending) and their offsets
            int kb = 0;
            if (lane < B) kb = __popc(lmask[lane]);
            const int koff = wave_incl_scan(kb, lane) - kb;
            if (lane < B) {
                loff[lane] = koff;
                unsigned m = lmask[lane];
                int i = 0;
                while (m) {
                    int s = __ffs(m) - 1;
                    m &= m - 1u;
                    slots[koff + i] = s;
                    for (int e = 0; e < 18; e++) hlp[(koff + i) * 18 + e] = (T)0;
                    i++;
                }
            }
            WAVE_SYNC();
            // ---- observations: residual, Huber, Jacobians; direct pose blocks into the tile; Hlp terms
            if (q_in) {
                const bool act = ob_act[lane] != 0;
                double e[2] = {0.0, 0.0}, w = 0.0, cst = 0.0, en2 = 0.0;
                T Jl[2][3], Jt[2][6], Jh[2][6];
                const int nd = q_cam == 0 ? 2 : 1;
                T *row = lrow + lane * 9;
                for (int i = 0; i < 9; i++) row[i] = (T)0;
                if (act) {
                    obs_eval<T>(sc64 + q_h * SC64_STRIDE, sc64 + q_t * SC64_STRIDE, sc + q_h * SC_STRIDE,
                                sc + q_t * SC_STRIDE, wc, wd, a, b, rho, q_cam, A.z[2 * (size_t)q_k],
                                A.z[2 * (size_t)q_k + 1], q_t == q_h, true, e, Jl, Jt, Jh);
                    const double isg = (double)A.isig[q_k];
                    for (int i = 0; i < nd; i++) {
                        e[i] *= isg;
                        en2 += e[i] * e[i];
                    }
                    huber(sqrt(en2), wd[5], &w, &cst);
                    const T isgT = A.isig[q_k], wT = (T)w;
                    T eT[2] = {(T)e[0], nd == 2 ? (T)e[1] : (T)0};
                    for (int r = 0; r < 2; r++) {
                        const T sgn = r < nd ? isgT : (T)0;
                        for (int j = 0; j < 3; j++) Jl[r][j] *= sgn;
                        for (int j = 0; j < 6; j++) {
                            Jt[r][j] *= sgn;
                            Jh[r][j] *= sgn;
                        }
                    }
                    // Hll (6) and bl (3) terms, summed per landmark in observation order below
                    row[0] = wT * (Jl[0][0] * Jl[0][0] + Jl[1][0] * Jl[1][0]);
                    row[1] = wT * (Jl[0][0] * Jl[0][1] + Jl[1][0] * Jl[1][1]);
                    row[2] = wT * (Jl[0][0] * Jl[0][2] + Jl[1][0] * Jl[1][2]);
                    row[3] = wT * (Jl[0][1] * Jl[0][1] + Jl[1][1] * Jl[1][1]);
                    row[4] = wT * (Jl[0][1] * Jl[0][2] + Jl[1][1] * Jl[1][2]);
                    row[5] = wT * (Jl[0][2] * Jl[0][2] + Jl[1][2] * Jl[1][2]);
                    for (int i = 0; i < 3; i++) row[6 + i] = wT * (Jl[0][i] * eT[0] + Jl[1][i] * eT[1]);
                    if (q_t != q_h) {
                        const unsigned m = lmask[q_b];
                        const int base = loff[q_b];
                        const int it = __popc(m & ((1u << q_t) - 1u)), ih = __popc(m & ((1u << q_h) - 1u));
                        T *At = hlp + (base + it) * 18, *Ah = hlp + (base + ih) * 18;
                        for (int i = 0; i < 3; i++)
                            for (int j = 0; j < 6; j++) {
                                lds_add<T>(At + 6 * i + j, wT * (Jl[0][i] * Jt[0][j] + Jl[1][i] * Jt[1][j]));
                                lds_add<T>(Ah + 6 * i + j, wT * (Jl[0][i] * Jh[0][j] + Jl[1][i] * Jh[1][j]));
                            }
                        // direct blocks (t,t), (h,h) lower triangles and the cross block
                        double *Btt = tile + (q_t * (q_t + 1) / 2 + q_t) * 36;
                        double *Bhh = tile + (q_h * (q_h + 1) / 2 + q_h) * 36;
                        for (int i = 0; i < 6; i++)
                            for (int j = 0; j <= i; j++) {
                                lds_add<double>(Btt + 6 * i + j,
                                                (double)(wT * (Jt[0][i] * Jt[0][j] + Jt[1][i] * Jt[1][j])));
                                lds_add<double>(Bhh + 6 * i + j,
                                                (double)(wT * (Jh[0][i] * Jh[0][j] + Jh[1][i] * Jh[1][j])));
                            }
                        if (q_t > q_h) {
                            double *Bx = tile + (q_t * (q_t + 1) / 2 + q_h) * 36;
                            for (int i = 0; i < 6; i++)
                                for (int j = 0; j < 6; j++)
                                    lds_add<double>(Bx + 6 * i + j,
                                                    (double)(wT * (Jt[0][i] * Jh[0][j] + Jt[1][i] * Jh[1][j])));
                        } else {
                            double *Bx = tile + (q_h * (q_h + 1) / 2 + q_t) * 36;
                            for (int i = 0; i < 6; i++)
                                for (int j = 0; j < 6; j++)
                                    lds_add<double>(Bx + 6 * i + j,
                                                    (double)(wT * (Jh[0][i] * Jt[0][j] + Jh[1][i] * Jt[1][j])));
                        }
                        for (int i = 0; i < 6; i++) {
                            lds_add<double>(btile + 6 * q_t + i, (double)(wT * (Jt[0][i] * eT[0] + Jt[1][i] * eT[1])));
                            lds_add<double>(btile + 6 * q_h + i, (double)(wT * (Jh[0][i] * eT[0] + Jh[1][i] * eT[1])));
                        }
                    }
                }
                rcq[lane] = act ? cst : 0.0;
                if (A.chi2_out) A.chi2_out[q_k] = act ? en2 : (double)INFINITY;
            }
            WAVE_SYNC();
            // ---- landmarks: Hll, bl in observation order, prior, damping, det guard, FP64 inverse
            double lcost = 0.0;
            if (lane < B) {
                const int l = lb + lane;
                const int li = A.lm_info[l];
                T *rec = lrec + lane * FB_LREC;
                rec[18] = (T)0;
                if ((li >> 12) & 1) {
                    const int q0 = lobs0[lane], n = lnobs[lane];
                    T s[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
                    int nact = 0;
                    for (int q = q0; q < q0 + n; q++) {
                        if (!ob_act[q]) continue;
                        nact++;
                        lcost += rcq[q];
                        for (int i = 0; i < 9; i++) s[i] += lrow[q * 9 + i];
                    }
                    double H6[6], bl[3];
                    for (int i = 0; i < 6; i++) H6[i] = (double)s[i];
                    for (int i = 0; i < 3; i++) bl[i] = (double)s[6 + i];
                    const double rho_l = A.lmp[3 * (size_t)l + 2];
                    const T ris = A.rho_isig[l];
                    if (ris > (T)0) {
                        double risd = (double)ris, r0 = (rho_l - A.rho_meas[l]) * risd, w, cc;
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
                    }
                    const double HiB0 = Hi[0] * bl[0] + Hi[1] * bl[1] + Hi[2] * bl[2];
                    const double HiB1 = Hi[3] * bl[0] + Hi[4] * bl[1] + Hi[5] * bl[2];
                    const double HiB2 = Hi[6] * bl[0] + Hi[7] * bl[1] + Hi[8] * bl[2];
                    const T hs[6] = {(T)Hi[0], (T)Hi[1], (T)Hi[2], (T)Hi[4], (T)Hi[5], (T)Hi[8]};
                    for (int i = 0; i < 6; i++) rec[9 + i] = hs[i];
                    rec[15] = (T)HiB0;
                    rec[16] = (T)HiB1;
                    rec[17] = (T)HiB2;
                    rec[18] = ok ? (T)1 : (T)0;
                    A.lmok[l] = ok ? 1 : 0;
                    if (A.ok_out) A.ok_out[l] = ok ? 1 : 0;
                    T *ho = A.lmHi + 6 * (size_t)l;
                    for (int i = 0; i < 6; i++) ho[i] = hs[i];
                    T *bo = A.lmbl + 3 * (size_t)l;
                    bo[0] = (T)bl[0];
                    bo[1] = (T)bl[1];
                    bo[2] = (T)bl[2];
                    if (A.Hinv_out)
                        for (int i = 0; i < 9; i++) A.Hinv_out[9 * (size_t)l + i] = ok ? Hi[i] : 0.0;
                } else {
                    A.lmok[l] = 0;
                }
            }
            const double pcost = wave_sum(lcost);
            if (lane == 0) wcst += pcost;
            WAVE_SYNC();
            // ---- Schur: every touched block pair of every landmark of the pass, lanes over the entries
            for (int bb = 0; bb < B; bb++) {
                const T *rec = lrec + bb * FB_LREC;
                if (rec[18] == (T)0) continue;
                const int k = __popc(lmask[bb]), base = loff[bb];
                const T h00 = rec[9], h01 = rec[10], h02 = rec[11], h11 = rec[12], h12 = rec[13], h22 = rec[14];
                const int npair = k * (k + 1) / 2;
                for (int o = lane; o < npair * 36; o += XBA_WAVE) {
                    const int p = o / 36, ij = o % 36, i = ij / 6, j = ij % 6;
                    int si = (int)((sqrtf(8.0f * p + 1.0f) - 1.0f) * 0.5f);
                    while ((si + 1) * (si + 2) / 2 <= p) si++;
                    while (si * (si + 1) / 2 > p) si--;
                    const int sj = p - si * (si + 1) / 2;
                    if (si == sj && j > i) continue;
                    const int s2 = slots[base + si], s1 = slots[base + sj];
                    const T *A2 = hlp + (base + si) * 18, *A1 = hlp + (base + sj) * 18;
                    const T c0 = A1[j], c1 = A1[6 + j], c2 = A1[12 + j];
                    const T t0 = h00 * c0 + h01 * c1 + h02 * c2;
                    const T t1 = h01 * c0 + h11 * c1 + h12 * c2;
                    const T t2 = h02 * c0 + h12 * c1 + h22 * c2;
                    const T v = A2[i] * t0 + A2[6 + i] * t1 + A2[12 + i] * t2;
                    lds_add<double>(tile + (s2 * (s2 + 1) / 2 + s1) * 36 + ij, -(double)v);
                }
                for (int o = lane; o < k * 6; o += XBA_WAVE) {
                    const int si = o / 6, i = o % 6, s = slots[base + si];
                    const T *A1 = hlp + (base + si) * 18;
                    const T v = A1[i] * rec[15] + A1[6 + i] * rec[16] + A1[12 + i] * rec[17];
                    lds_add<double>(btile + 6 * s + i, -(double)v);
                }
            }
            WAVE_SYNC();
            lb += B;
        }
        if (lane == 0) wcost[wv] = wcst;
        __syncthreads();
        double *pH = A.partH64 + ch.partH, *pb = A.partb64 + ch.partb;
        for (int i = tid; i < nblk * 36; i += blockDim.x) pH[i] = tile[i];
        for (int i = tid; i < 6 * R; i += blockDim.x) pb[i] = btile[i];
        if (tid == 0) {
            double cs = 0.0;
            for (int w = 0; w < NW; w++) cs += wcost[w];
            A.partc[c] = cs;
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------------------- k_trial_fast --
template <typename T>
__global__ __launch_bounds__(512) void k_trial_fast(KArgs<T> A, int Rmax) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int NW = blockDim.x / XBA_WAVE;
    const FastLayout FL = fast_layout(Rmax, (int)sizeof(T), NW, true);
    T *sc = (T *)(smem + FL.sc), *wc = (T *)(smem + FL.wc), *dps = (T *)(smem + FL.tile);
    double *sc64 = (double *)(smem + FL.sc64), *wd = (double *)(smem + FL.wd), *wcost = (double *)(smem + FL.wcost);
    const int tid = threadIdx.x, lane = tid % XBA_WAVE, wv = tid / XBA_WAVE;
    unsigned char *wb = smem + FL.wave + wv * FL.wave_stride;
    int *ob_lm = (int *)(wb + FL.o_oblm);
    double *rcq = (double *)(wb + FL.o_rcq), *ldbl = (double *)(wb + FL.o_ldbl);
    T *u = (T *)(wb + FL.o_u);
    int *lobs0 = (int *)(wb + FL.o_lobs0), *lnobs = (int *)(wb + FL.o_lnobs);

    for (int c = blockIdx.x; c < A.nchunks; c += gridDim.x) {
        const Chunk ch = A.chunks[c];
        const WinInfo W = A.wins[ch.win];
        if (A.wstate[(size_t)NWS * ch.win + WS_CHOLOK] == 0.0) {
            if (tid == 0) A.partc2[c] = 0.0;
            continue;
        }
        const double *wp = A.win_par + (size_t)XBA_WPAR * ch.win;
        const int R = ch.R;
        T *sc2 = sc + R * SC_STRIDE;
        double *sc642 = sc64 + R * SC64_STRIDE;
        if (wv == 0) {
            load_slots<T>(sc, sc64, A.Rwb, A.pwb, wp, W.slot0, ch.lo, R, lane);
            load_wconst<T>(wc, wp, lane);
            load_wd(wd, wp, lane);
        } else if (wv == 1 % NW) {
            load_slots<T>(sc2, sc642, A.R2, A.p2, wp, W.slot0, ch.lo, R, lane);
        }
        if (NW == 1) load_slots<T>(sc2, sc642, A.R2, A.p2, wp, W.slot0, ch.lo, R, lane);
        for (int i = tid; i < 6 * R; i += blockDim.x) {
            int s = i / 6, j = i % 6;
            dps[i] = (T)A.dx[W.st0 + A.slot_off[W.slot0 + ch.lo + s] + j];
        }
        __syncthreads();
        const int nl = ch.l1 - ch.l0;
        const int wbeg = ch.l0 + (int)(((int64_t)nl * wv) / NW), wend = ch.l0 + (int)(((int64_t)nl * (wv + 1)) / NW);
        double wcst = 0.0;
        for (int lb = wbeg; lb < wend;) {
            int nob;
            const int B = select_pass(A.ob_start, lb, wend, lane, lobs0, lnobs, ob_lm, &nob);
            WAVE_SYNC();
            // ---- observations at the current state: u_q = w Jl^T (Jt dp_t + Jh dp_h)
            int q_l = 0, q_k = 0, q_t = 0, q_h = 0, q_cam = 0, q_valid = 0;
            bool q_mask = false;
            const bool q_in = lane < nob;
            if (q_in) {
                const int qb = ob_lm[lane];
                q_l = lb + qb;
                q_k = A.ob_start[q_l] + (lane - lobs0[qb]);
                const int li = A.lm_info[q_l];
                q_mask = (li >> 12) & 1;
                q_h = (li & 0xfff) - ch.lo;
                const int info = A.ob_info[q_k];
                q_t = (info & 0xfff) - ch.lo;
                q_cam = (info >> 12) & 1;
                q_valid = (info >> 13) & 1;
                T u0 = 0, u1 = 0, u2 = 0;
                if (q_mask && q_valid && A.lmok[q_l] && q_t != q_h) {
                    const double a = A.lmp[3 * (size_t)q_l], b = A.lmp[3 * (size_t)q_l + 1],
                                 rho = A.lmp[3 * (size_t)q_l + 2];
                    double e[2];
                    T Jl[2][3], Jt[2][6], Jh[2][6];
                    if (obs_eval<T>(sc64 + q_h * SC64_STRIDE, sc64 + q_t * SC64_STRIDE, sc + q_h * SC_STRIDE,
                                    sc + q_t * SC_STRIDE, wc, wd, a, b, rho, q_cam, A.z[2 * (size_t)q_k],
                                    A.z[2 * (size_t)q_k + 1], false, true, e, Jl, Jt, Jh)) {
                        const int nd = q_cam == 0 ? 2 : 1;
                        const double isg = (double)A.isig[q_k];
                        double en2 = 0.0, w, cst;
                        for (int i = 0; i < nd; i++) en2 += (e[i] * isg) * (e[i] * isg);
                        huber(sqrt(en2), wd[5], &w, &cst);
                        const T isgT = A.isig[q_k];
                        T s[3] = {0, 0, 0};
                        for (int r = 0; r < nd; r++) {
                            T jd = 0;
                            for (int j = 0; j < 6; j++) jd += Jt[r][j] * dps[6 * q_t + j] + Jh[r][j] * dps[6 * q_h + j];
                            jd *= isgT * isgT;
                            for (int i = 0; i < 3; i++) s[i] += Jl[r][i] * jd;
                        }
                        u0 = (T)w * s[0];
                        u1 = (T)w * s[1];
                        u2 = (T)w * s[2];
                    }
                }
                u[3 * lane] = u0;
                u[3 * lane + 1] = u1;
                u[3 * lane + 2] = u2;
            }
            WAVE_SYNC();
            // ---- landmarks: back-substitution and the update rule; prior cost at the trial state
            double lcost = 0.0;
            if (lane < B) {
                const int l = lb + lane;
                const int li = A.lm_info[l];
                const double a = A.lmp[3 * (size_t)l], b = A.lmp[3 * (size_t)l + 1], rho = A.lmp[3 * (size_t)l + 2];
                double na = a, nb = b, nr = rho;
                if (((li >> 12) & 1) && A.lmok[l]) {
                    T v0 = 0, v1 = 0, v2 = 0;
                    for (int q = lobs0[lane]; q < lobs0[lane] + lnobs[lane]; q++) {
                        v0 += u[3 * q];
                        v1 += u[3 * q + 1];
                        v2 += u[3 * q + 2];
                    }
                    const T *Hi = A.lmHi + 6 * (size_t)l, *bl = A.lmbl + 3 * (size_t)l;
                    v0 += bl[0];
                    v1 += bl[1];
                    v2 += bl[2];
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
                ldbl[4 * lane] = na;
                ldbl[4 * lane + 1] = nb;
                ldbl[4 * lane + 2] = nr;
                if ((li >> 12) & 1) {
                    const T ris = A.rho_isig[l];
                    if (ris > (T)0) {
                        double r0 = (nr - A.rho_meas[l]) * (double)ris, w, cc;
                        huber(fabs(r0), wd[5], &w, &cc);
                        lcost += cc;
                    }
                }
            }
            WAVE_SYNC();
            // ---- observations at the trial state: robust cost (+1e6 behind the camera)
            if (q_in) {
                double cq = 0.0;
                if (q_mask && q_valid) {
                    const int qb = ob_lm[lane];
                    double e[2];
                    T Jl[2][3], Jt[2][6], Jh[2][6];
                    bool act = obs_eval<T>(sc642 + q_h * SC64_STRIDE, sc642 + q_t * SC64_STRIDE, sc2 + q_h * SC_STRIDE,
                                           sc2 + q_t * SC_STRIDE, wc, wd, ldbl[4 * qb], ldbl[4 * qb + 1],
                                           ldbl[4 * qb + 2], q_cam, A.z[2 * (size_t)q_k], A.z[2 * (size_t)q_k + 1],
                                           q_t == q_h, false, e, Jl, Jt, Jh);
                    if (!act) {
                        cq = 1e6;
                    } else {
                        double isg = (double)A.isig[q_k], en2 = e[0] * isg * (e[0] * isg), w;
                        if (q_cam == 0) en2 += e[1] * isg * (e[1] * isg);
                        huber(sqrt(en2), wd[5], &w, &cq);
                    }
                }
                rcq[lane] = cq;
            }
            WAVE_SYNC();
            if (lane < B)
                for (int q = lobs0[lane]; q < lobs0[lane] + lnobs[lane]; q++) lcost += rcq[q];
            const double pcost = wave_sum(lcost);
            if (lane == 0) wcst += pcost;
            WAVE_SYNC();
            lb += B;
        }
        if (lane == 0) wcost[wv] = wcst;
        __syncthreads();
        if (tid == 0) {
            double cs = 0.0;
            for (int w = 0; w < NW; w++) cs += wcost[w];
            A.partc2[c] = cs;
        }
        __syncthreads();
    }
}

template __global__ void k_lin_fast<float>(KArgs<float>, int);
template __global__ void k_lin_fast<double>(KArgs<double>, int);
template __global__ void k_trial_fast<float>(KArgs<float>, int);
template __global__ void k_trial_fast<double>(KArgs<double>, int);

}  // namespace xba
