// Per-window kernels of the window solve: the deterministic reduction of the chunk partials into each window's
// reduced system, the prior and IMU terms, the damped Cholesky solve (FP64, or FP32 with FP64 iterative
// refinement), the on-manifold pose update, and the device-side accept/reject.
//
// Reference arithmetic (paths relative to the reference's src/xclean/):
//   _full_system (scatter the 6P visual system, add prior and IMU)  vio/window.py:865-886
//   _imu_terms (Huber on the whitened norm with huber_k_dof(15))    vio/window.py:831-862
//   factor_core                                                     vio/imu_factor.py:771-875
//   Prior.terms / cost, box-minus                                   vio/marg.py:47-57, 102-127
//   damping Hd = H + diag(lam diag(H) + 1e-9), Cholesky, dx = -H^-1 b   vio/window.py:933-941
//   apply_update (R <- R Exp(dphi), p += dp; v, bg, ba additive)    vio/window.py:522-532, 945-955
//   accept iff c_new < c_tot: lam = max(lam/10, 1e-12), else lam *= 10   vio/window.py:967-981
#include "xba_internal.h"

namespace xba {

// ------------------------------------------------------------------------------------------ SO(3) helpers --
// vio/window.py:43-61
__device__ inline void so3_exp_nb(const double *w, double *R) {
    double th2 = w[0] * w[0] + w[1] * w[1] + w[2] * w[2], th = sqrt(th2), a, b;
    double K[9] = {0, -w[2], w[1], w[2], 0, -w[0], -w[1], w[0], 0};
    if (th < 1e-8) {
        a = 1.0 - th2 / 6.0;
        b = 0.5 - th2 / 24.0;
    } else {
        a = sin(th) / th;
        b = (1.0 - cos(th)) / th2;
    }
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            double kk = K[3 * i] * K[j] + K[3 * i + 1] * K[3 + j] + K[3 * i + 2] * K[6 + j];
            R[3 * i + j] = (i == j ? 1.0 : 0.0) + a * K[3 * i + j] + b * kk;
        }
}

// vio/window.py:64-105
__device__ inline void so3_log_nb(const double *R, double *w) {
    double c = (R[0] + R[4] + R[8] - 1.0) * 0.5;
    c = c > 1.0 ? 1.0 : (c < -1.0 ? -1.0 : c);
    double th = acos(c);
    double v0 = R[7] - R[5], v1 = R[2] - R[6], v2 = R[3] - R[1];
    if (th < 1e-6) {
        w[0] = 0.5 * v0; w[1] = 0.5 * v1; w[2] = 0.5 * v2;
        return;
    }
    if (th > M_PI - 1e-4) {
        double Bm[9];
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) Bm[3 * i + j] = (R[3 * i + j] + R[3 * j + i]) * 0.5;
        int i = 0;
        if (Bm[4] > Bm[4 * i]) i = 1;
        if (Bm[8] > Bm[4 * i]) i = 2;
        double den = sqrt(fmax((Bm[4 * i] - c) / (1.0 - c), 0.0)), ax[3];
        for (int j = 0; j < 3; j++) ax[j] = (Bm[3 * j + i] - (j == i ? c : 0.0)) / fmax(den * (1.0 - c), 1e-12);
        ax[i] = den;
        double n = sqrt(ax[0] * ax[0] + ax[1] * ax[1] + ax[2] * ax[2]), s = 1.0;
        if (ax[0] * v0 + ax[1] * v1 + ax[2] * v2 < 0) s = -1.0;
        for (int j = 0; j < 3; j++) w[j] = s * th * ax[j] / n;
        return;
    }
    double f = th / (2.0 * sin(th));
    w[0] = f * v0; w[1] = f * v1; w[2] = f * v2;
}

// series coefficients with the derived switch of core/lie.py:47-79 (TAYLOR_ANGLE = (7257600 eps)^0.1)
#define XBA_TAYLOR_ANGLE 0.12610312423896003
#define XBA_ANGLE_ZERO 3.650024149988857e-08
__device__ inline void abc(double t2, double *a, double *b, double *c) {
    double t = sqrt(t2);
    if (t < XBA_TAYLOR_ANGLE) {
        double x = t2;
        *a = 1.0 - x / 6.0 + x * x / 120.0 - x * x * x / 5040.0;
        *b = 0.5 - x / 24.0 + x * x / 720.0 - x * x * x / 40320.0;
        *c = 1.0 / 6.0 - x / 120.0 + x * x / 5040.0 - x * x * x / 362880.0;
    } else {
        double s = sin(t);
        *a = s / t;
        *b = (1.0 - cos(t)) / t2;
        *c = (t - s) / (t2 * t);
    }
}

// vio/imu_factor.py:214-241
__device__ inline void exp_jr(const double *v, double *R, double *Jr) {
    double x = v[0], y = v[1], z = v[2], a, b, c;
    abc(x * x + y * y + z * z, &a, &b, &c);
    double xy = x * y, xz = x * z, yz = y * z;
    double k00 = -(y * y + z * z), k11 = -(x * x + z * z), k22 = -(x * x + y * y);
    R[0] = 1.0 + b * k00; R[1] = -a * z + b * xy; R[2] = a * y + b * xz;
    R[3] = a * z + b * xy; R[4] = 1.0 + b * k11; R[5] = -a * x + b * yz;
    R[6] = -a * y + b * xz; R[7] = a * x + b * yz; R[8] = 1.0 + b * k22;
    Jr[0] = 1.0 + c * k00; Jr[1] = b * z + c * xy; Jr[2] = -b * y + c * xz;
    Jr[3] = -b * z + c * xy; Jr[4] = 1.0 + c * k11; Jr[5] = b * x + c * yz;
    Jr[6] = b * y + c * xz; Jr[7] = -b * x + c * yz; Jr[8] = 1.0 + c * k22;
}

// vio/imu_factor.py:244-263
__device__ inline void jr_inv(const double *v, double *o) {
    double x = v[0], y = v[1], z = v[2], t2 = x * x + y * y + z * z, t = sqrt(t2), co;
    if (t < XBA_TAYLOR_ANGLE)
        co = 1.0 / 12.0 + t2 / 720.0 + t2 * t2 / 30240.0 + t2 * t2 * t2 / 1209600.0;
    else
        co = (1.0 - (t * sin(t)) / (2.0 * (1.0 - cos(t)))) / t2;
    o[0] = 1.0 - co * (y * y + z * z); o[1] = -0.5 * z + co * x * y; o[2] = 0.5 * y + co * x * z;
    o[3] = 0.5 * z + co * x * y; o[4] = 1.0 - co * (x * x + z * z); o[5] = -0.5 * x + co * y * z;
    o[6] = -0.5 * y + co * x * z; o[7] = 0.5 * x + co * y * z; o[8] = 1.0 - co * (x * x + y * y);
}

// vio/imu_factor.py:266-330 (Shepperd's quaternion branch, atan2 angle)
__device__ inline void log_quat(const double *R, double *o) {
    double m00 = R[0], m11 = R[4], m22 = R[8], tr = m00 + m11 + m22, qw;
    int br = 0;
    if (!(tr > 0.0)) {
        double best = tr;
        if (m00 > best) { br = 1; best = m00; }
        if (m11 > best) { br = 2; best = m11; }
        if (m22 > best) { br = 3; best = m22; }
        if (br == 0) br = 1;
    }
    if (br == 0) {
        double s = sqrt(fmax(tr + 1.0, 0.0)) * 2.0;
        qw = 0.25 * s;
        o[0] = (R[7] - R[5]) / s; o[1] = (R[2] - R[6]) / s; o[2] = (R[3] - R[1]) / s;
    } else {
        int a = br - 1, c = (a + 1) % 3, d = (a + 2) % 3;
        double s = sqrt(fmax(1.0 + R[4 * a] - R[4 * c] - R[4 * d], 0.0)) * 2.0;
        qw = (R[3 * d + c] - R[3 * c + d]) / s;
        o[a] = 0.25 * s;
        o[c] = (R[3 * c + a] + R[3 * a + c]) / s;
        o[d] = (R[3 * d + a] + R[3 * a + d]) / s;
    }
    if (qw < 0.0) {
        qw = -qw;
        o[0] = -o[0]; o[1] = -o[1]; o[2] = -o[2];
    }
    double n = sqrt(o[0] * o[0] + o[1] * o[1] + o[2] * o[2]);
    double sc = (n < XBA_ANGLE_ZERO) ? 2.0 : 2.0 * atan2(n, qw) / n;
    o[0] *= sc; o[1] *= sc; o[2] *= sc;
}

__device__ inline void mm3(const double *A, const double *B, double *C) {
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) C[3 * i + j] = A[3 * i] * B[j] + A[3 * i + 1] * B[3 + j] + A[3 * i + 2] * B[6 + j];
}
__device__ inline void mtm3(const double *A, const double *B, double *C) {
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) C[3 * i + j] = A[i] * B[j] + A[3 + i] * B[3 + j] + A[6 + i] * B[6 + j];
}
__device__ inline void mmt3(const double *A, const double *B, double *C) {
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
            C[3 * i + j] = A[3 * i] * B[3 * j] + A[3 * i + 1] * B[3 * j + 1] + A[3 * i + 2] * B[3 * j + 2];
}

// ---------------------------------------------------------------------------------------- the IMU factor --
// Unwhitened residual r (15) and Jacobians Ji, Jj (15 x 15, row-major) of vio/imu_factor.py:771-846.
__device__ void imu_raw(const double *Ri, const double *pi, const double *vi, const double *bgi, const double *bai,
                        const double *Rj, const double *pj, const double *vj, const double *bgj, const double *baj,
                        const double *g, const double *rec, double *r, double *Ji, double *Jj) {
    const double *dR = rec, *dv = rec + 9, *dp = rec + 12, dt = rec[15], *J = rec + 16, *bg_lin = rec + 70,
                 *ba_lin = rec + 73;
    double dbg[3], dba[3], phib[3], rR[3], uv[3], up[3], Dv[3], Dp[3];
    double Eb[9], Jrb[9], dRc[9], Mij[9], E[9], Jri[9], T1[9], T2[9], T3[9], T4[9];
    for (int c = 0; c < 3; c++) {
        dbg[c] = bgi[c] - bg_lin[c];
        dba[c] = bai[c] - ba_lin[c];
    }
    for (int i = 0; i < 3; i++) phib[i] = J[6 * i] * dbg[0] + J[6 * i + 1] * dbg[1] + J[6 * i + 2] * dbg[2];
    exp_jr(phib, Eb, Jrb);
    mm3(dR, Eb, dRc);
    mtm3(Ri, Rj, Mij);
    mtm3(dRc, Mij, E);
    log_quat(E, rR);
    double dt2 = dt * dt;
    for (int c = 0; c < 3; c++) {
        Dv[c] = vj[c] - vi[c] - g[c] * dt;
        Dp[c] = pj[c] - pi[c] - vi[c] * dt - 0.5 * g[c] * dt2;
    }
    for (int i = 0; i < 3; i++) {
        uv[i] = Ri[i] * Dv[0] + Ri[3 + i] * Dv[1] + Ri[6 + i] * Dv[2];
        up[i] = Ri[i] * Dp[0] + Ri[3 + i] * Dp[1] + Ri[6 + i] * Dp[2];
    }
    for (int i = 0; i < 3; i++) {
        double cv = dv[i], cp = dp[i];
        for (int k = 0; k < 3; k++) {
            cv += J[6 * (3 + i) + k] * dbg[k] + J[6 * (3 + i) + 3 + k] * dba[k];
            cp += J[6 * (6 + i) + k] * dbg[k] + J[6 * (6 + i) + 3 + k] * dba[k];
        }
        r[i] = rR[i];
        r[3 + i] = uv[i] - cv;
        r[6 + i] = up[i] - cp;
        r[9 + i] = bgj[i] - bgi[i];
        r[12 + i] = baj[i] - bai[i];
    }
    if (!Ji) return;
    jr_inv(rR, Jri);
    mmt3(Jri, Mij, T1);
    mmt3(Jri, E, T2);
    mm3(T2, Jrb, T3);
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) T4[3 * i + j] = T3[3 * i] * J[j] + T3[3 * i + 1] * J[6 + j] + T3[3 * i + 2] * J[12 + j];
    for (int i = 0; i < 225; i++) Ji[i] = Jj[i] = 0.0;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            double RiT = Ri[3 * j + i];
            Ji[15 * i + j] = -T1[3 * i + j];
            Ji[15 * i + 9 + j] = -T4[3 * i + j];
            Jj[15 * i + j] = Jri[3 * i + j];
            Ji[15 * (3 + i) + 6 + j] = -RiT;
            Ji[15 * (3 + i) + 9 + j] = -J[6 * (3 + i) + j];
            Ji[15 * (3 + i) + 12 + j] = -J[6 * (3 + i) + 3 + j];
            Ji[15 * (6 + i) + 3 + j] = -RiT;
            Ji[15 * (6 + i) + 6 + j] = -RiT * dt;
            Ji[15 * (6 + i) + 9 + j] = -J[6 * (6 + i) + j];
            Ji[15 * (6 + i) + 12 + j] = -J[6 * (6 + i) + 3 + j];
            Jj[15 * (3 + i) + 6 + j] = RiT;
            Jj[15 * (6 + i) + 3 + j] = RiT;
        }
    Ji[15 * 3 + 1] = -uv[2]; Ji[15 * 3 + 2] = uv[1];
    Ji[15 * 4 + 0] = uv[2]; Ji[15 * 4 + 2] = -uv[0];
    Ji[15 * 5 + 0] = -uv[1]; Ji[15 * 5 + 1] = uv[0];
    Ji[15 * 6 + 1] = -up[2]; Ji[15 * 6 + 2] = up[1];
    Ji[15 * 7 + 0] = up[2]; Ji[15 * 7 + 2] = -up[0];
    Ji[15 * 8 + 0] = -up[1]; Ji[15 * 8 + 1] = up[0];
    for (int i = 0; i < 3; i++) {
        Ji[15 * (9 + i) + 9 + i] = -1.0;
        Ji[15 * (12 + i) + 12 + i] = -1.0;
        Jj[15 * (9 + i) + 9 + i] = 1.0;
        Jj[15 * (12 + i) + 12 + i] = 1.0;
    }
}

// IMU factor output record: blocks Ji^T Ji, Jj^T Jj, Ji^T Jj (3 x 225), gradients (2 x 15), cost.
#define IMU_O_B 675
#define IMU_O_COST 705
#define IMU_OUT 706

struct DArgs {
    int32_t nwin;
    const Chunk *chunks;
    const WinInfo *wins;
    const int32_t *colmap;      // (sum n): slot*8 + i for pose columns, -1 otherwise
    const int32_t *slot_off;    // (nslot)
    const int32_t *slot_dim;    // (nslot)
    const int32_t *imu_ij;      // (nimu,2)
    const int32_t *prior_slot, *prior_dim;
    const double *imu_rec, *prior_x0, *prior_S, *prior_r, *win_par;
    double *Rwb, *pwb, *vel, *bgs, *bas;  // state (nslot,...)
    double *R2, *p2, *v2, *g2, *a2;       // trial state
    double *H, *b, *dx, *fact64;          // (sum n^2), (sum n), (sum n), (sum n^2)
    float *fact32;                        // (sum n^2)
    double *imu_out;                      // (nimu, IMU_OUT) per-factor weighted blocks, gradients and cost
    const int32_t *imu_win;               // (nimu) window of every IMU factor
    const int32_t *imu_color;             // (nimu) colour of every IMU factor (no shared slot within a colour)
    double *wstate;
    const double *partc, *partc2;
    double *imu_chi2;                     // (nimu) or null: whitened chi2 of every IMU factor
    int32_t nref_max;
    int32_t fallback_on;                  // mixed precision: windows whose FP32 solve fails go to FP64
    int32_t only_fallback;                // k_reduce / k_terms: only the windows flagged WS_FALLBACK
    int32_t nbig;                         // windows solved by the tiled Cholesky
    const BigWin *bigs;                   // (nbig)
    float *bigF32;                        // tiled factors (sum np^2), mixed precision
    double *bigF64;                       // tiled factors (sum np^2), FP64
    double *bigS;                         // (sum np) Jacobi scaling of the tiled factors
    double *bigV;                         // (sum 3 np) refinement vectors of the tiled solves
    int32_t big64;                        // mixed precision: the tiled windows are factored in FP64 (one solve)
    const int32_t *chunk_slots;           // tile slot lists of the work items with smap >= 0
    // reduction lists: for every slot pair (s2 >= s1) of a window, the partial blocks (index into the partial
    // buffer in units of 36 entries) that hold it, in work-item order; for every slot, its partial gradient
    // segments (units of 6 entries)
    const int32_t *pair_start, *pair_ent, *bslot_start, *bslot_ent;
    // the FP64 partials of the multi-wavefront kernels: the mixed fallback (fallback_relin) reduces a flagged
    // window's re-assembled system inside its FP64 solve
    const double *partH64, *partb64;
    int32_t fallback_relin;
};

// ------------------------------------------------------------------------------------------- k_reduce --
// H[w] (n x n, FP64 out) = sum over the window's chunks, in chunk order, of their lower-triangular partial tiles
// (mirrored), with zeros in the non-pose columns; b likewise; the visual cost = sum of the chunk costs.  Every entry
// visits only the partial blocks that hold its slot pair (the reduction lists built at planning).
template <typename T>
struct Acc {  // FP64: plain sum; FP32: Kahan-compensated sum
    T s = 0, c = 0;
    __device__ inline void add(T v) {
        if (sizeof(T) == 4) {
            T y = v - c;
            T t = s + y;
            c = (t - s) - y;
            s = t;
        } else {
            s += v;
        }
    }
    __device__ inline double value() const { return (double)s - (double)c; }
};

// Deterministic sum over the block: thread t adds v[i] for i = a0 + t, a0 + t + nt, ... (and likewise over [b0, b1)) in
// order, then a butterfly inside each wavefront and the wavefronts' partials in order.  Called by every thread of the
// block; every thread gets the sum.  sred: blockDim.x / 64 doubles of LDS.
__device__ inline double block_sum_ranges(const double *v, int a0, int a1, int b0, int b1, double *sred) {
    const int tid = threadIdx.x, nt = blockDim.x;
    double s = 0.0;
    for (int c = a0 + tid; c < a1; c += nt) s += v[c];
    for (int c = b0 + tid; c < b1; c += nt) s += v[c];
    for (int d = 32; d >= 1; d >>= 1) s += __shfl_xor(s, d, 64);
    __syncthreads();
    if ((tid & 63) == 0) sred[tid >> 6] = s;
    __syncthreads();
    double t = 0.0;
    for (int q = 0; q < nt / 64; q++) t += sred[q];
    __syncthreads();
    return t;
}

template <typename T>
__device__ inline void reduce_window(const DArgs &D, int w, int64_t idx, const T *partH, const T *partb,
                                     const T *partH_lo, const T *partb_lo) {
    const WinInfo W = D.wins[w];
    const int n = W.n;
    const bool kah = sizeof(T) == 4;
    if (idx < (int64_t)n * n) {
        int a = (int)(idx / n), bcol = (int)(idx % n);
        int ca = D.colmap[W.st0 + a], cb = D.colmap[W.st0 + bcol];
        Acc<T> v;
        if (ca >= 0 && cb >= 0) {
            int s2 = ca >> 3, i = ca & 7, s1 = cb >> 3, j = cb & 7;
            if (s2 < s1 || (s2 == s1 && i < j)) {
                int t = s2; s2 = s1; s1 = t;
                t = i; i = j; j = t;
            }
            const int pr = W.pair0 + s2 * (s2 + 1) / 2 + s1;
            const int q0 = D.pair_start[pr], q1 = D.pair_start[pr + 1];
            // eight list entries' loads issued together, then added in list order
            int q = q0;
            for (; q + 8 <= q1; q += 8) {
                T h[8], hl[8];
#pragma unroll
                for (int u = 0; u < 8; u++) {
                    const int64_t e = (int64_t)D.pair_ent[q + u] * 36 + 6 * i + j;
                    h[u] = partH[e];
                    hl[u] = kah ? partH_lo[e] : (T)0;
                }
#pragma unroll
                for (int u = 0; u < 8; u++) {
                    v.add(h[u]);
                    if (kah) v.add(hl[u]);
                }
            }
            for (; q < q1; q++) {
                const int64_t e = (int64_t)D.pair_ent[q] * 36 + 6 * i + j;
                v.add(partH[e]);
                if (kah) v.add(partH_lo[e]);
            }
        }
        D.H[W.h0 + idx] = v.value();
    }
    if (idx < n) {
        int ca = D.colmap[W.st0 + idx];
        Acc<T> v;
        if (ca >= 0) {
            const int s = ca >> 3, i = ca & 7, gs = W.slot0 + s;
            const int q1 = D.bslot_start[gs + 1];
            int q = D.bslot_start[gs];
            for (; q + 8 <= q1; q += 8) {  // as above: eight entries' loads issued together, added in list order
                T h[8], hl[8];
#pragma unroll
                for (int u = 0; u < 8; u++) {
                    const int64_t e = (int64_t)D.bslot_ent[q + u] * 6 + i;
                    h[u] = partb[e];
                    hl[u] = kah ? partb_lo[e] : (T)0;
                }
#pragma unroll
                for (int u = 0; u < 8; u++) {
                    v.add(h[u]);
                    if (kah) v.add(hl[u]);
                }
            }
            for (; q < q1; q++) {
                const int64_t e = (int64_t)D.bslot_ent[q] * 6 + i;
                v.add(partb[e]);
                if (kah) v.add(partb_lo[e]);
            }
        }
        D.b[W.st0 + idx] = v.value();
    }
}

// The window's robust visual cost, the work items' pass-cost sums added over the block (fixed order).  Called by
// every thread of a block; thread 0 stores it.
__device__ inline void window_visual_cost(const DArgs &D, int w, double *sred) {
    const WinInfo W = D.wins[w];
    const double cs = block_sum_ranges(D.partc, W.c0, W.c1, W.c2, W.c3, sred);
    if (threadIdx.x == 0) D.wstate[(size_t)NWS * w + WS_COSTVIS] = cs;
}

// grid (ceil(max n^2 / 256), nwin): entry blockIdx.x * 256 + threadIdx.x of window blockIdx.y.  With
// D.only_fallback (grid (.., 1)): that entry of every window flagged WS_FALLBACK, the flags read by every block.
template <typename T>
__global__ void k_reduce(DArgs D, const T *partH, const T *partb, const T *partH_lo, const T *partb_lo) {
    __shared__ double sred[16];
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (!D.only_fallback) {
        reduce_window<T>(D, blockIdx.y, idx, partH, partb, partH_lo, partb_lo);
        if (blockIdx.x == 0) window_visual_cost(D, blockIdx.y, sred);
        return;
    }
    for (int w = 0; w < D.nwin; w++)
        if (D.wstate[(size_t)NWS * w + WS_FALLBACK] != 0.0) {
            reduce_window<T>(D, w, idx, partH, partb, partH_lo, partb_lo);
            if (blockIdx.x == 0) window_visual_cost(D, w, sred);
        }
}

// --------------------------------------------------------------------------------------------- k_imu --
// One workgroup per IMU factor: raw residual and Jacobians (vio/imu_factor.py:771-875), whitening by the factor's
// square-root information W, Huber at huber_k_dof(15) on the whitened norm (vio/window.py:831-862), and the factor's
// weighted blocks w Ji^T Ji, w Jj^T Jj, w Ji^T Jj, gradients w J^T r and cost into its output record (trial == 1:
// cost only, at the trial state).  k_terms adds them to the window's system in factor order.
__global__ __launch_bounds__(256) void k_imu(DArgs D, int trial) {
    __shared__ double sr[15], sJi[225], sJj[225];
    __shared__ double scr[465];  // unwhitened residual and Jacobians from the first thread
    __shared__ double se;
    const int fi = blockIdx.x, tid = threadIdx.x, nt = blockDim.x;
    const int w = D.imu_win[fi];
    const WinInfo W = D.wins[w];
    const double *R = trial ? D.R2 : D.Rwb, *p = trial ? D.p2 : D.pwb;
    const double *v = trial ? D.v2 : D.vel, *bg = trial ? D.g2 : D.bgs, *ba = trial ? D.a2 : D.bas;
    const double kd = D.win_par[(size_t)XBA_WPAR * w + 6];
    const double *g = D.win_par + (size_t)XBA_WPAR * w + 20;
    double *o = D.imu_out + (size_t)IMU_OUT * fi;
    const double *Wm = D.imu_rec + (size_t)XBA_IMU_REC * fi + 76;
    if (tid == 0) {
        const int i = W.slot0 + D.imu_ij[2 * fi], j = W.slot0 + D.imu_ij[2 * fi + 1];
        imu_raw(R + 9 * (size_t)i, p + 3 * (size_t)i, v + 3 * (size_t)i, bg + 3 * (size_t)i, ba + 3 * (size_t)i,
                R + 9 * (size_t)j, p + 3 * (size_t)j, v + 3 * (size_t)j, bg + 3 * (size_t)j, ba + 3 * (size_t)j, g,
                D.imu_rec + (size_t)XBA_IMU_REC * fi, scr, trial ? nullptr : scr + 15, trial ? nullptr : scr + 240);
    }
    __syncthreads();
    // whiten: X <- W X (W lower triangular)
    for (int e = tid; e < 15; e += nt) {
        double acc = 0.0;
        for (int k = 0; k <= e; k++) acc += Wm[15 * e + k] * scr[k];
        sr[e] = acc;
    }
    if (!trial)
        for (int e = tid; e < 450; e += nt) {
            const int which = e / 225, ic = e % 225, i = ic / 15, c = ic % 15;
            const double *X = scr + 15 + 225 * which;
            double acc = 0.0;
            for (int k = 0; k <= i; k++) acc += Wm[15 * i + k] * X[15 * k + c];
            (which ? sJj : sJi)[ic] = acc;
        }
    __syncthreads();
    if (tid == 0) {
        double e2 = 0.0;
        for (int k = 0; k < 15; k++) e2 += sr[k] * sr[k];
        se = e2;
        const double e = sqrt(e2);
        o[IMU_O_COST] = (e <= kd) ? 0.5 * e2 : kd * (e - 0.5 * kd);
        if (D.imu_chi2 && !trial) D.imu_chi2[fi] = e2;
    }
    __syncthreads();
    if (trial) return;
    const double e = sqrt(se), wgt = (e <= kd) ? 1.0 : kd / e;
    for (int q = tid; q < 3 * 225; q += nt) {
        const int which = q / 225, ac = q % 225, a = ac / 15, c = ac % 15;
        const double *X = which == 1 ? sJj : sJi, *Y = which == 0 ? sJi : sJj;
        double acc = 0.0;
        for (int k = 0; k < 15; k++) acc += X[15 * k + a] * Y[15 * k + c];
        o[q] = wgt * acc;
    }
    for (int q = tid; q < 30; q += nt) {
        const double *X = q < 15 ? sJi : sJj;
        const int a = q % 15;
        double acc = 0.0;
        for (int k = 0; k < 15; k++) acc += X[15 * k + a] * sr[k];
        o[IMU_O_B + q] = wgt * acc;
    }
}

// Sum and max over a wavefront (butterfly), the result in every lane.
__device__ inline double dwave_sum(double v) {
    for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
    return v;
}
__device__ inline double dwave_max(double v) {
    for (int d = 32; d >= 1; d >>= 1) v = fmax(v, __shfl_xor(v, d, 64));
    return v;
}

#define IMU_BATCH 8  // IMU block entries a thread scatters at a time

// ------------------------------------------------------------------------------------------- k_terms --
// The prior and IMU terms of one window (one workgroup per window).  trial == 0: add H, b and the cost at the
// current state, then c_tot = visual + rest (cost0 on the first call); trial == 1: cost only, at the trial state.
// The terms of window w by the calling workgroup; sd, sres (doubles) and scol (ints) are XBA_MAX_STATE entries of
// LDS scratch each.
__device__ inline void terms_window(const DArgs &D, int w, int trial, double *sd, double *sres, int *scol) {
    const int tid = threadIdx.x, nt = blockDim.x;
    const WinInfo W = D.wins[w];
    const int n = W.n;
    const double *R = trial ? D.R2 : D.Rwb, *p = trial ? D.p2 : D.pwb;
    const double *v = trial ? D.v2 : D.vel, *bg = trial ? D.g2 : D.bgs, *ba = trial ? D.a2 : D.bas;
    double *H = D.H + W.h0, *b = D.b + W.st0;
    double cost = 0.0;  // thread 0
    // ---- prior: d = x [-] x0 per block, res = S d + r
    if (W.pblk1 > W.pblk0) {
        int np_ = W.np, m = W.prow1 - W.prow0;
        if (tid == 0) {
            int o = 0;
            for (int k = W.pblk0; k < W.pblk1; k++) {
                int s = D.prior_slot[k], dim = D.prior_dim[k], gs = W.slot0 + s;
                const double *x0 = D.prior_x0 + 21 * (size_t)k;
                double M[9];
                mtm3(x0, R + 9 * (size_t)gs, M);
                so3_log_nb(M, sd + o);
                for (int i = 0; i < 3; i++) sd[o + 3 + i] = p[3 * (size_t)gs + i] - x0[9 + i];
                if (dim == 15)
                    for (int i = 0; i < 3; i++) {
                        sd[o + 6 + i] = v[3 * (size_t)gs + i] - x0[12 + i];
                        sd[o + 9 + i] = bg[3 * (size_t)gs + i] - x0[15 + i];
                        sd[o + 12 + i] = ba[3 * (size_t)gs + i] - x0[18 + i];
                    }
                for (int i = 0; i < dim; i++) scol[o + i] = D.slot_off[gs] + i;
                o += dim;
            }
        }
        __syncthreads();
        const double *S = D.prior_S + W.prS0;
        for (int r = tid; r < m; r += nt) {
            double acc = 0.0;
            for (int c = 0; c < np_; c++) acc += S[(size_t)r * np_ + c] * sd[c];
            sres[r] = acc + D.prior_r[W.prow0 + r];
        }
        __syncthreads();
        if (tid == 0) {
            double cs = 0.0;
            for (int r = 0; r < m; r++) cs += sres[r] * sres[r];
            cost += 0.5 * cs;
        }
        if (!trial) {
            for (int e = tid; e < np_ * np_; e += nt) {
                int a = e / np_, c = e % np_;
                double acc = 0.0;
                for (int r = 0; r < m; r++) acc += S[(size_t)r * np_ + a] * S[(size_t)r * np_ + c];
                H[(size_t)scol[a] * n + scol[c]] += acc;
            }
            for (int a = tid; a < np_; a += nt) {
                double acc = 0.0;
                for (int r = 0; r < m; r++) acc += S[(size_t)r * np_ + a] * sres[r];
                b[scol[a]] += acc;
            }
        }
        __syncthreads();
    }
    // ---- IMU factors (vio/window.py:831-862): their weighted whitened blocks come from k_imu.  The cost is summed
    // in a fixed order; the blocks are added one colour at a time (factors of one colour share no slot, colours
    // assigned in factor order at planning), so every H entry receives its terms in a fixed order.
    const int nf = W.imu1 - W.imu0;
    if (tid < 64) {  // the first wavefront sums the factors' costs in a fixed order: lane-strided, then the lanes
        double ic = 0.0;
        for (int f = tid; f < nf; f += 64) ic += D.imu_out[(size_t)IMU_OUT * (W.imu0 + f) + IMU_O_COST];
        ic = dwave_sum(ic);
        if (tid == 0) cost += ic;
    }
    if (!trial && nf > 0) {
        // per factor its colour and its two slots' columns, staged in LDS (scol is free after the prior) when they
        // fit, so that the scatter below needs one global round trip per batch
        const bool staged = 3 * nf <= XBA_MAX_STATE;
        int *fcol = scol, *fci = scol + nf, *fcj = scol + 2 * nf;
        if (staged) {
            for (int f = tid; f < nf; f += nt) {
                const int fi = W.imu0 + f;
                fcol[f] = D.imu_color[fi];
                fci[f] = D.slot_off[W.slot0 + D.imu_ij[2 * fi]];
                fcj[f] = D.slot_off[W.slot0 + D.imu_ij[2 * fi + 1]];
            }
            __syncthreads();
        }
        for (int col = 0; col < W.imu_ncol; col++) {
            // IMU_BATCH entries per thread at a time: their targets and values loaded first, then the sums stored
            // (each target entry receives one term per colour, so the order within it is free)
            for (int q0 = tid; q0 < nf * 705; q0 += IMU_BATCH * nt) {
                double val[IMU_BATCH], t1[IMU_BATCH], t2[IMU_BATCH];
                double *d1[IMU_BATCH], *d2[IMU_BATCH];
#pragma unroll
                for (int u = 0; u < IMU_BATCH; u++) {
                    const int q = q0 + u * nt;
                    d1[u] = d2[u] = nullptr;
                    val[u] = 0.0;
                    if (q >= nf * 705) continue;
                    const int f = q / 705, e = q % 705, fi = W.imu0 + f;
                    int ci, cj;
                    if (staged) {
                        if (fcol[f] != col) continue;
                        ci = fci[f];
                        cj = fcj[f];
                    } else {
                        if (D.imu_color[fi] != col) continue;
                        ci = D.slot_off[W.slot0 + D.imu_ij[2 * fi]];
                        cj = D.slot_off[W.slot0 + D.imu_ij[2 * fi + 1]];
                    }
                    val[u] = D.imu_out[(size_t)IMU_OUT * fi + e];
                    if (e >= IMU_O_B) {
                        d1[u] = b + (e - IMU_O_B < 15 ? ci : cj) + (e - IMU_O_B) % 15;
                        continue;
                    }
                    const int which = e / 225, ac = e % 225, a = ac / 15, c = ac % 15;
                    if (which == 0) {
                        d1[u] = H + (size_t)(ci + a) * n + ci + c;
                    } else if (which == 1) {
                        d1[u] = H + (size_t)(cj + a) * n + cj + c;
                    } else {
                        d1[u] = H + (size_t)(ci + a) * n + cj + c;
                        d2[u] = H + (size_t)(cj + c) * n + ci + a;
                    }
                }
#pragma unroll
                for (int u = 0; u < IMU_BATCH; u++) {
                    if (d1[u]) t1[u] = *d1[u];
                    if (d2[u]) t2[u] = *d2[u];
                }
#pragma unroll
                for (int u = 0; u < IMU_BATCH; u++) {
                    if (d1[u]) *d1[u] = t1[u] + val[u];
                    if (d2[u]) *d2[u] = t2[u] + val[u];
                }
            }
            __syncthreads();
        }
    }
    if (tid == 0) {
        double *ws = D.wstate + (size_t)NWS * w;
        if (trial) {
            ws[WS_COSTTRIAL_REST] = cost;
        } else {
            ws[WS_COSTREST] = cost;
            double ct = ws[WS_COSTVIS] + cost;
            ws[WS_CTOT] = ct;
            ws[WS_CUR] = ct;
            if (ws[WS_HAVE0] == 0.0) {
                ws[WS_COST0] = ct;
                ws[WS_HAVE0] = 1.0;
            }
        }
    }
}

__global__ __launch_bounds__(1024) void k_terms(DArgs D, int trial) {
    __shared__ double sd[XBA_MAX_STATE], sres[XBA_MAX_STATE];
    __shared__ int scol[XBA_MAX_STATE];
    terms_window(D, blockIdx.x, trial, sd, sres, scol);
}

// ---------------------------------------------------------------------------------------- dense Cholesky --
// Orders the LDS accesses of one wavefront's lanes (a lane reads what another lane of the wavefront wrote).
#define CH_WAVE_SYNC()                                         \
    do {                                                       \
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront"); \
        __builtin_amdgcn_s_waitcnt(0xc07f);                    \
        __builtin_amdgcn_wave_barrier();                       \
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront"); \
    } while (0)

// Broadcast of lane jj's value (jj uniform) through the scalar unit.
__device__ inline float lane_bcast(float v, int jj) { return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), jj)); }
__device__ inline double lane_bcast(double v, int jj) {
    const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)u, jj);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(u >> 32), jj);
    return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}

// In-place lower Cholesky of the n x n row-major A (leading dimension ld) by one workgroup (a multiple of 64
// threads), right-looking in panels of CH_BS columns:
//   1. the panel's diagonal block in the first wavefront's registers (lane i holds row i of the block): per
//      column the pivot and then each entry of the scaled column are broadcast from the lanes' registers through
//      the scalar unit (no LDS round trip in the column loop); reciprocal pivots into sinv;
//   2. the rows below solved against it, one row per thread in registers (the block read at wave-uniform
//      addresses, so every read is a broadcast), written to the lower triangle and, transposed, to the unused
//      upper triangle (row j0 + m holds column m of the panel, contiguous in the row index);
//   3. the trailing lower triangle updated in 4 x 4 register tiles from the transposed copy: a tile's 4 rows
//      and 4 columns are 4 consecutive entries of one row, one 4-wide load each when ld is a multiple of 4
//      (consecutive threads then read consecutive 16-byte groups), and each loaded entry feeds 4 products.
// The lower triangle holds the factor on return; the upper triangle is scratch.  Every thread reads the pivot
// status, so a failure exits uniformly.  Returns false on a non-positive or non-finite pivot.  Inlined, so an
// LDS-resident A gets LDS instructions.  sinv: CH_SCRATCH entries of LDS (the pivots' inverse square roots in the
// first CH_BS).
#define CH_BS 32
#define CH_SCRATCH (2 * CH_BS)
template <typename T>
__device__ __attribute__((always_inline)) inline bool chol_wg(T *A, int n, int *sflag, T *sinv, int ld) {
    const int tid = threadIdx.x, nt = blockDim.x, lane = tid & 63, wv = tid >> 6;
    for (int j0 = 0; j0 < n; j0 += CH_BS) {
        const int nb = min(CH_BS, n - j0), j1 = j0 + nb;
        if (wv == 0) {
            // lane i < nb: row j0 + i of the block, whole (its entries right of the diagonal are scratch: they only
            // ever reach entries right of the diagonal); lanes past nb hold zeros and columns past nb take a unit
            // pivot, so those steps change nothing.  No per-lane masks: every condition is on uniform values.
            const bool in = lane < nb;
            T row[CH_BS];
#pragma unroll
            for (int jj = 0; jj < CH_BS; jj++) row[jj] = (jj < nb && in) ? A[(size_t)(j0 + lane) * ld + j0 + jj] : (T)0;
            int ok = 1;
#pragma unroll
            for (int jj = 0; jj < CH_BS; jj++) {
                T d = jj < nb ? lane_bcast(row[jj], jj) : (T)1;
                if (!(d > (T)0 && isfinite(d))) {  // (uniform) recorded; the sweep goes on with a unit pivot
                    ok = 0;
                    d = (T)1;
                }
                // l = L[lane][jj] for lane > jj; lane jj gets d / sqrt(d), the pivot (within an ulp of sqrt(d)).  The
                // FP32 factor (refined in FP64 against the FP64 matrix) takes 1/sqrt(d) from the hardware
                // approximation (within an ulp), the FP64 factor's is correctly rounded.
                T s;
                if constexpr (sizeof(T) == 4)
                    s = __builtin_amdgcn_rsqf(d);
                else
                    s = (T)1 / sqrt(d);
                const T l = row[jj] * s;
                row[jj] = l;
                if (lane == 0) sinv[jj] = s;
                // row[k] -= L[lane][jj] L[k][jj], L[k][jj] broadcast from lane k's register; entries right of the
                // diagonal are scratch, never stored.  The broadcasts are issued in groups of 8 ahead of their
                // multiply-adds (a broadcast's scalar result read by the next instruction stalls the wavefront).
#pragma unroll
                for (int k0 = jj + 1; k0 < CH_BS; k0 += 8) {
                    T bc[8];
#pragma unroll
                    for (int e = 0; e < 8; e++)
                        if (k0 + e < CH_BS) bc[e] = lane_bcast(l, k0 + e);
                    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                    for (int e = 0; e < 8; e++)
                        if (k0 + e < CH_BS) {
                            row[k0 + e] -= l * bc[e];
                            // the update happens here: otherwise the compiler defers each row[k] to its use at
                            // step k and keeps every earlier column's values live until then
                            asm volatile("" : "+v"(row[k0 + e]));
                        }
                }
            }
            // whole rows back: the part right of the diagonal is never read as factor
#pragma unroll
            for (int jj = 0; jj < CH_BS; jj++)
                if (jj < nb && in) A[(size_t)(j0 + lane) * ld + j0 + jj] = row[jj];
            if (lane == 0) *sflag = ok;
        }
        __syncthreads();
        if (!*sflag) return false;
        // rows below the block (present only when the block is full, nb = CH_BS): L[i][j0:j1] = A[i][j0:j1] L^-T
        for (int i = j1 + tid; i < n; i += nt) {
            T *Ai = A + (size_t)i * ld + j0;
            T x[CH_BS];
#pragma unroll
            for (int jj = 0; jj < CH_BS; jj++) x[jj] = Ai[jj];
#pragma unroll
            for (int jj = 0; jj < CH_BS; jj++) {
                const T *Lj = A + (size_t)(j0 + jj) * ld + j0;
                T v = x[jj];
#pragma unroll
                for (int m = 0; m < jj; m++) v -= x[m] * Lj[m];
                x[jj] = v * sinv[jj];
            }
#pragma unroll
            for (int jj = 0; jj < CH_BS; jj++) {
                Ai[jj] = x[jj];
                A[(size_t)(j0 + jj) * ld + i] = x[jj];
            }
        }
        __syncthreads();
        // trailing lower triangle: tile (bi, bk), bk <= bi, covers rows j1 + 4 bi.. and columns j1 + 4 bk..
        const int M = n - j1, mt = (M + 3) >> 2, ntile = mt * (mt + 1) / 2;
        for (int t = tid; t < ntile; t += nt) {
            int bi = (int)((sqrtf(8.0f * (float)t + 1.0f) - 1.0f) * 0.5f);
            while ((bi + 1) * (bi + 2) / 2 <= t) bi++;
            while (bi * (bi + 1) / 2 > t) bi--;
            const int bk = t - bi * (bi + 1) / 2;
            const int i0 = j1 + 4 * bi, k0 = j1 + 4 * bk;
            T acc[4][4];
#pragma unroll
            for (int r = 0; r < 4; r++)
#pragma unroll
                for (int c = 0; c < 4; c++) acc[r][c] = (T)0;
            if ((ld & 3) == 0) {
                // j1 is a multiple of CH_BS here (a panel with rows below is full), so i0, k0 and the row starts are
                // multiples of 4: aligned 4-wide reads.  Entries past n lie in the row padding (ld >= n rounded up
                // to 4); they only reach rows / columns that are not written.
                typedef T T4 __attribute__((ext_vector_type(4)));
                for (int m = 0; m < nb; m++) {
                    const T *Pm = A + (size_t)(j0 + m) * ld;
                    const T4 a = *(const T4 *)(Pm + i0), b = *(const T4 *)(Pm + k0);
                    const T av[4] = {a.x, a.y, a.z, a.w}, bv[4] = {b.x, b.y, b.z, b.w};
#pragma unroll
                    for (int r = 0; r < 4; r++)
#pragma unroll
                        for (int c = 0; c < 4; c++) acc[r][c] += av[r] * bv[c];
                }
            } else {
                for (int m = 0; m < nb; m++) {
                    const T *Pm = A + (size_t)(j0 + m) * ld;
                    T av[4], bv[4];
#pragma unroll
                    for (int r = 0; r < 4; r++) {
                        av[r] = Pm[min(i0 + r, n - 1)];  // rows past n: not written
                        bv[r] = Pm[min(k0 + r, n - 1)];
                    }
#pragma unroll
                    for (int r = 0; r < 4; r++)
#pragma unroll
                        for (int c = 0; c < 4; c++) acc[r][c] += av[r] * bv[c];
                }
            }
#pragma unroll
            for (int r = 0; r < 4; r++)
#pragma unroll
                for (int c = 0; c < 4; c++) {
                    const int i = i0 + r, k = k0 + c;
                    if (i < n && k <= i) A[(size_t)i * ld + k] -= acc[r][c];
                }
        }
        __syncthreads();
    }
    return true;
}

// x <- (L L^T)^-1 x with the lower triangle of Lm (leading dimension ld; the upper triangle is not read as
// factor): forward then back substitution in blocks of 32 columns.  The triangle of each block is solved by the
// first wavefront with its row (forward) or column (backward) of L in registers: each step's solution entry is
// broadcast through the scalar unit and stored by lane 0, and every lane subtracts its multiple of it (a lane whose
// entry is already solved then only updates a value that is not read again, so no per-lane masks are needed).
// The rest of the vector is updated by every thread.
template <typename T>
__device__ __attribute__((always_inline)) inline void chol_solve_wg(const T *Lm, int n, T *x, int ld) {
    const int tid = threadIdx.x, nt = blockDim.x, lane = tid & 63, wv = tid >> 6;
    constexpr int BS = 32;
    for (int j0 = 0; j0 < n; j0 += BS) {
        const int nb = min(BS, n - j0);
        if (wv == 0) {
            const bool in = lane < nb;
            T xi = in ? x[j0 + lane] : (T)0;
            const T dinv = in ? (T)1 / Lm[(size_t)(j0 + lane) * ld + j0 + lane] : (T)0;
            T row[BS];
#pragma unroll
            for (int jj = 0; jj < BS; jj++) row[jj] = (jj < nb && in) ? Lm[(size_t)(j0 + lane) * ld + j0 + jj] : (T)0;
#pragma unroll
            for (int jj = 0; jj < BS; jj++) {
                if (jj < nb) {
                    const T xj = lane_bcast(xi, jj) * lane_bcast(dinv, jj);
                    if (lane == 0) x[j0 + jj] = xj;
                    xi -= row[jj] * xj;
                }
            }
        }
        __syncthreads();
        for (int i = j0 + nb + tid; i < n; i += nt) {
            const T *Li = Lm + (size_t)i * ld + j0;
            T acc = x[i];
            for (int jj = 0; jj < nb; jj++) acc -= Li[jj] * x[j0 + jj];
            x[i] = acc;
        }
        __syncthreads();
    }
    for (int j1 = n; j1 > 0; j1 -= BS) {
        const int j0 = max(0, j1 - BS), nb = j1 - j0;
        if (wv == 0) {
            const bool in = lane < nb;
            T xi = in ? x[j0 + lane] : (T)0;
            const T dinv = in ? (T)1 / Lm[(size_t)(j0 + lane) * ld + j0 + lane] : (T)0;
            T col[BS];
#pragma unroll
            for (int jj = 0; jj < BS; jj++) col[jj] = (jj < nb && in) ? Lm[(size_t)(j0 + jj) * ld + j0 + lane] : (T)0;
#pragma unroll
            for (int jj = BS - 1; jj >= 0; jj--) {
                if (jj < nb) {
                    const T xj = lane_bcast(xi, jj) * lane_bcast(dinv, jj);
                    if (lane == 0) x[j0 + jj] = xj;
                    xi -= col[jj] * xj;
                }
            }
        }
        __syncthreads();
        for (int i = tid; i < j0; i += nt) {
            T acc = x[i];
            for (int jj = 0; jj < nb; jj++) acc -= Lm[(size_t)(j0 + jj) * ld + i] * x[j0 + jj];
            x[i] = acc;
        }
        __syncthreads();
    }
}

// The trial state of a window from its update dx (vio/window.py:522-532, 945-955): R <- R Exp(dphi), p += dp, and
// v, bg, ba additive on 15-dimensional slots.
__device__ inline void trial_state(const DArgs &D, const WinInfo &W, int tid, int nt) {
    for (int s = tid; s < W.P; s += nt) {
        int gs = W.slot0 + s, o = D.slot_off[gs];
        const double *dxs = D.dx + W.st0 + o;
        double E[9];
        so3_exp_nb(dxs, E);
        mm3(D.Rwb + 9 * (size_t)gs, E, D.R2 + 9 * (size_t)gs);
        for (int i = 0; i < 3; i++) D.p2[3 * (size_t)gs + i] = D.pwb[3 * (size_t)gs + i] + dxs[3 + i];
        if (W.vi) {
            bool full = D.slot_dim[gs] == 15;
            for (int i = 0; i < 3; i++) {
                D.v2[3 * (size_t)gs + i] = D.vel[3 * (size_t)gs + i] + (full ? dxs[6 + i] : 0.0);
                D.g2[3 * (size_t)gs + i] = D.bgs[3 * (size_t)gs + i] + (full ? dxs[9 + i] : 0.0);
                D.a2[3 * (size_t)gs + i] = D.bas[3 * (size_t)gs + i] + (full ? dxs[12 + i] : 0.0);
            }
        }
    }
}

// ------------------------------------------------------------------------------------------- k_solve --
// Damped Cholesky solve of every window, dx = -Hd^-1 b, then the trial state.  FP64 (T = double) factors Hd
// itself; MIXED (T = float) factors the Jacobi-scaled Hd in FP32 and refines x in FP64 against Hd until the last
// correction is at most XBA_REFINE_TOL of x or the residual at most 1e-15 of b (at most nref_max solves).  A mixed
// window whose FP32 factor fails or whose refinement does not meet that test within the budget is flagged
// (WS_FALLBACK) and left to the FP64 fallback: k_solve<double> launched with `fallback` = 1 after every mixed solve
// of the iteration (one-workgroup and tiled), which solves exactly the flagged windows, small or tiled, in FP64
// (in LDS when the matrix fits `lds_cap`, else in the window's slot of the FP64 factor buffer).
template <typename T>
__global__ __launch_bounds__(512) void k_solve(DArgs D, int lds_cap, int fallback) {
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ double sx[XBA_MAX_STATE], sr[XBA_MAX_STATE], ss[XBA_MAX_STATE];
    __shared__ T sy[XBA_MAX_STATE];
    __shared__ T sinv[CH_SCRATCH];
    __shared__ double sred[32];
    __shared__ int sflag;
    const int w = blockIdx.x, tid = threadIdx.x, nt = blockDim.x;
    const int lane = tid & 63, wv = tid >> 6, nw = nt >> 6;
    const WinInfo W = D.wins[w];
    double *ws = D.wstate + (size_t)NWS * w;
    if (fallback) {
        if (ws[WS_FALLBACK] == 0.0) return;  // written by this iteration's mixed solve of the window
        if (D.fallback_relin) {
            // the window's system re-assembled with FP64 records: its reduction and prior / IMU terms here
            for (int64_t idx = tid; idx < (int64_t)W.n * W.n; idx += nt)
                reduce_window<double>(D, w, idx, D.partH64, D.partb64, (const double *)nullptr,
                                      (const double *)nullptr);
            window_visual_cost(D, w, sred);
            __syncthreads();
            terms_window(D, w, 0, sx, sr, (int *)ss);
            __syncthreads();
        }
    } else if (W.big >= 0) {
        return;  // factored by the tiled kernels (k_big_*)
    }
    const int n = W.n;
    const double lam = ws[WS_LAM];
    const double *H = D.H + W.h0, *bv = D.b + W.st0;
    // in LDS the rows are padded to a multiple of 4 entries
    const int ldp = (n + 3) & ~3;
    const bool in_lds = (size_t)n * ldp * sizeof(T) <= (size_t)lds_cap;
    const int ld = in_lds ? ldp : n;
    T *Fl = (T *)smem, *Fg = sizeof(T) == 8 ? (T *)(D.fact64 + W.h0) : (T *)(D.fact32 + W.h0);
    T *F = in_lds ? Fl : Fg;
    // scaling (mixed) and the damped matrix in the working precision
    for (int i = tid; i < n; i += nt) {
        double hd = H[(size_t)i * n + i] + (lam * H[(size_t)i * n + i] + 1e-9);
        ss[i] = (sizeof(T) == 8) ? 1.0 : (hd > 0.0 ? 1.0 / sqrt(hd) : 1.0);
    }
    __syncthreads();
    for (int i = tid >> 6; i < n; i += nt >> 6)
        for (int j = tid & 63; j < n; j += 64) {
            double h = H[(size_t)i * n + j];
            if (i == j) h += lam * h + 1e-9;
            F[(size_t)i * ld + j] = (T)(h * ss[i] * ss[j]);
        }
    __syncthreads();
    // the factor and solves inlined once per memory space
    bool ok = in_lds ? chol_wg<T>(Fl, n, &sflag, sinv, ld) : chol_wg<T>(Fg, n, &sflag, sinv, ld);
    int nref = 0;
    bool conv = sizeof(T) == 8;
    if (ok) {
        for (int i = tid; i < n; i += nt) {
            sx[i] = 0.0;
            sr[i] = bv[i];
        }
        __syncthreads();
        const int maxit = (sizeof(T) == 8) ? 1 : D.nref_max;
        for (int it = 0; it < maxit; it++) {
            for (int i = tid; i < n; i += nt) sy[i] = (T)(sr[i] * ss[i]);
            __syncthreads();
            if (in_lds)
                chol_solve_wg<T>(Fl, n, sy, ld);
            else
                chol_solve_wg<T>(Fg, n, sy, ld);
            double cmax = 0.0, xmax = 0.0;
            for (int i = tid; i < n; i += nt) {
                const double c = (double)sy[i] * ss[i], xi = sx[i] + c;
                sx[i] = xi;
                cmax = fmax(cmax, fabs(c));
                xmax = fmax(xmax, fabs(xi));
            }
            __syncthreads();
            nref = it + 1;
            if (sizeof(T) == 8) break;
            // the correction against the solution (block maxima; NaN compares false and never converges)
            cmax = dwave_max(cmax);
            xmax = dwave_max(xmax);
            if (lane == 0) {
                sred[2 * wv] = cmax;
                sred[2 * wv + 1] = xmax;
            }
            __syncthreads();
            double cn = 0.0, xn = 0.0;
            for (int q = 0; q < nw; q++) {
                cn = fmax(cn, sred[2 * q]);
                xn = fmax(xn, sred[2 * q + 1]);
            }
            __syncthreads();
            if (cn <= XBA_REFINE_TOL * xn) {
                conv = true;
                break;
            }
            if (it + 1 == maxit) break;
            // residual r = b - Hd x in FP64, entry i by thread i from column i of the symmetric H (row j's entry i:
            // the threads of a wavefront read consecutive addresses)
            double rmax = 0.0, bmax = 0.0;
            for (int i = tid; i < n; i += nt) {
                // 16 column entries in flight per thread, summed in column order
                double acc = 0.0;
                int j = 0;
                for (; j + 16 <= n; j += 16) {
                    double h[16];
#pragma unroll
                    for (int q = 0; q < 16; q++) h[q] = H[(size_t)(j + q) * n + i];
#pragma unroll
                    for (int q = 0; q < 16; q++) acc += h[q] * sx[j + q];
                }
                for (; j < n; j++) acc += H[(size_t)j * n + i] * sx[j];
                const double ri = bv[i] - acc - (lam * H[(size_t)i * n + i] + 1e-9) * sx[i];
                sr[i] = ri;
                rmax = fmax(rmax, fabs(ri));
                bmax = fmax(bmax, fabs(bv[i]));
            }
            rmax = dwave_max(rmax);
            bmax = dwave_max(bmax);
            if (lane == 0) {
                sred[2 * wv] = rmax;
                sred[2 * wv + 1] = bmax;
            }
            __syncthreads();
            double rn = 0.0, bn = 0.0;
            for (int q = 0; q < nw; q++) {
                rn = fmax(rn, sred[2 * q]);
                bn = fmax(bn, sred[2 * q + 1]);
            }
            __syncthreads();
            if (rn <= 1e-15 * bn) {
                conv = true;
                break;
            }
        }
    }
    // mixed: a failed factor or an unconverged refinement leaves the window to the FP64 fallback
    const bool fb = sizeof(T) == 4 && D.fallback_on && !(ok && conv);
    if (!fb) {
        if (ok) {
            for (int i = tid; i < n; i += nt) D.dx[W.st0 + i] = -sx[i];
            __syncthreads();
            trial_state(D, W, tid, nt);
        } else {
            for (int i = tid; i < n; i += nt) D.dx[W.st0 + i] = NAN;
        }
    }
    if (tid == 0) {
        ws[WS_CHOLOK] = ok ? 1.0 : 0.0;
        if (!fallback) {  // the fallback leaves the mixed solve's refinement record as it was
            ws[WS_NREF] = nref;
            ws[WS_CONV] = conv ? 1.0 : 0.0;
            ws[WS_FALLBACK] = fb ? 1.0 : 0.0;
            if (fb) ws[WS_NFALLBACK] += 1.0;
        }
    }
}

// ------------------------------------------------------------------------------------------- k_accept --
// c_new = trial visual cost (chunk sums in order) + prior/IMU trial cost; accept iff c_new < c_tot.
__global__ __launch_bounds__(256) void k_accept(DArgs D) {
    __shared__ int sacc;
    __shared__ double sred[16];
    const int w = blockIdx.x, tid = threadIdx.x, nt = blockDim.x;
    const WinInfo W = D.wins[w];
    double *ws = D.wstate + (size_t)NWS * w;
    const bool chol_ok = ws[WS_CHOLOK] != 0.0;  // (uniform)
    const double cvis = chol_ok ? block_sum_ranges(D.partc2, W.c0, W.c1, W.c2, W.c3, sred) : 0.0;
    if (tid == 0) {
        int acc = 0;
        if (chol_ok) {
            double cn = cvis;
            cn += ws[WS_COSTTRIAL_REST];
            ws[WS_CNEW] = cn;
            acc = cn < ws[WS_CTOT];
        }
        if (acc) {
            ws[WS_LAM] = fmax(ws[WS_LAM] / 10.0, 1e-12);
            ws[WS_ACC] += 1.0;
            ws[WS_CUR] = ws[WS_CNEW];
        } else {
            ws[WS_LAM] *= 10.0;
        }
        ws[WS_ACCEPT] = acc ? 1.0 : 0.0;
        ws[WS_ITERS] += 1.0;
        sacc = acc;
    }
    __syncthreads();
    if (!sacc) return;
    for (int s = tid; s < W.P; s += nt) {
        size_t gs = W.slot0 + s;
        for (int i = 0; i < 9; i++) D.Rwb[9 * gs + i] = D.R2[9 * gs + i];
        for (int i = 0; i < 3; i++) D.pwb[3 * gs + i] = D.p2[3 * gs + i];
        if (W.vi)
            for (int i = 0; i < 3; i++) {
                D.vel[3 * gs + i] = D.v2[3 * gs + i];
                D.bgs[3 * gs + i] = D.g2[3 * gs + i];
                D.bas[3 * gs + i] = D.a2[3 * gs + i];
            }
    }
}

template __global__ void k_reduce<float>(DArgs, const float *, const float *, const float *, const float *);
template __global__ void k_reduce<double>(DArgs, const double *, const double *, const double *, const double *);
template __global__ void k_solve<float>(DArgs, int, int);
template __global__ void k_solve<double>(DArgs, int, int);

}  // namespace xba
