// Landmark kernels of the window solve: linearise + Schur (k_lin), back-substitution + trial cost (k_trial),
// landmark commit, and the adjoint's per-observation gradient.  One wavefront per work item (a landmark chunk of
// one window); inside a chunk the landmarks are processed one after another with one lane per observation, and
// the reduced-system contributions accumulate in an LDS tile that covers the chunk's slot range.  Every sum runs
// in a fixed order, so results are bitwise reproducible run to run.
//
// Reference arithmetic (paths relative to the reference's src/xclean/):
//   residual / Jacobians  vio/window.py:125-146, 164-263
//   Huber                 vio/window.py:284-288
//   per-landmark blocks, inverse-depth prior, damping, det guard, closed-form inverse, Schur   vio/window.py:324-462
//   back-substitution     vio/window.py:496-519;  landmark update rule   vio/window.py:533-540
//   trial cost            vio/window.py:466-493
#include "xba_internal.h"

namespace xba {

// ---------------------------------------------------------------------------------------------- LDS layout --
enum {
    F_W = 0,   // Huber weight (0: observation not evaluable)
    F_E = 1,   // whitened residual (2)
    F_JL = 3,  // whitened d r / d(a, b, rho)   (2 x 3)
    F_JT = 9,  // whitened d r / d target pose  (2 x 6)
    F_JH = 21, // whitened d r / d host pose    (2 x 6)
    F_C = 33,  // robust cost of the observation
    F_ND = 34, // residual dimension
    NF = 35
};
#define SC_STRIDE 24  // per slot: Rwc 9, pwc 3, Rwb 9, pad 3
#define SC64_STRIDE 12  // per slot: Rwc 9, pwc 3 (FP64)
#define LM_REC 32     // Hll 9, bl 3, Hi 9, HiB 3, ok, nobs, (pad)

struct LdsLayout {
    int sc, sc64, wc, wd, rec, rc, rt, flags, tl, hlp, t1, lm, hacc, bacc, hcmp, bcmp, total;
};

__host__ __device__ inline int align16(int x) { return (x + 15) & ~15; }

__host__ __device__ inline LdsLayout lds_layout(int R, int ts) {
    LdsLayout L;
    int o = 0;
    L.sc = o;    o = align16(o + 2 * R * SC_STRIDE * ts);   // current and trial slot caches (working precision)
    L.sc64 = o;  o = align16(o + 2 * R * SC64_STRIDE * 8);  // the same camera poses in FP64 (residuals)
    L.wc = o;    o = align16(o + 32 * ts);                   // Rcb 9, Px 9, fx fy cx cy B hk
    L.wd = o;    o = align16(o + 8 * 8);                     // fx fy cx cy B hk in FP64
    L.rec = o;   o = align16(o + NF * XBA_WAVE * ts);
    L.rc = o;    o = align16(o + (XBA_WAVE + 8) * 8);        // per-observation robust cost (FP64) + scratch
    L.rt = o;    o = align16(o + XBA_WAVE * 4);
    L.flags = o; o = align16(o + XBA_WAVE * 4);
    L.tl = o;    o = align16(o + XBA_WAVE * 4);
    L.hlp = o;   o = align16(o + R * 18 * ts);
    L.t1 = o;    o = align16(o + R * 18 * ts);
    L.lm = o;    o = align16(o + LM_REC * ts);
    L.hacc = o;  o = align16(o + (R * (R + 1) / 2) * 36 * ts);
    L.bacc = o;  o = align16(o + 6 * R * ts);
    // FP32 tiles are accumulated with Kahan compensation (compensation terms here)
    const int kc = ts == 4 ? 1 : 0;
    L.hcmp = o;  o = align16(o + kc * (R * (R + 1) / 2) * 36 * ts);
    L.bcmp = o;  o = align16(o + kc * 6 * R * ts);
    L.total = o;
    return L;
}

// --------------------------------------------------------------------------------------- per-slot caches --
// Camera poses of the R tile slots of a work item of window w (slots lo + t, or map[t] when map is given):
// Rwc = Rwb Rbc, pwc = pwb + Rwb pbc (vio/window.py:110-122), computed in FP64 and stored in the working precision
// (sc) and, unless sc64 is null, in FP64.
template <typename T>
__device__ void load_slots(T *sc, double *sc64, const double *Rwb, const double *pwb, const double *wp, int slot0,
                           int lo, int R, int lane, const int32_t *map = nullptr) {
    const double *Rbc = wp + 8, *pbc = wp + 17;
    for (int s = lane; s < R; s += XBA_WAVE) {
        const int ws = slot0 + (map ? map[s] : lo + s);
        const double *Rb = Rwb + 9 * (size_t)ws;
        const double *pb = pwb + 3 * (size_t)ws;
        T *c = sc + s * SC_STRIDE;
        double *c64 = sc64 ? sc64 + s * SC64_STRIDE : nullptr;
        for (int i = 0; i < 3; i++) {
            for (int j = 0; j < 3; j++) {
                double acc = 0.0;
                for (int k = 0; k < 3; k++) acc += Rb[3 * i + k] * Rbc[3 * k + j];
                c[3 * i + j] = (T)acc;
                if (sc64) c64[3 * i + j] = acc;
            }
            double acc = pb[i];
            for (int k = 0; k < 3; k++) acc += Rb[3 * i + k] * pbc[k];
            c[9 + i] = (T)acc;
            if (sc64) c64[9 + i] = acc;
        }
        for (int i = 0; i < 9; i++) c[12 + i] = (T)Rb[i];
    }
}

template <typename T>
__device__ void load_wconst(T *wc, const double *wp, int lane) {
    if (lane < 9) {
        int i = lane / 3, j = lane % 3;
        wc[lane] = (T)wp[8 + 3 * j + i];  // Rcb = Rbc^T
    } else if (lane < 18) {
        const double *p = wp + 17;
        int q = lane - 9;
        double v = 0.0;
        // Px = [pbc]x, row-major
        switch (q) {
            case 1: v = -p[2]; break;
            case 2: v = p[1]; break;
            case 3: v = p[2]; break;
            case 5: v = -p[0]; break;
            case 6: v = -p[1]; break;
            case 7: v = p[0]; break;
            default: v = 0.0;
        }
        wc[lane] = (T)v;
    } else if (lane < 24) {
        wc[lane] = (T)wp[lane - 18];  // fx fy cx cy B huber_k
    }
}

__device__ inline void load_wd(double *wd, const double *wp, int lane) {
    if (lane < 6) wd[lane] = wp[lane];  // fx fy cx cy B huber_k
}

// ------------------------------------------------------------------------------------ one observation --
// Residual of one observation (vio/window.py:125-146), formed in FP64 from the FP64 camera poses and state, and,
// when `jac`, its Jacobians in the working precision T (vio/window.py:164-263): Jl wrt (a, b, rho), Jt / Jh wrt the
// target / host body pose blocks [dphi, dp] (zero when t == h).
template <typename T>
__device__ inline bool obs_eval(const double *sh64, const double *st64, const T *sh, const T *st, const T *wc,
                                const double *wd, double a, double b, double rho, int cam, double z0, double z1,
                                bool same, bool jac, double e[2], T Jl[2][3], T Jt[2][6], T Jh[2][6]) {
    const double *Rh = sh64, *ph = sh64 + 9, *Rt = st64, *pt = st64 + 9;
    double d0 = ph[0] - pt[0], d1 = ph[1] - pt[1], d2 = ph[2] - pt[2];
    double Y0 = Rh[0] * a + Rh[1] * b + Rh[2] + rho * d0;
    double Y1 = Rh[3] * a + Rh[4] * b + Rh[5] + rho * d1;
    double Y2 = Rh[6] * a + Rh[7] * b + Rh[8] + rho * d2;
    double q0 = Rt[0] * Y0 + Rt[3] * Y1 + Rt[6] * Y2;
    double q1 = Rt[1] * Y0 + Rt[4] * Y1 + Rt[7] * Y2;
    double q2 = Rt[2] * Y0 + Rt[5] * Y1 + Rt[8] * Y2;
    if (!(q2 > 1e-9)) return false;
    const double fxd = wd[0], fyd = wd[1], cxd = wd[2], cyd = wd[3], Bd = wd[4];
    // both entries written once, after the branch (stores in the branches get merged into one store through a
    // selected address, which keeps the caller's array in scratch memory)
    double e0, e1;
    if (cam == 0) {
        e0 = fxd * q0 / q2 + cxd - z0;
        e1 = fyd * q1 / q2 + cyd - z1;
    } else {
        e0 = fxd * (q0 - rho * Bd) / q2 + cxd - z0;
        e1 = 0.0;
    }
    e[0] = e0;
    e[1] = e1;
    if (!jac) return true;
    const T fx = wc[18], fy = wc[19], B = wc[22];
    const T *RhT = sh, *RtT = st;
    const T qa = (T)q0, qb = (T)q1, qc = (T)q2, rT = (T)rho, aT = (T)a, bT = (T)b;
    T iq = (T)1 / qc, Dp[2][3];
    if (cam == 0) {
        Dp[0][0] = fx * iq; Dp[0][1] = 0; Dp[0][2] = -fx * qa * iq * iq;
        Dp[1][0] = 0; Dp[1][1] = fy * iq; Dp[1][2] = -fy * qb * iq * iq;
    } else {
        Dp[0][0] = fx * iq; Dp[0][1] = 0; Dp[0][2] = -fx * (qa - rT * B) * iq * iq;
        Dp[1][0] = 0; Dp[1][1] = 0; Dp[1][2] = 0;
    }
    T RcR[9];  // R_ct R_wc,h
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
            RcR[3 * i + j] = RtT[i] * RhT[j] + RtT[3 + i] * RhT[3 + j] + RtT[6 + i] * RhT[6 + j];
    const T dT0 = (T)d0, dT1 = (T)d1, dT2 = (T)d2;
    T r0 = RtT[0] * dT0 + RtT[3] * dT1 + RtT[6] * dT2;
    T r1 = RtT[1] * dT0 + RtT[4] * dT1 + RtT[7] * dT2;
    T r2 = RtT[2] * dT0 + RtT[5] * dT1 + RtT[8] * dT2;
    if (cam == 1) r0 -= B;
    for (int r = 0; r < 2; r++) {
        Jl[r][0] = Dp[r][0] * RcR[0] + Dp[r][1] * RcR[3] + Dp[r][2] * RcR[6];
        Jl[r][1] = Dp[r][0] * RcR[1] + Dp[r][1] * RcR[4] + Dp[r][2] * RcR[7];
        Jl[r][2] = Dp[r][0] * r0 + Dp[r][1] * r1 + Dp[r][2] * r2;
    }
    if (same) {
        // (stores in the order of the general branch's last ones, Jt then Jh: the compiler merges the two branches'
        // final stores, and stores to different entries would leave the caller's arrays in scratch memory)
        for (int r = 0; r < 2; r++)
            for (int j = 0; j < 6; j++) {
                Jt[r][j] = (T)0;
                Jh[r][j] = (T)0;
            }
        return true;
    }
    const T *Rcb = wc, *Px = wc + 9, *Rwbt = st + 12, *Rwbh = sh + 12;
    // Ah = RcR [m]x with m = (a, b, 1)
    T Mx[9] = {0, -1, bT, 1, 0, -aT, -bT, aT, 0};
    T Qx[9] = {0, -qc, qb, qc, 0, -qa, -qb, qa, 0};
    T Ah[9], Wt[9], Wh[9], T1[9], T2[9];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            Ah[3 * i + j] = RcR[3 * i] * Mx[j] + RcR[3 * i + 1] * Mx[3 + j] + RcR[3 * i + 2] * Mx[6 + j];
            Wt[3 * i + j] = Rwbt[3 * i] * Px[j] + Rwbt[3 * i + 1] * Px[3 + j] + Rwbt[3 * i + 2] * Px[6 + j];
            Wh[3 * i + j] = Rwbh[3 * i] * Px[j] + Rwbh[3 * i + 1] * Px[3 + j] + Rwbh[3 * i + 2] * Px[6 + j];
        }
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            // Rct = Rt^T: Rct[i][k] = Rt[3k + i]
            T qr = Qx[3 * i] * Rcb[j] + Qx[3 * i + 1] * Rcb[3 + j] + Qx[3 * i + 2] * Rcb[6 + j];
            T rw = RtT[i] * Wt[j] + RtT[3 + i] * Wt[3 + j] + RtT[6 + i] * Wt[6 + j];
            T1[3 * i + j] = qr + rT * rw;
            T ar = Ah[3 * i] * Rcb[j] + Ah[3 * i + 1] * Rcb[3 + j] + Ah[3 * i + 2] * Rcb[6 + j];
            T rh = RtT[i] * Wh[j] + RtT[3 + i] * Wh[3 + j] + RtT[6 + i] * Wh[6 + j];
            T2[3 * i + j] = -ar - rT * rh;
        }
    for (int r = 0; r < 2; r++)
        for (int j = 0; j < 3; j++) {
            Jt[r][j] = Dp[r][0] * T1[j] + Dp[r][1] * T1[3 + j] + Dp[r][2] * T1[6 + j];
            Jh[r][j] = Dp[r][0] * T2[j] + Dp[r][1] * T2[3 + j] + Dp[r][2] * T2[6 + j];
            T v = Dp[r][0] * RtT[3 * j] + Dp[r][1] * RtT[3 * j + 1] + Dp[r][2] * RtT[3 * j + 2];
            Jt[r][3 + j] = -rT * v;
            Jh[r][3 + j] = rT * v;
        }
    return true;
}

// The same observation's residual, its landmark Jacobian Jl and, instead of the pose Jacobians, their product with
// the pose steps [dphi, dp] of the target and host slots: jd = Jt [dphi_t, dp_t] + Jh [dphi_h, dp_h] (t != h), all in
// FP64.  The camera-frame point moves by
//     dq = [q]x Rcb dphi_t + R_ct (rho (R_wb,t [pbc]x dphi_t - dp_t - R_wb,h [pbc]x dphi_h + dp_h) - R_wc,h [m]x Rcb dphi_h),
// obs_eval's T1 / T2 / translation columns applied to the steps, so jd = Dp dq.  Each slot's step enters through
// v = [Rcb dphi, R_wb [pbc]x dphi - dp] (dir_slot), vt / vh of the target / host.  The world-frame vectors are
// formed first and rotated by R_ct = R_wc,t^T last, so each pose is read once; Jl = Dp R_ct [R_wc,h e0, R_wc,h e1,
// p_wc,h - p_wc,t] (minus the baseline in x for the right camera) is taken as (R_wc,t Dp^T)^T applied to those.
__device__ inline bool obs_dir(const double *sh, const double *st, const double *wd, double a, double b, double rho,
                               int cam, double z0, double z1, const double *vt, const double *vh, double e[2],
                               double Jl[2][3], double jd[2]) {
    const double *Rh = sh, *ph = sh + 9, *Rt = st, *pt = st + 9;
    const double d0 = ph[0] - pt[0], d1 = ph[1] - pt[1], d2 = ph[2] - pt[2];
    const double m0 = b * vh[2] - vh[1], m1 = vh[0] - a * vh[2], m2 = a * vh[1] - b * vh[0];
    const double dd[3] = {d0, d1, d2};
    double Y[3], u[3];
    for (int i = 0; i < 3; i++) {
        Y[i] = Rh[3 * i] * a + Rh[3 * i + 1] * b + Rh[3 * i + 2] + rho * dd[i];
        u[i] = rho * (vt[3 + i] - vh[3 + i]) - (Rh[3 * i] * m0 + Rh[3 * i + 1] * m1 + Rh[3 * i + 2] * m2);
    }
    const double q0 = Rt[0] * Y[0] + Rt[3] * Y[1] + Rt[6] * Y[2];
    const double q1 = Rt[1] * Y[0] + Rt[4] * Y[1] + Rt[7] * Y[2];
    const double q2 = Rt[2] * Y[0] + Rt[5] * Y[1] + Rt[8] * Y[2];
    if (!(q2 > 1e-9)) return false;
    const double fx = wd[0], fy = wd[1], cx = wd[2], cy = wd[3], B = wd[4];
    if (cam == 0) {
        e[0] = fx * q0 / q2 + cx - z0;
        e[1] = fy * q1 / q2 + cy - z1;
    } else {
        e[0] = fx * (q0 - rho * B) / q2 + cx - z0;
        e[1] = 0.0;
    }
    const double iq = 1.0 / q2;
    double Dp[2][3];
    if (cam == 0) {
        Dp[0][0] = fx * iq; Dp[0][1] = 0; Dp[0][2] = -fx * q0 * iq * iq;
        Dp[1][0] = 0; Dp[1][1] = fy * iq; Dp[1][2] = -fy * q1 * iq * iq;
    } else {
        Dp[0][0] = fx * iq; Dp[0][1] = 0; Dp[0][2] = -fx * (q0 - rho * B) * iq * iq;
        Dp[1][0] = 0; Dp[1][1] = 0; Dp[1][2] = 0;
    }
    // [q]x Rcb dphi_t, in the target camera
    const double c0 = q1 * vt[2] - q2 * vt[1], c1 = q2 * vt[0] - q0 * vt[2], c2 = q0 * vt[1] - q1 * vt[0];
    for (int r = 0; r < 2; r++) {
        double D[3];  // R_wc,t Dp[r]^T
        for (int k = 0; k < 3; k++) D[k] = Rt[3 * k] * Dp[r][0] + Rt[3 * k + 1] * Dp[r][1] + Rt[3 * k + 2] * Dp[r][2];
        Jl[r][0] = D[0] * Rh[0] + D[1] * Rh[3] + D[2] * Rh[6];
        Jl[r][1] = D[0] * Rh[1] + D[1] * Rh[4] + D[2] * Rh[7];
        Jl[r][2] = D[0] * d0 + D[1] * d1 + D[2] * d2 - (cam == 1 ? Dp[r][0] * B : 0.0);
        jd[r] = Dp[r][0] * c0 + Dp[r][1] * c1 + Dp[r][2] * c2 + D[0] * u[0] + D[1] * u[1] + D[2] * u[2];
    }
    return true;
}

// One slot's step [dphi, dp] (in d, overwritten) as obs_dir reads it: [Rcb dphi, R_wb [pbc]x dphi - dp], with the
// slot record c (R_wb at c + 12) and the window constants wc (Rcb at wc, [pbc]x row-major at wc + 9).
template <typename T>
__device__ inline void dir_slot(T *d, const T *c, const T *wc) {
    const T *Rcb = wc, *P = wc + 9, *Rwb = c + 12;
    const T f0 = d[0], f1 = d[1], f2 = d[2], t0 = d[3], t1 = d[4], t2 = d[5];
    const T x0 = P[1] * f1 + P[2] * f2, x1 = P[3] * f0 + P[5] * f2, x2 = P[6] * f0 + P[7] * f1;
    T v[6];
    for (int i = 0; i < 3; i++) v[i] = Rcb[3 * i] * f0 + Rcb[3 * i + 1] * f1 + Rcb[3 * i + 2] * f2;
    v[3] = Rwb[0] * x0 + Rwb[1] * x1 + Rwb[2] * x2 - t0;
    v[4] = Rwb[3] * x0 + Rwb[4] * x1 + Rwb[5] * x2 - t1;
    v[5] = Rwb[6] * x0 + Rwb[7] * x1 + Rwb[8] * x2 - t2;
    for (int i = 0; i < 6; i++) d[i] = v[i];
}

__device__ inline void huber(double en, double k, double *w, double *c) {
    double ww, cc;  // each output stored once, after the branch (see obs_eval)
    if (en <= k) {
        ww = 1.0;
        cc = 0.5 * en * en;
    } else {
        ww = k / en;
        cc = k * (en - 0.5 * k);
    }
    *w = ww;
    *c = cc;
}

// Tile accumulation: FP64 adds directly; FP32 adds with Kahan compensation (c holds the negated lost low part).
template <typename T>
__device__ inline void tile_add(T *acc, T *cmp, int i, T v) {
    if (sizeof(T) == 4) {
        T y = v - cmp[i];
        T t = acc[i] + y;
        cmp[i] = (t - acc[i]) - y;
        acc[i] = t;
    } else {
        acc[i] += v;
    }
}

// ----------------------------------------------------------------- one landmark: observations into LDS --
// Lanes over the landmark's observations: whitened residual (FP64), Huber weight and cost (FP64), whitened
// Jacobians (computed in T), stored per lane in the wave's record in the record type RT.
// COMPACT (the multi-wave kernels' record): no residual-dimension field, and of the host block only its rotation
// columns (its translation columns are the negated target ones, see obs_eval), at F_JH + 3 r + j.
template <typename T, typename RT = T, bool COMPACT = false>
__device__ inline void lin_observations(const KArgs<T> &A, RT *rec, double *rc, int *rt, const double *sc64,
                                        const T *sc, const T *wc, const double *wd, int o0, int no, int lo, int h,
                                        double a, double b, double rho, int lane, bool jac) {
    if (lane < no) {
        int k = o0 + lane;
        int info = A.ob_info[k];
        int t = (info & 0xfff) - lo, cam = (info >> 12) & 1, valid = (info >> 13) & 1;
        double e[2] = {0.0, 0.0};
        T Jl[2][3], Jt[2][6], Jh[2][6];
        bool act = false;
        int nd = cam == 0 ? 2 : 1;
        double w = 0, c = 0, en2 = 0;
        if (valid) {
            act = obs_eval<T>(sc64 + h * SC64_STRIDE, sc64 + t * SC64_STRIDE, sc + h * SC_STRIDE, sc + t * SC_STRIDE,
                              wc, wd, a, b, rho, cam, A.z[2 * (size_t)k], A.z[2 * (size_t)k + 1], t == h, jac, e, Jl,
                              Jt, Jh);
        }
        if (act) {
            const double isg = (double)A.isig[k];
            for (int i = 0; i < nd; i++) {
                e[i] *= isg;
                en2 += e[i] * e[i];
            }
            huber(sqrt(en2), wd[5], &w, &c);
            if (A.ob_w) {
                w *= A.ob_w[k];
                c *= A.ob_w[k];
            }
            if (jac) {
                const T isgT = A.isig[k];
                for (int i = 0; i < nd; i++) {
                    for (int j = 0; j < 3; j++) Jl[i][j] *= isgT;
                    for (int j = 0; j < 6; j++) {
                        Jt[i][j] *= isgT;
                        Jh[i][j] *= isgT;
                    }
                }
            }
        }
        if (A.chi2_out) A.chi2_out[k] = act ? en2 : (double)INFINITY;
        rec[F_W * XBA_WAVE + lane] = act ? (RT)w : (RT)0;
        rc[lane] = act ? c : 0.0;
        if (!COMPACT) rec[F_ND * XBA_WAVE + lane] = (RT)nd;
        rt[lane] = t;
        if (act) {
            rec[(F_E + 0) * XBA_WAVE + lane] = (RT)e[0];
            rec[(F_E + 1) * XBA_WAVE + lane] = nd == 2 ? (RT)e[1] : (RT)0;
            if (jac)
                for (int r = 0; r < 2; r++) {
                    bool on = r < nd;
                    for (int j = 0; j < 3; j++) rec[(F_JL + 3 * r + j) * XBA_WAVE + lane] = on ? (RT)Jl[r][j] : (RT)0;
                    for (int j = 0; j < 6; j++) {
                        rec[(F_JT + 6 * r + j) * XBA_WAVE + lane] = on ? (RT)Jt[r][j] : (RT)0;
                        if (!COMPACT)
                            rec[(F_JH + 6 * r + j) * XBA_WAVE + lane] = on ? (RT)Jh[r][j] : (RT)0;
                        else if (j < 3)
                            rec[(F_JH + 3 * r + j) * XBA_WAVE + lane] = on ? (RT)Jh[r][j] : (RT)0;
                    }
                }
        }
    }
}

#define REC(f, q) rec[(f) * XBA_WAVE + (q)]

// Row stride inside the record for a field group (e: 1, Jl: 3, Jt/Jh: 6).
__device__ inline int fstride(int f) { return f >= F_JT ? 6 : (f >= F_JL ? 3 : 1); }

// Sum over the landmark's active observations q, in order, of w_q * sum_r X_q[r] Y_q[r] for the record fields
// fx, fy, restricted by `mode`: 0 all active, 1 active with t != h, 2 active with t == s != h.

template <typename T>
__device__ inline T obs_sum2(const T *rec, const int *rt, int no, int h, int s, int fx, int fy, int mode) {
    T acc = 0;
    int sx = fstride(fx), sy = fstride(fy);
    for (int q = 0; q < no; q++) {
        T w = REC(F_W, q);
        if (w == (T)0) continue;
        int t = rt[q];
        if (mode == 1 && t == h) continue;
        if (mode == 2 && (t != s || t == h)) continue;
        T v = REC(fx, q) * REC(fy, q);
        if ((int)REC(F_ND, q) == 2) v += REC(fx + sx, q) * REC(fy + sy, q);
        acc += w * v;
    }
    return acc;
}

// --------------------------------------------------------------------------------------- k_lin --
// build_system for every chunk: per landmark Hll, bl, Hlp per touched slot, the inverse-depth prior, damping and
// the det guard, the closed-form inverse, then the direct pose-pose blocks and the Schur complement into the
// chunk's LDS tile; the tile is flushed to the chunk's partial at the end.
template <typename T>
__global__ __launch_bounds__(XBA_WAVE) void k_lin(KArgs<T> A, int Rmax) {
    extern __shared__ __align__(16) unsigned char smem[];
    const LdsLayout LL = lds_layout(Rmax, (int)sizeof(T));
    T *sc = (T *)(smem + LL.sc), *wc = (T *)(smem + LL.wc), *rec = (T *)(smem + LL.rec);
    double *sc64 = (double *)(smem + LL.sc64), *wd = (double *)(smem + LL.wd), *rc = (double *)(smem + LL.rc);
    int *rt = (int *)(smem + LL.rt), *flags = (int *)(smem + LL.flags), *tl = (int *)(smem + LL.tl);
    T *hlp = (T *)(smem + LL.hlp), *t1 = (T *)(smem + LL.t1), *lm = (T *)(smem + LL.lm);
    T *hacc = (T *)(smem + LL.hacc), *bacc = (T *)(smem + LL.bacc);
    T *hcmp = (T *)(smem + LL.hcmp), *bcmp = (T *)(smem + LL.bcmp);
    const int lane = threadIdx.x;
    const bool kah = sizeof(T) == 4;

    for (int c = blockIdx.x; c < A.nchunks; c += gridDim.x) {
        const Chunk ch = A.chunks[c];
        const WinInfo W = A.wins[ch.win];
        const double *wp = A.win_par + (size_t)XBA_WPAR * ch.win;
        const int R = ch.R, nblk = R * (R + 1) / 2;
        const double lamd = A.lam_override >= 0.0 ? A.lam_override : A.wstate[(size_t)NWS * ch.win + WS_LAM];
        load_slots<T>(sc, sc64, A.Rwb, A.pwb, wp, W.slot0, ch.lo, R, lane);
        load_wconst<T>(wc, wp, lane);
        load_wd(wd, wp, lane);
        for (int i = lane; i < nblk * 36; i += XBA_WAVE) {
            hacc[i] = 0;
            if (kah) hcmp[i] = 0;
        }
        for (int i = lane; i < 6 * R; i += XBA_WAVE) {
            bacc[i] = 0;
            if (kah) bcmp[i] = 0;
        }
        double cost = 0.0;  // lane 0
        __syncthreads();
        for (int l = ch.l0; l < ch.l1; l++) {
            const int li = A.lm_info[l];
            if (!((li >> 12) & 1)) continue;  // lm_mask
            const int h = (li & 0xfff) - ch.lo;
            const int o0 = A.ob_start[l], no = A.ob_start[l + 1] - o0;
            const double a = A.lmp[3 * (size_t)l], b = A.lmp[3 * (size_t)l + 1], rho = A.lmp[3 * (size_t)l + 2];
            if (lane < XBA_WAVE) flags[lane] = 0;
            lin_observations<T>(A, rec, rc, rt, sc64, sc, wc, wd, o0, no, ch.lo, h, a, b, rho, lane, true);
            __syncthreads();
            // touched slots: host + every target of an active observation with t != h
            if (lane < no && REC(F_W, lane) != (T)0 && rt[lane] != h) {
                flags[rt[lane]] = 1;
                flags[h] = 1;
            }
            __syncthreads();
            const int f = lane < R ? flags[lane] : 0;
            const uint64_t bal = __ballot(f);
            const int k = __popcll(bal);
            if (f) tl[__popcll(bal & ((1ull << lane) - 1ull))] = lane;
            // landmark blocks: Hll (lanes 0-8), bl (9-11), observation cost sum and count (lane 12)
            if (lane < 9) {
                int i = lane / 3, j = lane % 3;
                lm[lane] = obs_sum2<T>(rec, rt, no, h, 0, F_JL + i, F_JL + j, 0);
            } else if (lane < 12) {
                int i = lane - 9;
                lm[lane] = obs_sum2<T>(rec, rt, no, h, 0, F_JL + i, F_E, 0);
            } else if (lane == 12) {
                double cs = 0.0;
                int nact = 0;
                for (int q = 0; q < no; q++)
                    if (REC(F_W, q) != (T)0) {
                        cs += rc[q];
                        nact++;
                    }
                lm[25] = (T)nact;
                rc[XBA_WAVE] = cs;
            }
            __syncthreads();
            // Hlp per touched slot: Hlp[s][i][j] = sum_q w (Jl_q^T J_{q,s})[i][j]
            for (int o = lane; o < k * 18; o += XBA_WAVE) {
                int si = o / 18, ij = o % 18, i = ij / 6, j = ij % 6, s = tl[si];
                T v = (s == h) ? obs_sum2<T>(rec, rt, no, h, s, F_JL + i, F_JH + j, 1)
                               : obs_sum2<T>(rec, rt, no, h, s, F_JL + i, F_JT + j, 2);
                hlp[si * 18 + ij] = v;
            }
            // lane 0: inverse-depth prior, damping, det guard, closed-form inverse (vio/window.py:407-434), in FP64
            if (lane == 0) {
                double Hll[9], bl[3];
                for (int i = 0; i < 9; i++) Hll[i] = (double)lm[i];
                for (int i = 0; i < 3; i++) bl[i] = (double)lm[9 + i];
                double lcost = rc[XBA_WAVE];
                const T ris = A.rho_isig[l];
                if (ris > (T)0) {
                    double risd = (double)ris, r0 = (rho - A.rho_meas[l]) * risd, w, cc;
                    huber(fabs(r0), wd[5], &w, &cc);
                    lcost += cc;
                    Hll[8] += w * risd * risd;
                    bl[2] += w * risd * r0;
                }
                cost += lcost;
                for (int i = 0; i < 3; i++) Hll[4 * i] += lamd * Hll[4 * i] + 1e-12;
                double det = (Hll[0] * (Hll[4] * Hll[8] - Hll[5] * Hll[7]) - Hll[1] * (Hll[3] * Hll[8] - Hll[5] * Hll[6]) +
                              Hll[2] * (Hll[3] * Hll[7] - Hll[4] * Hll[6]));
                double tr = Hll[0] + Hll[4] + Hll[8];
                bool ok = ((int)lm[25] > 0) && (det > 1e-18 * tr * tr * tr) && isfinite(det);
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
                for (int i = 0; i < 9; i++) lm[12 + i] = (T)Hi[i];
                for (int i = 0; i < 3; i++)
                    lm[21 + i] = (T)(Hi[3 * i] * bl[0] + Hi[3 * i + 1] * bl[1] + Hi[3 * i + 2] * bl[2]);
                lm[24] = ok ? (T)1 : (T)0;
                A.lmok[l] = ok ? 1 : 0;
                if (A.ok_out) A.ok_out[l] = ok ? 1 : 0;
                T *ho = A.lmHi + 6 * (size_t)l;
                ho[0] = (T)Hi[0]; ho[1] = (T)Hi[1]; ho[2] = (T)Hi[2]; ho[3] = (T)Hi[4]; ho[4] = (T)Hi[5];
                ho[5] = (T)Hi[8];
                T *bo = A.lmbl + 3 * (size_t)l;
                bo[0] = (T)bl[0]; bo[1] = (T)bl[1]; bo[2] = (T)bl[2];
                if (A.Hinv_out)
                    for (int i = 0; i < 9; i++) A.Hinv_out[9 * (size_t)l + i] = ok ? Hi[i] : 0.0;
            }
            __syncthreads();
            const bool ok = lm[24] != (T)0;
            // T1 = Hi Hlp per touched slot
            if (ok)
                for (int o = lane; o < k * 18; o += XBA_WAVE) {
                    int si = o / 18, i = (o % 18) / 6, j = o % 6;
                    const T *A1 = hlp + si * 18;
                    t1[o] = lm[12 + 3 * i] * A1[j] + lm[12 + 3 * i + 1] * A1[6 + j] + lm[12 + 3 * i + 2] * A1[12 + j];
                }
            __syncthreads();
            // reduced system: lower-triangular touched block pairs (s2 >= s1)
            const int npair = k * (k + 1) / 2;
            for (int o = lane; o < npair * 36; o += XBA_WAVE) {
                int p = o / 36, ij = o % 36, i = ij / 6, j = ij % 6;
                int si = (int)((sqrtf(8.0f * p + 1.0f) - 1.0f) * 0.5f);
                while ((si + 1) * (si + 2) / 2 <= p) si++;
                while (si * (si + 1) / 2 > p) si--;
                int sj = p - si * (si + 1) / 2;
                int s2 = tl[si], s1 = tl[sj];
                T v = 0;
                if (s2 == s1) {
                    v = (s2 == h) ? obs_sum2<T>(rec, rt, no, h, s2, F_JH + i, F_JH + j, 1)
                                  : obs_sum2<T>(rec, rt, no, h, s2, F_JT + i, F_JT + j, 2);
                } else if (s1 == h) {
                    v = obs_sum2<T>(rec, rt, no, h, s2, F_JT + i, F_JH + j, 2);
                } else if (s2 == h) {
                    v = obs_sum2<T>(rec, rt, no, h, s1, F_JH + i, F_JT + j, 2);
                }
                if (ok) {
                    const T *A2 = hlp + si * 18, *B1 = t1 + sj * 18;
                    v -= A2[i] * B1[j] + A2[6 + i] * B1[6 + j] + A2[12 + i] * B1[12 + j];
                }
                tile_add<T>(hacc, hcmp, (s2 * (s2 + 1) / 2 + s1) * 36 + ij, v);
            }
            for (int o = lane; o < k * 6; o += XBA_WAVE) {
                int si = o / 6, i = o % 6, s = tl[si];
                T v = (s == h) ? obs_sum2<T>(rec, rt, no, h, s, F_JH + i, F_E, 1)
                               : obs_sum2<T>(rec, rt, no, h, s, F_JT + i, F_E, 2);
                if (ok) {
                    const T *A1 = hlp + si * 18;
                    v -= A1[i] * lm[21] + A1[6 + i] * lm[22] + A1[12 + i] * lm[23];
                }
                tile_add<T>(bacc, bcmp, 6 * s + i, v);
            }
            __syncthreads();
        }
        // flush the tile
        T *pH = A.partH + ch.partH, *pb = A.partb + ch.partb;
        for (int i = lane; i < nblk * 36; i += XBA_WAVE) {
            pH[i] = hacc[i];
            if (kah) A.partH_lo[ch.partH + i] = -hcmp[i];
        }
        for (int i = lane; i < 6 * R; i += XBA_WAVE) {
            pb[i] = bacc[i];
            if (kah) A.partb_lo[ch.partb + i] = -bcmp[i];
        }
        if (lane == 0) A.partc[c] = cost;
        __syncthreads();
    }
}

// ------------------------------------------------------------------------------------- k_trial --
// back_substitute + apply_update's landmark rule + visual_cost at the trial state, per chunk.  Windows whose
// Cholesky failed skip the chunk (their trial cost is unused).
template <typename T>
__global__ __launch_bounds__(XBA_WAVE) void k_trial(KArgs<T> A, int Rmax) {
    extern __shared__ __align__(16) unsigned char smem[];
    const LdsLayout LL = lds_layout(Rmax, (int)sizeof(T));
    T *sc = (T *)(smem + LL.sc), *wc = (T *)(smem + LL.wc), *rec = (T *)(smem + LL.rec);
    double *sc64 = (double *)(smem + LL.sc64), *wd = (double *)(smem + LL.wd), *rc = (double *)(smem + LL.rc);
    int *rt = (int *)(smem + LL.rt);
    T *lm = (T *)(smem + LL.lm);
    T *dps = (T *)(smem + LL.hlp);  // dp per slot of the range (6R)
    const int lane = threadIdx.x;
    for (int c = blockIdx.x; c < A.nchunks; c += gridDim.x) {
        const Chunk ch = A.chunks[c];
        const WinInfo W = A.wins[ch.win];
        if (A.wstate[(size_t)NWS * ch.win + WS_CHOLOK] == 0.0) {
            if (lane == 0) A.partc2[c] = 0.0;
            continue;
        }
        const double *wp = A.win_par + (size_t)XBA_WPAR * ch.win;
        const int R = ch.R;
        T *sc2 = sc + R * SC_STRIDE;
        double *sc642 = sc64 + R * SC64_STRIDE;
        load_slots<T>(sc, sc64, A.Rwb, A.pwb, wp, W.slot0, ch.lo, R, lane);
        load_slots<T>(sc2, sc642, A.R2, A.p2, wp, W.slot0, ch.lo, R, lane);
        load_wconst<T>(wc, wp, lane);
        load_wd(wd, wp, lane);
        for (int i = lane; i < 6 * R; i += XBA_WAVE) {
            int s = i / 6, j = i % 6;
            dps[i] = (T)A.dx[W.st0 + A.slot_off[W.slot0 + ch.lo + s] + j];
        }
        double cost = 0.0;
        __syncthreads();
        for (int l = ch.l0; l < ch.l1; l++) {
            const int li = A.lm_info[l];
            const double a = A.lmp[3 * (size_t)l], b = A.lmp[3 * (size_t)l + 1], rho = A.lmp[3 * (size_t)l + 2];
            if (!((li >> 12) & 1)) {  // masked out: the trial state is the current one
                if (lane == 0) {
                    A.lmp_trial[3 * (size_t)l] = a;
                    A.lmp_trial[3 * (size_t)l + 1] = b;
                    A.lmp_trial[3 * (size_t)l + 2] = rho;
                }
                continue;
            }
            const int h = (li & 0xfff) - ch.lo;
            const int o0 = A.ob_start[l], no = A.ob_start[l + 1] - o0;
            const bool ok = A.lmok[l] != 0;
            // v = bl + sum_s Hlp[l,s] dp_s = bl + sum_q w_q Jl_q^T (Jt_q dp_t + Jh_q dp_h)
            if (ok) lin_observations<T>(A, rec, rc, rt, sc64, sc, wc, wd, o0, no, ch.lo, h, a, b, rho, lane, true);
            __syncthreads();
            if (ok && lane < 3) {
                int i = lane;
                T acc = 0;
                for (int q = 0; q < no; q++) {
                    T w = REC(F_W, q);
                    int t = rt[q];
                    if (w == (T)0 || t == h) continue;
                    int nd = (int)REC(F_ND, q);
                    T s = 0;
                    for (int r = 0; r < nd; r++) {
                        T jd = 0;
                        for (int j = 0; j < 6; j++)
                            jd += REC(F_JT + 6 * r + j, q) * dps[6 * t + j] + REC(F_JH + 6 * r + j, q) * dps[6 * h + j];
                        s += REC(F_JL + 3 * r + i, q) * jd;
                    }
                    acc += w * s;
                }
                lm[i] = acc;
            }
            __syncthreads();
            double na = a, nb = b, nr = rho;
            if (ok) {
                const T *Hi = A.lmHi + 6 * (size_t)l, *bl = A.lmbl + 3 * (size_t)l;
                T v0 = bl[0] + lm[0], v1 = bl[1] + lm[1], v2 = bl[2] + lm[2];
                T d0 = -(Hi[0] * v0 + Hi[1] * v1 + Hi[2] * v2);
                T d1 = -(Hi[1] * v0 + Hi[3] * v1 + Hi[4] * v2);
                T d2 = -(Hi[2] * v0 + Hi[4] * v1 + Hi[5] * v2);
                double r = rho + (double)d2;
                if (r > 1e-9 && isfinite(r)) {
                    na = a + (double)d0;
                    nb = b + (double)d1;
                    nr = r;
                }
            }
            if (lane == 0) {
                A.lmp_trial[3 * (size_t)l] = na;
                A.lmp_trial[3 * (size_t)l + 1] = nb;
                A.lmp_trial[3 * (size_t)l + 2] = nr;
            }
            // trial cost of the landmark's observations (a point behind the camera adds 1e6)
            double cq = 0.0;
            if (lane < no) {
                int k = o0 + lane;
                int info = A.ob_info[k];
                int t = (info & 0xfff) - ch.lo, cam = (info >> 12) & 1, valid = (info >> 13) & 1;
                if (valid) {
                    double e[2];
                    T Jl[2][3], Jt[2][6], Jh[2][6];
                    bool act = obs_eval<T>(sc642 + h * SC64_STRIDE, sc642 + t * SC64_STRIDE, sc2 + h * SC_STRIDE,
                                           sc2 + t * SC_STRIDE, wc, wd, na, nb, nr, cam, A.z[2 * (size_t)k],
                                           A.z[2 * (size_t)k + 1], t == h, false, e, Jl, Jt, Jh);
                    if (!act) {
                        cq = 1e6;
                    } else {
                        double isg = (double)A.isig[k], en2 = e[0] * isg * (e[0] * isg), w;
                        if (cam == 0) en2 += e[1] * isg * (e[1] * isg);
                        huber(sqrt(en2), wd[5], &w, &cq);
                        if (A.ob_w) cq *= A.ob_w[k];
                    }
                }
            }
            rc[lane] = cq;
            __syncthreads();
            if (lane == 0) {
                double cs = 0.0;
                for (int q = 0; q < no; q++) cs += rc[q];
                const T ris = A.rho_isig[l];
                if (ris > (T)0) {
                    double r0 = (nr - A.rho_meas[l]) * (double)ris, w, cc;
                    huber(fabs(r0), wd[5], &w, &cc);
                    cs += cc;
                }
                cost += cs;
            }
            __syncthreads();
        }
        if (lane == 0) A.partc2[c] = cost;
        __syncthreads();
    }
}

// Landmark commit of accepted windows (vio/window.py:967-970).
template <typename T>
__global__ void k_commit_lm(KArgs<T> A, const int32_t *lm_win) {
    for (int l = blockIdx.x * blockDim.x + threadIdx.x; l < A.nlm; l += gridDim.x * blockDim.x) {
        int w = lm_win[l];
        if (A.wstate[(size_t)NWS * w + WS_ACCEPT] != 0.0) {
            A.lmp[3 * (size_t)l] = A.lmp_trial[3 * (size_t)l];
            A.lmp[3 * (size_t)l + 1] = A.lmp_trial[3 * (size_t)l + 1];
            A.lmp[3 * (size_t)l + 2] = A.lmp_trial[3 * (size_t)l + 2];
        }
    }
}

template __global__ void k_lin<float>(KArgs<float>, int);
template __global__ void k_lin<double>(KArgs<double>, int);
template __global__ void k_trial<float>(KArgs<float>, int);
template __global__ void k_trial<double>(KArgs<double>, int);
template __global__ void k_commit_lm<float>(KArgs<float>, const int32_t *);
template __global__ void k_commit_lm<double>(KArgs<double>, const int32_t *);

}  // namespace xba
