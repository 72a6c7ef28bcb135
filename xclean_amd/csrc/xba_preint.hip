// IMU preintegration of many intervals at once, FP64 (vio/imu_factor.py:392-704 _step / preint_core, :707-746
// sqrt_info_core, :1008-1048 preintegrate; paths relative to the reference's src/xclean/).
//
// Sixteen lanes own one interval, four intervals share a wavefront (one wavefront per workgroup).  The sub-intervals
// are produced sixteen at a time, one per lane and in parallel: edges (t0, the released stamps inside, min(t1,
// t_end)), the bracket of the midpoint, the interpolated and bias-corrected rates, the sample-support scale and the
// two Exp / Jr evaluations, all into an LDS batch; so the sequential chain only reads its inputs from LDS.  Inside a
// sub-interval the group splits the 15x15 covariance propagation by columns: lane c keeps column c of P in
// registers, forms (F P)[:, c] with F's blocks rebuilt in registers from the sub-interval's N, T and midpoint
// attitude, the group transposes through LDS, lane c forms (F P F^T)[:, c] = F (F P)[c, :]^T, the symmetrisation
// goes through LDS once more, and lane c adds column c of the closed-form noise.  Lanes 0-5 keep the columns of the
// bias Jacobians; every lane keeps the mean.  The order of the chain, the rules (the bounded bracket, zero-length
// sub-intervals skipped, the hold past the last released sample, the refusal codes) and the arithmetic follow
// preint_core step for step.
#include "xba_internal.h"

namespace xba {

#define PRE_G 16   // lanes per interval, and sub-intervals per batch
#define PRE_NG 4   // intervals per workgroup (one wavefront)
#define PRE_SBW 63 // doubles per batch entry

enum { PI_STATUS, PI_NSUB, PI_NINNER, PI_EXTRAP, PI_TEND, PI_GAPMAX, PI_NREL, PI_TNOM, PI_SMAX };
enum { PI_OK = 0, PI_BAD_INTERVAL = 1, PI_NO_DATA = 2, PI_BEFORE_FIRST = 3, PI_NONFINITE = 4, PI_NO_RATE = 5 };
enum { PP_T0, PP_T1, PP_HZ, PP_BG, PP_BA = 6, PP_SG2 = 9, PP_SA2, PP_SWG2, PP_SWA2, PP_TNOM, PP_KAPPA, PP_SBG2,
       PP_SBA2, PP_N };
// batch entry: h, a (bias-corrected accelerometer, 3), white-noise scale, bracket gap, flags, Exp(w h), Jr(w h),
// Exp(w h / 2), Jr(w h / 2), the noise coefficients of the sub-interval (pre_qcoef)
enum { SB_H = 0, SB_A = 1, SB_SW = 4, SB_GAP = 5, SB_FL = 6, SB_DRK = 10, SB_JRF = 19, SB_RH = 28, SB_JRH = 37,
       SB_Q = 46 };
enum { FL_STEP = 1, FL_NONFINITE = 2, FL_SAMPLE = 4, FL_HOLD = 8 };

struct PreLds {
    double SB[PRE_G * PRE_SBW];
    double M[9], MMt[9], Rm[9], T[9];
    double Y[225];  // column c at Y[15 c + i]
};

// y = F x for one column x of the internal covariance (vio/imu_factor.py:366-389).  F's blocks are formed from
// Exp(w h) and Jr(w h) (LDS) and N, T, Rm (registers) exactly as _step forms them: Rt = Exp^T, F1 = -N h,
// F2 = -0.5 N h^2, G0 = Jr h, G1 = -T (0.5 h^2), G2 = -T (0.25 h^3), A1 = Rm h, A2 = 0.5 Rm h^2.
__device__ inline void pre_fcol(const double *dRk, const double *Jrf, const double *N, const double *T,
                                const double *Rm, double h, const double *x, double *y) {
    const double h2 = h * h, h3 = h2 * h;
    for (int i = 0; i < 3; i++) {
        double s0 = 0.0, s1 = 0.0, s2 = 0.0;
        for (int k = 0; k < 3; k++) {
            const int e = 3 * i + k;
            const double xph = x[k], xg = x[9 + k], xa = x[12 + k];
            s0 += dRk[3 * k + i] * xph + (Jrf[e] * h) * xg;
            s1 += (-N[e] * h) * xph + (-T[e] * (0.5 * h2)) * xg + (Rm[e] * h) * xa;
            s2 += (-0.5 * N[e] * h2) * xph + (-T[e] * (0.25 * h3)) * xg + (0.5 * Rm[e] * h2) * xa;
        }
        y[i] = s0;
        y[3 + i] = s1 + x[3 + i];
        y[6 + i] = s2 + h * x[3 + i] + x[6 + i];
    }
    for (int i = 9; i < 15; i++) y[i] = x[i];
}

// The sub-interval's noise coefficients (vio/imu_factor.py:459-500) from h and the scaled densities, in the order
// the entries use them: 0 qpp, 1 qvp1, 2 qpp1, 3 qvv_i, 4 qvv_m, 5 qpv_i, 6 qpv_m, 7 qpp_i, 8 qpp_m, 9 swg2 h^2 / 2,
// 10 swg2 h^3 / 6, 11 swg2 h^4 / 24, 12 swg2 h, 13 swa2 h^2 / 2, 14 swa2 h^3 / 6, 15 swa2 h, 16 zero.
__device__ inline void pre_qcoef(double h, double sg2s, double sa2s, double swg2, double swa2, double *q) {
    const double h2 = h * h, h3 = h2 * h, h4 = h3 * h, h5 = h4 * h, h6 = h5 * h, h7 = h6 * h;
    q[0] = sg2s * h + swg2 * h3 / 3.0;
    q[1] = sg2s * h2 / 2.0 + swg2 * h4 / 8.0;
    q[2] = sg2s * h3 / 6.0 + swg2 * h5 / 30.0;
    q[3] = sa2s * h + swa2 * h3 / 3.0;
    q[4] = sg2s * h3 / 3.0 + swg2 * h5 / 20.0;
    q[5] = sa2s * h2 / 2.0 + swa2 * h4 / 8.0;
    q[6] = sg2s * h4 / 8.0 + swg2 * h6 / 72.0;
    q[7] = sa2s * h3 / 3.0 + swa2 * h5 / 20.0;
    q[8] = sg2s * h5 / 20.0 + swg2 * h7 / 252.0;
    q[9] = swg2 * h2 / 2.0;
    q[10] = swg2 * h3 / 6.0;
    q[11] = swg2 * h4 / 24.0;
    q[12] = swg2 * h;
    q[13] = swa2 * h2 / 2.0;
    q[14] = swa2 * h3 / 6.0;
    q[15] = swa2 * h;
    q[16] = 0.0;
}

// Column c of the sub-interval's noise added into p (c = 3 cb + j; blocks phi v p eps_g eps_a).  Every entry is
// q[dq] d + q[mm] MMt - q[mq] M + q[rq] Rm with d = [i == j] and the in-block indices of M and Rm ordered (v/p
// block, other block); for each pair of blocks the coefficient indices (16: the zero) come from the entry table of
// vio/imu_factor.py:459-500, with the block above the diagonal read as the transpose of the one below.  Branch-free:
// the lanes of a group hold different column blocks.
__device__ inline void pre_qcol(int c, const double *q, const PreLds &S, double *p) {
    const int cb = c / 3, j = c % 3;
    for (int rb = 0; rb < 5; rb++) {
        const bool lo = rb >= cb;
        const int A = lo ? rb : cb, Bk = lo ? cb : rb, cs = 5 * A + Bk;
        const int dq = cs == 0 ? 0 : cs == 6 ? 3 : cs == 11 ? 5 : cs == 12 ? 7 : cs == 15 ? 9 : cs == 18 ? 12
                     : cs == 24 ? 15 : 16;
        const int mq = cs == 5 ? 1 : cs == 10 ? 2 : cs == 16 ? 10 : cs == 17 ? 11 : 16;
        const int mmq = cs == 6 ? 4 : cs == 11 ? 6 : cs == 12 ? 8 : 16;
        const int rq = cs == 21 ? 13 : cs == 22 ? 14 : 16;
        const double qd = q[dq], qm = q[mq], qmm = q[mmq], qr = q[rq];
        const bool mrow = lo ? (A < 3) : (A == 3);  // M[i][j] (else M[j][i])
        for (int ii = 0; ii < 3; ii++) {
            const int mi = mrow ? 3 * ii + j : 3 * j + ii, ri = lo ? 3 * j + ii : 3 * ii + j;
            const double d = ii == j ? 1.0 : 0.0;
            p[3 * rb + ii] += ((qd * d - qm * S.M[mi]) + qmm * S.MMt[3 * ii + j]) + qr * S.Rm[ri];
        }
    }
}

// A <- kappa A + J_bg sbg2 J_bg^T + J_ba sba2 J_ba^T on the nine preintegration rows (vio/imu_factor.py:707-722),
// A row-major 15 x 15, J row-major 9 x 6 (columns bg, ba).
__device__ inline void pre_info_matrix(double *A, const double *J, double kappa, double sbg2, double sba2) {
    for (int i = 0; i < 225; i++) A[i] = kappa * A[i];
    if (sbg2 > 0.0 || sba2 > 0.0)
        for (int i = 0; i < 9; i++)
            for (int j = 0; j < 9; j++) {
                double s = 0.0;
                for (int q = 0; q < 3; q++)
                    s += sbg2 * J[6 * i + q] * J[6 * j + q] + sba2 * J[6 * i + 3 + q] * J[6 * j + 3 + q];
                A[15 * i + j] += s;
            }
}

// The whitening of vio/imu_factor.py:723-746 in place: the lower Cholesky factor L of A (lower triangle read), then
// W = L^-1 column by column over L's own storage (column j of W needs L(i, q) for q >= j only, and L(i, j) is read
// before W(i, j) replaces it); the upper triangle is zeroed.  False (A then undefined) when A is not positive
// definite.
__device__ inline bool pre_whiten_inplace(double *A) {
    for (int j = 0; j < 15; j++) {
        double s = A[15 * j + j];
        for (int q = 0; q < j; q++) s -= A[15 * j + q] * A[15 * j + q];
        if (!(s > 0.0)) return false;
        const double dj = sqrt(s);
        A[15 * j + j] = dj;
        for (int i = j + 1; i < 15; i++) {
            s = A[15 * i + j];
            for (int q = 0; q < j; q++) s -= A[15 * i + q] * A[15 * j + q];
            A[15 * i + j] = s / dj;
        }
    }
    for (int j = 0; j < 15; j++) {
        const double ljj = A[15 * j + j];
        A[15 * j + j] = 1.0 / ljj;
        for (int i = j + 1; i < 15; i++) {
            double s = 0.0;
            for (int q = j; q < i; q++) s += A[15 * i + q] * A[15 * q + j];
            A[15 * i + j] = -s / A[15 * i + i];
        }
        for (int i = 0; i < j; i++) A[15 * i + j] = 0.0;
    }
    return true;
}

__global__ __launch_bounds__(64) void k_preint(const double *__restrict__ ts, const double *__restrict__ gyr,
                                                const double *__restrict__ acc, const int64_t *__restrict__ s0,
                                                const int32_t *__restrict__ ns, const double *__restrict__ par, int nf,
                                                double *oR, double *ov, double *op, double *oJ, double *oC,
                                                double *oinfo, double *oW, int32_t *wok, double *orec) {
    __shared__ PreLds lds[PRE_NG];
    const int g = threadIdx.x / PRE_G, c = threadIdx.x % PRE_G;
    const int f = blockIdx.x * PRE_NG + g;
    PreLds &S = lds[g];
    const bool live = f < nf;
    const double *pr = par + (size_t)PP_N * (live ? f : 0);
    const double *t = ts, *gy = gyr, *ac = acc;
    int n = 0;
    if (live) {
        t = ts + s0[f];
        gy = gyr + 3 * s0[f];
        ac = acc + 3 * s0[f];
        n = ns[f];
    }
    const double t0 = pr[PP_T0], t1 = pr[PP_T1], hz = pr[PP_HZ], tnom = pr[PP_TNOM];
    const double bg[3] = {pr[PP_BG], pr[PP_BG + 1], pr[PP_BG + 2]}, ba[3] = {pr[PP_BA], pr[PP_BA + 1], pr[PP_BA + 2]};
    const double sg2 = pr[PP_SG2], sa2 = pr[PP_SA2], swg2 = pr[PP_SWG2], swa2 = pr[PP_SWA2];
    double info[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    double dR[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1}, dv[3] = {0, 0, 0}, dp[3] = {0, 0, 0};
    double P[15], J[9];
    for (int i = 0; i < 15; i++) P[i] = 0.0;
    for (int i = 0; i < 9; i++) J[i] = 0.0;
    // released count and refusals (stamps only)
    int lo = 0, hi = n;
    while (lo < hi) {
        const int mid = (lo + hi) / 2;
        if (t[mid] <= hz) lo = mid + 1; else hi = mid;
    }
    const int nrel = lo;
    info[PI_NREL] = nrel;
    int status = PI_OK;
    const bool rate_ok = tnom > 0.0 && isfinite(tnom);
    if (!(t1 > t0) || !isfinite(t0) || !isfinite(t1)) status = PI_BAD_INTERVAL;
    else if (nrel == 0) status = PI_NO_DATA;
    else if (t0 < t[0]) status = PI_BEFORE_FIRST;
    else if ((sg2 > 0.0 || sa2 > 0.0) && !rate_ok) status = PI_NO_RATE;
    const double te = (status == PI_OK) ? t[nrel - 1] : 0.0;
    if (status == PI_OK) {
        info[PI_TNOM] = rate_ok ? tnom : 0.0;
        info[PI_TEND] = te;
    }
    // phase 0: interior sub-intervals (when te > t0), 1: the hold, 2: no more batches
    int phase = (!live || status != PI_OK) ? 2 : (te > t0 ? 0 : 1);
    const double tie = t1 < te ? t1 : te;
    int kc = 0;  // next released stamp that can end an interior sub-interval
    double e0 = t0, smax = 1.0, gapmax = 0.0;
    int nsub = 0, ninner = 0, bj = 0, bn = 0;
    if (phase == 0) {
        lo = 0;
        hi = nrel;
        while (lo < hi) {
            const int mid = (lo + hi) / 2;
            if (t[mid] <= t0) lo = mid + 1; else hi = mid;
        }
        kc = lo;
    }
    for (;;) {
        // ---- groups with an exhausted batch produce the next one, one sub-interval per lane
        const bool need = phase < 2 && bj >= bn;
        if (need) {
            double *E = S.SB + PRE_SBW * c;
            double h = 0.0, wv[3] = {0, 0, 0}, av[3] = {0, 0, 0}, sw = 1.0, gap = 0.0;
            int fl = 0;
            bool mine = false;
            if (phase == 0) {
                // lane c: the sub-interval ending at stamp kc + c, or at tie if that stamp is not before tie; it
                // exists when every lane before it ended at a stamp
                const int kj = kc + c;
                const bool samp = kj < nrel && t[kj] < tie;
                const unsigned gb = (unsigned)((__ballot(samp) >> (PRE_G * g)) & 0xffffu);
                const int last = gb == 0xffffu ? PRE_G - 1 : __ffs(~gb) - 1;  // the first lane ending at tie
                bn = last + 1;
                mine = c <= last;
                if (mine) {
                    const double ea = c == 0 ? e0 : t[kj - 1], eb = samp ? t[kj] : tie;
                    h = eb - ea;
                    fl = samp ? FL_SAMPLE : 0;
                    if (h > 0.0) {
                        // the bracket of the midpoint: the last released stamp <= m, at most nrel - 2 (the walk of
                        // preint_core from the previous sub-interval ends there)
                        const double m = 0.5 * (ea + eb);
                        const int left = kj - 1 < nrel - 2 ? kj - 1 : nrel - 2;
                        const double ta = t[left];
                        gap = t[left + 1] - ta;
                        for (int q = 0; q < 3; q++) {
                            double sl = (gy[3 * (left + 1) + q] - gy[3 * left + q]) / gap;
                            wv[q] = sl * (m - ta) + gy[3 * left + q] - bg[q];
                            sl = (ac[3 * (left + 1) + q] - ac[3 * left + q]) / gap;
                            av[q] = sl * (m - ta) + ac[3 * left + q] - ba[q];
                        }
                        const bool fin = isfinite(wv[0]) && isfinite(wv[1]) && isfinite(wv[2]) && isfinite(av[0]) &&
                                         isfinite(av[1]) && isfinite(av[2]);
                        fl |= fin ? FL_STEP : FL_NONFINITE;
                        if (rate_ok) {
                            const double u = (m - ta) / gap;
                            sw = 2.0 * ((1.0 - u) * (1.0 - u) + u * u) * h / tnom;
                            if (sw < 1.0) sw = 1.0;
                        }
                    }
                }
                if (gb == 0xffffu) {
                    kc += PRE_G;
                    e0 = t[kc - 1];
                } else {
                    phase = 1;
                }
            } else {
                // the hold of the last released sample over (t_end, t1]
                phase = 2;
                bn = t1 > te ? 1 : 0;
                mine = c == 0 && bn == 1;
                if (mine) {
                    h = t1 - (t0 > te ? t0 : te);
                    for (int q = 0; q < 3; q++) {
                        wv[q] = gy[3 * (nrel - 1) + q] - bg[q];
                        av[q] = ac[3 * (nrel - 1) + q] - ba[q];
                    }
                    const bool fin = isfinite(wv[0]) && isfinite(wv[1]) && isfinite(wv[2]) && isfinite(av[0]) &&
                                     isfinite(av[1]) && isfinite(av[2]);
                    fl = FL_HOLD | (fin ? FL_STEP : FL_NONFINITE);
                    if (rate_ok) {
                        sw = (te > t0 ? 2.0 : 1.0) * h / tnom;
                        if (sw < 1.0) sw = 1.0;
                    }
                }
            }
            bj = 0;
            if (mine) {
                E[SB_H] = h;
                for (int q = 0; q < 3; q++) E[SB_A + q] = av[q];
                E[SB_SW] = sw;
                E[SB_GAP] = gap;
                E[SB_FL] = (double)fl;
                if (fl & FL_STEP) {
                    pre_qcoef(h, sg2 * sw, sa2 * sw, swg2, swa2, E + SB_Q);
                    const double wh[3] = {wv[0] * h, wv[1] * h, wv[2] * h};
                    const double wm[3] = {0.5 * wh[0], 0.5 * wh[1], 0.5 * wh[2]};
                    double R[9], Jr[9];
                    exp_jr(wh, R, Jr);
                    for (int q = 0; q < 9; q++) {
                        E[SB_DRK + q] = R[q];
                        E[SB_JRF + q] = Jr[q];
                    }
                    exp_jr(wm, R, Jr);
                    for (int q = 0; q < 9; q++) {
                        E[SB_RH + q] = R[q];
                        E[SB_JRH + q] = Jr[q];
                    }
                }
            }
        }
        WAVE_SYNC();
        // ---- the group's next sub-interval with a step, in order
        bool have = false;
        while (bj < bn) {
            const double *E = S.SB + PRE_SBW * bj;
            const int fl = (int)E[SB_FL];
            if (fl & FL_SAMPLE) ninner++;
            if (fl & FL_NONFINITE) {
                status = PI_NONFINITE;
                phase = 2;
                bn = 0;
                break;
            }
            if (fl & FL_HOLD) {
                if (te > t0) ninner++;
                info[PI_EXTRAP] = E[SB_H];
            }
            bj++;
            if (fl & FL_STEP) {
                have = true;
                break;
            }
        }
        if (__ballot(have) == 0) {
            if (__ballot(phase < 2) == 0) break;
            continue;
        }
        if (!have) continue;
        // ---- one sub-interval (vio/imu_factor.py:392-542), the whole group
        const double *E = S.SB + PRE_SBW * (bj - 1);
        const double h = E[SB_H], sw = E[SB_SW], a[3] = {E[SB_A], E[SB_A + 1], E[SB_A + 2]};
        const double *dRk = E + SB_DRK, *Jrf = E + SB_JRF, *Rh = E + SB_RH, *Jrh = E + SB_JRH;
        nsub++;
        if (E[SB_GAP] > gapmax) gapmax = E[SB_GAP];
        if (sw > smax) smax = sw;
        double Rm[9], M[9], N[9], T[9], am[3];
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++)
                Rm[3 * i + j] = dR[3 * i] * Rh[j] + dR[3 * i + 1] * Rh[3 + j] + dR[3 * i + 2] * Rh[6 + j];
        for (int i = 0; i < 3; i++) {
            const double *r = Rm + 3 * i;
            am[i] = r[0] * a[0] + r[1] * a[1] + r[2] * a[2];
            M[3 * i + 0] = r[1] * a[2] - r[2] * a[1];
            M[3 * i + 1] = -r[0] * a[2] + r[2] * a[0];
            M[3 * i + 2] = r[0] * a[1] - r[1] * a[0];
        }
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) {
                double nn = 0.0, tt = 0.0;
                for (int q = 0; q < 3; q++) {
                    nn += M[3 * i + q] * Rh[3 * j + q];
                    tt += M[3 * i + q] * Jrh[3 * q + j];
                }
                N[3 * i + j] = nn;
                T[3 * i + j] = tt;
            }
        if (c == 0)
            for (int e = 0; e < 9; e++) {
                const int i = e / 3, j = e % 3;
                S.M[e] = M[e];
                S.MMt[e] = M[3 * i] * M[3 * j] + M[3 * i + 1] * M[3 * j + 1] + M[3 * i + 2] * M[3 * j + 2];
                S.Rm[e] = Rm[e];
                S.T[e] = T[e];
            }
        const double h2 = h * h, h3 = h2 * h;
        // covariance: column c of F P, transposed through LDS, column c of F (F P)^T, symmetrised, plus the noise
        double y[15];
        if (c < 15) {
            pre_fcol(dRk, Jrf, N, T, Rm, h, P, y);
            for (int i = 0; i < 15; i++) S.Y[15 * c + i] = y[i];
        }
        WAVE_SYNC();
        if (c < 15) {
            double z[15];
            for (int i = 0; i < 15; i++) z[i] = S.Y[15 * i + c];  // row c of F P
            pre_fcol(dRk, Jrf, N, T, Rm, h, z, y);
        }
        WAVE_SYNC();
        if (c < 15)
            for (int i = 0; i < 15; i++) S.Y[15 * c + i] = y[i];
        WAVE_SYNC();
        if (c < 15) {
            for (int i = 0; i < 15; i++) P[i] = (i == c) ? y[i] : 0.5 * (y[i] + S.Y[15 * i + c]);
            pre_qcol(c, E + SB_Q, S, P);
        }
        // bias Jacobians: J <- A9 J - B (lanes 0-5, column c)
        if (c < 6) {
            double jn[9];
            for (int i = 0; i < 3; i++) {
                double a0 = 0.0, a1 = 0.0, a2 = 0.0;
                for (int q = 0; q < 3; q++) {
                    const int e = 3 * i + q;
                    a0 += dRk[3 * q + i] * J[q];
                    a1 += (-N[e] * h) * J[q];
                    a2 += (-0.5 * N[e] * h2) * J[q];
                }
                jn[i] = a0;
                jn[3 + i] = a1 + J[3 + i];
                jn[6 + i] = a2 + h * J[3 + i] + J[6 + i];
            }
            const int cc = c % 3;
            for (int i = 0; i < 3; i++) {
                const int e = 3 * i + cc;
                if (c < 3) {
                    jn[i] -= Jrf[e] * h;
                    jn[3 + i] -= -S.T[e] * (0.5 * h2);
                    jn[6 + i] -= -S.T[e] * (0.25 * h3);
                } else {
                    jn[3 + i] -= S.Rm[e] * h;
                    jn[6 + i] -= 0.5 * S.Rm[e] * h2;
                }
            }
            for (int i = 0; i < 9; i++) J[i] = jn[i];
        }
        // the mean
        for (int i = 0; i < 3; i++) {
            dp[i] = dp[i] + dv[i] * h + 0.5 * am[i] * h2;
            dv[i] = dv[i] + am[i] * h;
        }
        double Rn[9];
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++)
                Rn[3 * i + j] = dR[3 * i] * dRk[j] + dR[3 * i + 1] * dRk[3 + j] + dR[3 * i + 2] * dRk[6 + j];
        for (int i = 0; i < 9; i++) dR[i] = Rn[i];
        WAVE_SYNC();  // LDS is rewritten by the next sub-interval or batch
    }
    if (!live) return;
    info[PI_STATUS] = status;
    double *io = oinfo + 9 * (size_t)f;
    if (status != PI_OK) {
        // nothing but info is meaningful; outputs zeroed like the reference's initialisation
        if (c == 0) {
            for (int i = 0; i < 9; i++) {
                io[i] = info[i];
                oR[9 * (size_t)f + i] = (i % 4 == 0) ? 1.0 : 0.0;
            }
            for (int i = 0; i < 3; i++) ov[3 * (size_t)f + i] = op[3 * (size_t)f + i] = 0.0;
            if (wok) wok[f] = 0;
            if (oW)
                for (int i = 0; i < 225; i++) oW[225 * (size_t)f + i] = 0.0;
            if (orec)
                for (int i = 0; i < XBA_IMU_REC; i++) orec[(size_t)XBA_IMU_REC * f + i] = 0.0;
        }
        if (c < 15)
            for (int i = 0; i < 15; i++) oC[225 * (size_t)f + 15 * i + c] = 0.0;
        if (c < 6)
            for (int i = 0; i < 9; i++) oJ[54 * (size_t)f + 6 * i + c] = 0.0;
        return;
    }
    // one Newton-Schulz polar step
    {
        double Tm[9], Rn[9];
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++)
                Tm[3 * i + j] = dR[i] * dR[j] + dR[3 + i] * dR[3 + j] + dR[6 + i] * dR[6 + j];
        for (int i = 0; i < 9; i++) Tm[i] = (i % 4 == 0 ? 1.5 : 0.0) - 0.5 * Tm[i];
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++)
                Rn[3 * i + j] = dR[3 * i] * Tm[j] + dR[3 * i + 1] * Tm[3 + j] + dR[3 * i + 2] * Tm[6 + j];
        for (int i = 0; i < 9; i++) dR[i] = Rn[i];
    }
    info[PI_NSUB] = nsub;
    info[PI_NINNER] = ninner;
    info[PI_GAPMAX] = gapmax;
    info[PI_SMAX] = smax;
    // residual covariance: the preintegration / bias cross block flips sign
    double *Co = oC + 225 * (size_t)f;
    if (c < 15)
        for (int i = 0; i < 15; i++) {
            const double v = ((i < 9) != (c < 9)) ? -P[i] : P[i];
            Co[15 * i + c] = v;
            S.Y[15 * i + c] = v;  // row-major copy for the whitening
        }
    if (c < 6)
        for (int i = 0; i < 9; i++) oJ[54 * (size_t)f + 6 * i + c] = J[i];
    if (c == 0) {
        for (int i = 0; i < 9; i++) {
            io[i] = info[i];
            oR[9 * (size_t)f + i] = dR[i];
        }
        for (int i = 0; i < 3; i++) {
            ov[3 * (size_t)f + i] = dv[i];
            op[3 * (size_t)f + i] = dp[i];
        }
    }
    if (!oW && !orec) return;
    // whitening (vio/imu_factor.py:707-746): W = L^-1, L L^T = kappa C + BIASCOV; J is needed by lane 0
    WAVE_SYNC();
    double *Js = S.SB;  // the batch is no longer needed
    if (c < 6)
        for (int i = 0; i < 9; i++) Js[6 * i + c] = J[i];
    WAVE_SYNC();
    if (c == 0) {
        double *A = S.Y;
        pre_info_matrix(A, Js, pr[PP_KAPPA], pr[PP_SBG2], pr[PP_SBA2]);
        const bool ok = pre_whiten_inplace(A);
        // W (or zero when kappa C + BIASCOV is not positive definite) to the output and into the solver's record
        double *Wd = oW ? oW + 225 * (size_t)f : nullptr;
        double *rec = orec ? orec + (size_t)XBA_IMU_REC * f : nullptr;
        double *Wr = rec ? rec + 76 : nullptr;
        for (int i = 0; i < 225; i++) {
            const double v = ok ? A[i] : 0.0;
            if (Wd) Wd[i] = v;
            if (Wr) Wr[i] = v;
        }
        if (wok) wok[f] = ok ? 1 : 0;
        if (rec) {  // the window solver's IMU record: dR 9, dv 3, dp 3, dt 1, J 54, bg_lin 3, ba_lin 3, W 225
            for (int i = 0; i < 9; i++) rec[i] = dR[i];
            for (int i = 0; i < 3; i++) {
                rec[9 + i] = dv[i];
                rec[12 + i] = dp[i];
                rec[70 + i] = bg[i];
                rec[73 + i] = ba[i];
            }
            rec[15] = t1 - t0;
            for (int i = 0; i < 54; i++) rec[16 + i] = Js[i];
        }
    }
}

// vio/imu_factor.py:1051-1063 `sqrt_information` for intervals already preintegrated: one thread per interval,
// W = L^-1 with L L^T = kappa C + BIASCOV formed and factored in W's own storage (C, J in the Preint layout:
// row-major 15 x 15 residual-order covariance and 9 x 6 bias Jacobians; par (nf, 3): kappa, sbg2, sba2).  w_ok 0 and
// W zero when the matrix is not positive definite.
__global__ __launch_bounds__(64) void k_sqrt_info(int nf, const double *__restrict__ C, const double *__restrict__ J,
                                                  const double *__restrict__ par, double *W, int32_t *w_ok) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nf) return;
    double *A = W + 225 * (size_t)f;
    const double *Cf = C + 225 * (size_t)f;
    for (int i = 0; i < 225; i++) A[i] = Cf[i];
    pre_info_matrix(A, J + 54 * (size_t)f, par[3 * (size_t)f], par[3 * (size_t)f + 1], par[3 * (size_t)f + 2]);
    const bool ok = pre_whiten_inplace(A);
    if (!ok)
        for (int i = 0; i < 225; i++) A[i] = 0.0;
    w_ok[f] = ok ? 1 : 0;
}

}  // namespace xba
