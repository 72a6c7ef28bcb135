// IMU preintegration of many intervals at once, FP64 (vio/imu_factor.py:392-704 _step / preint_core, :707-746
// sqrt_info_core, :1008-1048 preintegrate; paths relative to the reference's src/xclean/).
//
// Sixteen lanes own one interval, four intervals share a wavefront (one wavefront per workgroup).  The chain of
// sub-intervals is sequential; inside one sub-interval the group splits the 15x15 covariance propagation by columns:
// lane c keeps column c of P in registers, forms (F P)[:, c], the group transposes through LDS, lane c forms
// (F P F^T)[:, c] = F (F P)[c, :]^T, the symmetrisation goes through LDS once more, and lane c adds column c of the
// sub-interval's closed-form noise.  Lanes 0-5 keep the columns of the bias Jacobians.  The mean (dR, dv, dp) and
// the sub-interval's 3x3 blocks are formed by every lane of the group (the same arithmetic, so the same values);
// lane 0 publishes the blocks in LDS for the column updates.  The interval rules (edges, midpoint interpolation,
// the bounded bracket, the hold past the last released sample, the sample-support scale, the refusals) follow
// preint_core step for step; every lane of a group runs them on the same data.
#include "xba_internal.h"

namespace xba {

#define PRE_G 16  // lanes per interval
#define PRE_NG 4  // intervals per workgroup (one wavefront)

enum { PI_STATUS, PI_NSUB, PI_NINNER, PI_EXTRAP, PI_TEND, PI_GAPMAX, PI_NREL, PI_TNOM, PI_SMAX };
enum { PI_OK = 0, PI_BAD_INTERVAL = 1, PI_NO_DATA = 2, PI_BEFORE_FIRST = 3, PI_NONFINITE = 4, PI_NO_RATE = 5 };
enum { PP_T0, PP_T1, PP_HZ, PP_BG, PP_BA = 6, PP_SG2 = 9, PP_SA2, PP_SWG2, PP_SWA2, PP_TNOM, PP_KAPPA, PP_SBG2,
       PP_SBA2, PP_N };

struct PreLds {
    double B[72];  // the transition's 3x3 blocks: Rt F1 F2 G0 G1 G2 A1 A2 (row-major each)
    double M[9], MMt[9], Rm[9];
    double Y[225];  // column c at Y[15 c + i]
};

// Y[:, c] = F X[:, c] for the sub-interval transition (vio/imu_factor.py:366-389), X[:, c] = x.
__device__ inline void pre_fcol(const double *B, double h, const double *x, double *y) {
    const double *Rt = B, *F1 = B + 9, *F2 = B + 18, *G0 = B + 27, *G1 = B + 36, *G2 = B + 45, *A1 = B + 54,
                 *A2 = B + 63;
    for (int i = 0; i < 3; i++) {
        double s0 = 0.0, s1 = 0.0, s2 = 0.0;
        for (int k = 0; k < 3; k++) {
            const double xph = x[k], xg = x[9 + k], xa = x[12 + k];
            s0 += Rt[3 * i + k] * xph + G0[3 * i + k] * xg;
            s1 += F1[3 * i + k] * xph + G1[3 * i + k] * xg + A1[3 * i + k] * xa;
            s2 += F2[3 * i + k] * xph + G2[3 * i + k] * xg + A2[3 * i + k] * xa;
        }
        y[i] = s0;
        y[3 + i] = s1 + x[3 + i];
        y[6 + i] = s2 + h * x[3 + i] + x[6 + i];
    }
    for (int i = 9; i < 15; i++) y[i] = x[i];
}

// Entry (rb i, cb j) of the sub-interval's noise for rb >= cb (block index, in-block index; blocks phi v p eps_g
// eps_a), vio/imu_factor.py:459-500.  q: 0 qpp, 1 qvp1, 2 qpp1, 3 qvv_i, 4 qvv_m, 5 qpv_i, 6 qpv_m, 7 qpp_i, 8 qpp_m,
// 9 swg2, 10 swa2, 11 h, 12 h^2, 13 h^3, 14 h^4; M, MMt, Rm (the midpoint attitude) in LDS.
__device__ inline double pre_qlo(int rb, int i, int cb, int j, const double *q, const PreLds &S) {
    const double d = i == j ? 1.0 : 0.0;
    switch (rb * 5 + cb) {
        case 0: return q[0] * d;                                          // (phi, phi)
        case 5: return -S.M[3 * i + j] * q[1];                            // (v, phi)
        case 6: return q[3] * d + q[4] * S.MMt[3 * i + j];                // (v, v)
        case 10: return -S.M[3 * i + j] * q[2];                           // (p, phi)
        case 11: return q[5] * d + q[6] * S.MMt[3 * i + j];               // (p, v)
        case 12: return q[7] * d + q[8] * S.MMt[3 * i + j];               // (p, p)
        case 15: return q[9] * q[12] / 2.0 * d;                           // (eps_g, phi)
        case 16: return -S.M[3 * j + i] * q[9] * q[13] / 6.0;             // (eps_g, v) = (v, eps_g)^T
        case 17: return -S.M[3 * j + i] * q[9] * q[14] / 24.0;            // (eps_g, p)
        case 18: return q[9] * q[11] * d;                                 // (eps_g, eps_g)
        case 21: return S.Rm[3 * j + i] * q[10] * q[12] / 2.0;            // (eps_a, v) = (v, eps_a)^T
        case 22: return S.Rm[3 * j + i] * q[10] * q[13] / 6.0;            // (eps_a, p)
        case 24: return q[10] * q[11] * d;                                // (eps_a, eps_a)
        default: return 0.0;
    }
}

__device__ inline double pre_qe(int r, int c, const double *q, const PreLds &S) {
    const int rb = r / 3, cb = c / 3;
    return rb >= cb ? pre_qlo(rb, r % 3, cb, c % 3, q, S) : pre_qlo(cb, c % 3, rb, r % 3, q, S);
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
    // the edge walk: phase 0 interior sub-intervals (when te > t0), phase 1 the hold, phase 2 done
    int phase = (!live || status != PI_OK) ? 2 : (te > t0 ? 0 : 1);
    const double tie = t1 < te ? t1 : te;
    int k = 0, left = 0;
    double e0 = t0, smax = 1.0, gapmax = 0.0;
    int nsub = 0, ninner = 0;
    if (phase == 0) {
        lo = 0;
        hi = nrel;
        while (lo < hi) {
            const int mid = (lo + hi) / 2;
            if (t[mid] <= t0) lo = mid + 1; else hi = mid;
        }
        k = lo;
        left = k - 1;
    }
    for (;;) {
        // ---- the group's next sub-interval (h, w, a, white-noise scale), or none
        bool have = false;
        double h = 0.0, w[3] = {0, 0, 0}, a[3] = {0, 0, 0}, sw = 1.0;
        while (phase == 0 && !have) {
            const bool samp = k < nrel && t[k] < tie;
            const double e1 = samp ? t[k] : tie;
            h = e1 - e0;
            if (h > 0.0) {
                const double m = 0.5 * (e0 + e1);
                while (left + 2 < nrel && t[left + 1] <= m) left++;
                const double ta = t[left], gap = t[left + 1] - ta;
                if (gap > gapmax) gapmax = gap;
                for (int q = 0; q < 3; q++) {
                    double s = (gy[3 * (left + 1) + q] - gy[3 * left + q]) / gap;
                    w[q] = s * (m - ta) + gy[3 * left + q] - bg[q];
                    s = (ac[3 * (left + 1) + q] - ac[3 * left + q]) / gap;
                    a[q] = s * (m - ta) + ac[3 * left + q] - ba[q];
                }
                if (!(isfinite(w[0]) && isfinite(w[1]) && isfinite(w[2]) && isfinite(a[0]) && isfinite(a[1]) &&
                      isfinite(a[2]))) {
                    status = PI_NONFINITE;
                    phase = 2;
                    break;
                }
                if (rate_ok) {
                    const double u = (m - ta) / gap;
                    sw = 2.0 * ((1.0 - u) * (1.0 - u) + u * u) * h / tnom;
                    if (sw < 1.0) sw = 1.0;
                    if (sw > smax) smax = sw;
                }
                have = true;
                nsub++;
            }
            if (samp) {
                ninner++;
                e0 = e1;
                k++;
            } else {
                phase = 1;
            }
        }
        if (!have && phase == 1) {
            phase = 2;
            if (t1 > te) {
                const double tsx = t0 > te ? t0 : te;
                h = t1 - tsx;
                for (int q = 0; q < 3; q++) {
                    w[q] = gy[3 * (nrel - 1) + q] - bg[q];
                    a[q] = ac[3 * (nrel - 1) + q] - ba[q];
                }
                if (!(isfinite(w[0]) && isfinite(w[1]) && isfinite(w[2]) && isfinite(a[0]) && isfinite(a[1]) &&
                      isfinite(a[2]))) {
                    status = PI_NONFINITE;
                } else {
                    if (rate_ok) {
                        sw = (te > t0 ? 2.0 : 1.0) * h / tnom;
                        if (sw < 1.0) sw = 1.0;
                        if (sw > smax) smax = sw;
                    }
                    have = true;
                    nsub++;
                    if (te > t0) ninner++;
                    info[PI_EXTRAP] = h;
                }
            }
        }
        if (__ballot(have) == 0) break;
        if (!have) continue;
        // ---- one sub-interval (vio/imu_factor.py:392-542), the whole group
        const double wh[3] = {w[0] * h, w[1] * h, w[2] * h}, wm[3] = {0.5 * wh[0], 0.5 * wh[1], 0.5 * wh[2]};
        double dRk[9], Jrf[9], Rh[9], Jrh[9], Rm[9], am[3], M[9];
        exp_jr(wh, dRk, Jrf);
        exp_jr(wm, Rh, Jrh);
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
        const double h2 = h * h, h3 = h2 * h;
        if (c == 0) {
            double *B = S.B;
            for (int i = 0; i < 3; i++)
                for (int j = 0; j < 3; j++) {
                    const int e = 3 * i + j;
                    double Nij = 0.0, Tij = 0.0, MM = 0.0;
                    for (int q = 0; q < 3; q++) {
                        Nij += M[3 * i + q] * Rh[3 * j + q];
                        Tij += M[3 * i + q] * Jrh[3 * q + j];
                        MM += M[3 * i + q] * M[3 * j + q];
                    }
                    B[e] = dRk[3 * j + i];
                    B[9 + e] = -Nij * h;
                    B[18 + e] = -0.5 * Nij * h2;
                    B[27 + e] = Jrf[e] * h;
                    B[36 + e] = -Tij * (0.5 * h2);
                    B[45 + e] = -Tij * (0.25 * h3);
                    B[54 + e] = Rm[e] * h;
                    B[63 + e] = 0.5 * Rm[e] * h2;
                    S.M[e] = M[e];
                    S.MMt[e] = MM;
                    S.Rm[e] = Rm[e];
                }
        }
        WAVE_SYNC();
        // covariance: column c of F P, transposed through LDS, column c of F (F P)^T, symmetrised, plus the noise
        double y[15];
        if (c < 15) {
            pre_fcol(S.B, h, P, y);
            for (int i = 0; i < 15; i++) S.Y[15 * c + i] = y[i];
        }
        WAVE_SYNC();
        if (c < 15) {
            double z[15];
            for (int i = 0; i < 15; i++) z[i] = S.Y[15 * i + c];  // row c of F P
            pre_fcol(S.B, h, z, y);
        }
        WAVE_SYNC();
        if (c < 15)
            for (int i = 0; i < 15; i++) S.Y[15 * c + i] = y[i];
        WAVE_SYNC();
        if (c < 15) {
            const double sg2s = sg2 * sw, sa2s = sa2 * sw;
            const double h4 = h3 * h, h5 = h4 * h, h6 = h5 * h, h7 = h6 * h;
            double q[15];
            q[0] = sg2s * h + swg2 * h3 / 3.0;
            q[1] = sg2s * h2 / 2.0 + swg2 * h4 / 8.0;
            q[2] = sg2s * h3 / 6.0 + swg2 * h5 / 30.0;
            q[3] = sa2s * h + swa2 * h3 / 3.0;
            q[4] = sg2s * h3 / 3.0 + swg2 * h5 / 20.0;
            q[5] = sa2s * h2 / 2.0 + swa2 * h4 / 8.0;
            q[6] = sg2s * h4 / 8.0 + swg2 * h6 / 72.0;
            q[7] = sa2s * h3 / 3.0 + swa2 * h5 / 20.0;
            q[8] = sg2s * h5 / 20.0 + swg2 * h7 / 252.0;
            q[9] = swg2;
            q[10] = swa2;
            q[11] = h;
            q[12] = h2;
            q[13] = h3;
            q[14] = h4;
            for (int i = 0; i < 15; i++) {
                const double v = (i == c) ? y[i] : 0.5 * (y[i] + S.Y[15 * i + c]);
                P[i] = v + pre_qe(i, c, q, S);
            }
        }
        // bias Jacobians: J <- A9 J - B (lanes 0-5, column c)
        if (c < 6) {
            const double *Rt = S.B, *F1 = S.B + 9, *F2 = S.B + 18;
            double jn[9];
            for (int i = 0; i < 3; i++) {
                double s0 = 0.0, s1 = 0.0, s2 = 0.0;
                for (int q = 0; q < 3; q++) {
                    s0 += Rt[3 * i + q] * J[q];
                    s1 += F1[3 * i + q] * J[q];
                    s2 += F2[3 * i + q] * J[q];
                }
                jn[i] = s0;
                jn[3 + i] = s1 + J[3 + i];
                jn[6 + i] = s2 + h * J[3 + i] + J[6 + i];
            }
            const int cc = c % 3;
            for (int i = 0; i < 3; i++) {
                if (c < 3) {
                    jn[i] -= S.B[27 + 3 * i + cc];
                    jn[3 + i] -= S.B[36 + 3 * i + cc];
                    jn[6 + i] -= S.B[45 + 3 * i + cc];
                } else {
                    jn[3 + i] -= S.B[54 + 3 * i + cc];
                    jn[6 + i] -= S.B[63 + 3 * i + cc];
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
        WAVE_SYNC();  // the blocks and Y are rewritten by the next sub-interval
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
    double *Js = S.B;  // the blocks are no longer needed
    if (c < 6)
        for (int i = 0; i < 9; i++) Js[6 * i + c] = J[i];
    WAVE_SYNC();
    if (c == 0) {
        const double kappa = pr[PP_KAPPA], sbg2 = pr[PP_SBG2], sba2 = pr[PP_SBA2];
        double *A = S.Y;
        for (int i = 0; i < 225; i++) A[i] = kappa * A[i];
        if (sbg2 > 0.0 || sba2 > 0.0)
            for (int i = 0; i < 9; i++)
                for (int j = 0; j < 9; j++) {
                    double s = 0.0;
                    for (int q = 0; q < 3; q++)
                        s += sbg2 * Js[6 * i + q] * Js[6 * j + q] + sba2 * Js[6 * i + 3 + q] * Js[6 * j + 3 + q];
                    A[15 * i + j] += s;
                }
        // lower Cholesky in place (below and on the diagonal of A)
        bool ok = true;
        for (int j = 0; j < 15 && ok; j++) {
            double s = A[15 * j + j];
            for (int q = 0; q < j; q++) s -= A[15 * j + q] * A[15 * j + q];
            if (!(s > 0.0)) {
                ok = false;
                break;
            }
            const double dj = sqrt(s);
            A[15 * j + j] = dj;
            for (int i = j + 1; i < 15; i++) {
                s = A[15 * i + j];
                for (int q = 0; q < j; q++) s -= A[15 * i + q] * A[15 * j + q];
                A[15 * i + j] = s / dj;
            }
        }
        // W = L^-1, column by column (W's column j depends on L only), into the upper triangle's mirror: W is
        // written to the output directly
        double *Wd = oW ? oW + 225 * (size_t)f : nullptr;
        double *rec = orec ? orec + (size_t)XBA_IMU_REC * f : nullptr;
        double *Wr = rec ? rec + 76 : nullptr;
        for (int i = 0; i < 225; i++) {
            if (Wd) Wd[i] = 0.0;
            if (Wr) Wr[i] = 0.0;
        }
        if (ok) {
            for (int j = 0; j < 15; j++) {
                double col[15];
                col[j] = 1.0 / A[15 * j + j];
                for (int i = j + 1; i < 15; i++) {
                    double s = 0.0;
                    for (int q = j; q < i; q++) s += A[15 * i + q] * col[q];
                    col[i] = -s / A[15 * i + i];
                }
                for (int i = j; i < 15; i++) {
                    if (Wd) Wd[15 * i + j] = col[i];
                    if (Wr) Wr[15 * i + j] = col[i];
                }
            }
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

}  // namespace xba
