/*
 * xba_oracle.c -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C, single-threaded, FP64 restatement of the reference's bundle-adjustment solve, used as the parity
 * checker for the HIP path and as the CPU baseline that bench.py times.  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load this library; the product path (xclean_amd/) never does.
 *
 * What it restates (reference file:line; paths relative to the reference checkout's src/xclean/):
 *   - vio/window.py:43-61    so3_exp_nb                -> o_so3_exp_nb
 *   - vio/window.py:64-105   so3_log_nb                -> o_so3_log_nb
 *   - vio/window.py:110-122  _cam_poses                -> cam_poses
 *   - vio/window.py:125-146  obs_residual              -> obs_residual
 *   - vio/window.py:164-263  obs_jacobians_ws          -> obs_jacobians
 *   - vio/window.py:284-288  _huber                    -> huber
 *   - vio/window.py:291-463  build_system              -> ow_build
 *   - vio/window.py:466-493  visual_cost               -> ow_visual_cost
 *   - vio/window.py:496-519  back_substitute           -> back_substitute
 *   - vio/window.py:522-540  apply_update              -> apply_update
 *   - vio/window.py:550-765  _pf_block_fast, posefree_cost, posefree_step -> pf_block, pf_cost, pf_step
 *   - vio/window.py:831-886  _imu_terms, _full_system  -> imu_terms, ow_full_system
 *   - vio/window.py:889-982  solve                     -> ow_solve
 *   - vio/window.py:985-1008 final_linearisation       -> ow_final_linearisation
 *   - vio/imu_factor.py:195-330 _abc, _exp_jr, _jr_inv, _log -> abc, exp_jr, jr_inv, log_quat
 *   - vio/imu_factor.py:771-875 factor_core            -> o_imu_factor
 *   - vio/imu_factor.py:366-542 _F_left, _step         -> pre_fleft, pre_step
 *   - vio/imu_factor.py:545-704 preint_core            -> o_preint
 *   - vio/imu_factor.py:707-746 sqrt_info_core         -> o_sqrt_info
 *   - vio/marg.py:47-57, 102-127 _boxminus, Prior._d/terms/cost -> prior_d, prior_terms
 *   - core/lie.py:47-79, 108-141, 201-242, 301-335     series thresholds, coefficients, so3_project, se3_exp, se3_inv
 *   - backend/ba.py:644-697  _rot_factor_blocks        -> rot_factor_blocks
 *   - backend/ba.py:713-869  _solve_np                 -> onp_solve
 *
 * Dense linear algebra the reference delegates to numpy/LAPACK (cholesky, solve, inv, svd) is restated with
 * textbook algorithms (Cholesky-Banachiewicz, Gaussian elimination with partial pivoting, polar iteration); the
 * tests pin each against numpy on the same matrices.
 */
#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define IMU_REC 301 /* dR 9, dv 3, dp 3, dt 1, J 54, bg_lin 3, ba_lin 3, W 225 */

/* ------------------------------------------------------------------------------------------------- problem -- */
typedef struct {
    int32_t P, L, M;
    double *Rwb, *pwb;          /* (P,3,3) (P,3), updated in place on accept */
    const double *Rbc, *pbc;    /* (3,3) (3) */
    double *lmp;                /* (L,3) (a, b, rho) */
    const int32_t *lm_host;     /* (L) */
    const uint8_t *lm_mask;     /* (L) */
    const double *rho_meas, *rho_isig;
    const int32_t *ob_start;    /* (L+1) */
    const int32_t *ob_pose;     /* (M) */
    const uint8_t *ob_cam, *ob_valid;
    const double *ob_z;         /* (M,2) */
    const double *ob_isig;      /* (M) */
    double fx, fy, cx, cy, B, huber_k;
    const int32_t *dims;        /* (P) 6 or 15; NULL: all 6 */
    double *vel, *bgs, *bas;    /* (P,3) or NULL */
    const double *g_w;          /* (3) */
    int32_t n_imu;
    const int32_t *imu_ij;      /* (n_imu,2) */
    const double *imu_rec;      /* (n_imu,IMU_REC) */
    double imu_k;               /* huber_k_dof(15) */
    int32_t prior_nb, prior_m;
    const int32_t *prior_slot, *prior_dim;
    const double *prior_x0;     /* (nb,21): R0 9, p0 3, v0 3, bg0 3, ba0 3 */
    const double *prior_S;      /* (m, sum prior dims) */
    const double *prior_r;      /* (m) */
    int32_t n_pf;
    const int32_t *pf_host;
    const double *pf_z;         /* (n_pf,2) */
    double pf_isg;
    const double *pf_rmeas, *pf_ris;
    double *pf_lmp;             /* (n_pf,3) */
    /* optional per-observation weight multiplying the robust weight and cost of each visual row (NULL: 1).  Not
     * in the reference: it is the input the adjoint differentiates with respect to. */
    const double *ob_w;         /* (M) */
} ow_problem;

/* ------------------------------------------------------------------------------------------- small helpers -- */
static void mm3(const double *A, const double *B, double *C) {
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) C[3 * i + j] = A[3 * i] * B[j] + A[3 * i + 1] * B[3 + j] + A[3 * i + 2] * B[6 + j];
}
static void mtm3(const double *A, const double *B, double *C) { /* C = A^T B */
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) C[3 * i + j] = A[i] * B[j] + A[3 + i] * B[3 + j] + A[6 + i] * B[6 + j];
}
static void mmt3(const double *A, const double *B, double *C) { /* C = A B^T */
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
            C[3 * i + j] = A[3 * i] * B[3 * j] + A[3 * i + 1] * B[3 * j + 1] + A[3 * i + 2] * B[3 * j + 2];
}

/* core/lie.py:47-79 -- derived series thresholds */
static double taylor_angle(void) { return pow(7257600.0 * DBL_EPSILON, 0.1); }
static double angle_zero(void) { return sqrt(6.0 * DBL_EPSILON); }

/* vio/window.py:43-61 */
void o_so3_exp_nb(const double *w, double *R) {
    double th2 = w[0] * w[0] + w[1] * w[1] + w[2] * w[2], th = sqrt(th2), a, b;
    double K[9] = {0, -w[2], w[1], w[2], 0, -w[0], -w[1], w[0], 0}, KK[9];
    if (th < 1e-8) {
        a = 1.0 - th2 / 6.0;
        b = 0.5 - th2 / 24.0;
    } else {
        a = sin(th) / th;
        b = (1.0 - cos(th)) / th2;
    }
    mm3(K, K, KK);
    for (int i = 0; i < 9; i++) R[i] = ((i % 4 == 0) ? 1.0 : 0.0) + a * K[i] + b * KK[i];
}

/* vio/window.py:64-105 */
void o_so3_log_nb(const double *R, double *w) {
    double c = (R[0] + R[4] + R[8] - 1.0) * 0.5;
    if (c > 1.0) c = 1.0;
    if (c < -1.0) c = -1.0;
    double th = acos(c);
    double v0 = R[7] - R[5], v1 = R[2] - R[6], v2 = R[3] - R[1];
    if (th < 1e-6) {
        w[0] = 0.5 * v0;
        w[1] = 0.5 * v1;
        w[2] = 0.5 * v2;
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
    w[0] = f * v0;
    w[1] = f * v1;
    w[2] = f * v2;
}

/* vio/imu_factor.py:195-211 (core/lie.py _coeff_a/b/c with the derived switch) */
static void abc(double t2, double *a, double *b, double *c) {
    double t = sqrt(t2);
    if (t < taylor_angle()) {
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

/* vio/imu_factor.py:214-241: R = Exp(v), Jr = right Jacobian */
void o_exp_jr(const double *v, double *R, double *Jr) {
    double x = v[0], y = v[1], z = v[2], a, b, c;
    abc(x * x + y * y + z * z, &a, &b, &c);
    double xy = x * y, xz = x * z, yz = y * z;
    double k00 = -(y * y + z * z), k11 = -(x * x + z * z), k22 = -(x * x + y * y);
    R[0] = 1.0 + b * k00; R[1] = -a * z + b * xy; R[2] = a * y + b * xz;
    R[3] = a * z + b * xy; R[4] = 1.0 + b * k11; R[5] = -a * x + b * yz;
    R[6] = -a * y + b * xz; R[7] = a * x + b * yz; R[8] = 1.0 + b * k22;
    if (Jr) {
        Jr[0] = 1.0 + c * k00; Jr[1] = b * z + c * xy; Jr[2] = -b * y + c * xz;
        Jr[3] = -b * z + c * xy; Jr[4] = 1.0 + c * k11; Jr[5] = b * x + c * yz;
        Jr[6] = b * y + c * xz; Jr[7] = -b * x + c * yz; Jr[8] = 1.0 + c * k22;
    }
}

/* vio/imu_factor.py:244-263 (core/lie.py so3_jr_inv) */
void o_jr_inv(const double *v, double *out) {
    double x = v[0], y = v[1], z = v[2], t2 = x * x + y * y + z * z, t = sqrt(t2), co;
    if (t < taylor_angle())
        co = 1.0 / 12.0 + t2 / 720.0 + t2 * t2 / 30240.0 + t2 * t2 * t2 / 1209600.0;
    else
        co = (1.0 - (t * sin(t)) / (2.0 * (1.0 - cos(t)))) / t2;
    out[0] = 1.0 - co * (y * y + z * z); out[1] = -0.5 * z + co * x * y; out[2] = 0.5 * y + co * x * z;
    out[3] = 0.5 * z + co * x * y; out[4] = 1.0 - co * (x * x + z * z); out[5] = -0.5 * x + co * y * z;
    out[6] = -0.5 * y + co * x * z; out[7] = 0.5 * x + co * y * z; out[8] = 1.0 - co * (x * x + y * y);
}

/* vio/imu_factor.py:266-330 (core/lie.py so3_log through Shepperd's quaternion) */
void o_log_quat(const double *R, double *out) {
    double m00 = R[0], m11 = R[4], m22 = R[8], tr = m00 + m11 + m22, qw;
    int br = 0;
    if (!(tr > 0.0)) {
        double best = tr;
        br = 0;
        if (m00 > best) { br = 1; best = m00; }
        if (m11 > best) { br = 2; best = m11; }
        if (m22 > best) { br = 3; best = m22; }
        if (br == 0) br = 1;
    }
    if (br == 0) {
        double s = sqrt(fmax(tr + 1.0, 0.0)) * 2.0;
        qw = 0.25 * s;
        out[0] = (R[7] - R[5]) / s;
        out[1] = (R[2] - R[6]) / s;
        out[2] = (R[3] - R[1]) / s;
    } else {
        int a, c, d;
        if (br == 1) { a = 0; c = 1; d = 2; }
        else if (br == 2) { a = 1; c = 2; d = 0; }
        else { a = 2; c = 0; d = 1; }
        double s = sqrt(fmax(1.0 + R[4 * a] - R[4 * c] - R[4 * d], 0.0)) * 2.0;
        qw = (R[3 * d + c] - R[3 * c + d]) / s;
        out[a] = 0.25 * s;
        out[c] = (R[3 * c + a] + R[3 * a + c]) / s;
        out[d] = (R[3 * d + a] + R[3 * a + d]) / s;
    }
    if (qw < 0.0) {
        qw = -qw;
        out[0] = -out[0]; out[1] = -out[1]; out[2] = -out[2];
    }
    double n = sqrt(out[0] * out[0] + out[1] * out[1] + out[2] * out[2]);
    double sc = (n < angle_zero()) ? 2.0 : 2.0 * atan2(n, qw) / n;
    out[0] *= sc; out[1] *= sc; out[2] *= sc;
}

/* core/lie.py:301-314 se3_exp: xi = [rho, phi] -> 4x4 (row-major) */
void o_se3_exp(const double *xi, double *T) {
    const double *rho = xi, *phi = xi + 3;
    double a, b, c, K[9] = {0, -phi[2], phi[1], phi[2], 0, -phi[0], -phi[1], phi[0], 0}, KK[9];
    abc(phi[0] * phi[0] + phi[1] * phi[1] + phi[2] * phi[2], &a, &b, &c);
    mm3(K, K, KK);
    double R[9], V[9];
    for (int i = 0; i < 9; i++) {
        double e = (i % 4 == 0) ? 1.0 : 0.0;
        R[i] = e + a * K[i] + b * KK[i];
        V[i] = e + b * K[i] + c * KK[i];
    }
    memset(T, 0, 16 * sizeof(double));
    for (int i = 0; i < 3; i++) {
        for (int j = 0; j < 3; j++) T[4 * i + j] = R[3 * i + j];
        T[4 * i + 3] = V[3 * i] * rho[0] + V[3 * i + 1] * rho[1] + V[3 * i + 2] * rho[2];
    }
    T[15] = 1.0;
}

/* core/lie.py:328-335 se3_inv */
static void se3_inv(const double *T, double *out) {
    memset(out, 0, 16 * sizeof(double));
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) out[4 * i + j] = T[4 * j + i];
    for (int i = 0; i < 3; i++) out[4 * i + 3] = -(T[i] * T[3] + T[4 + i] * T[7] + T[8 + i] * T[11]);
    out[15] = 1.0;
}

static int inv3(const double *A, double *X) { /* cofactor inverse; 0 when singular */
    double c00 = A[4] * A[8] - A[5] * A[7], c01 = A[5] * A[6] - A[3] * A[8], c02 = A[3] * A[7] - A[4] * A[6];
    double det = A[0] * c00 + A[1] * c01 + A[2] * c02;
    if (det == 0.0 || !isfinite(det)) return 0;
    double id = 1.0 / det;
    X[0] = c00 * id; X[1] = (A[2] * A[7] - A[1] * A[8]) * id; X[2] = (A[1] * A[5] - A[2] * A[4]) * id;
    X[3] = c01 * id; X[4] = (A[0] * A[8] - A[2] * A[6]) * id; X[5] = (A[2] * A[3] - A[0] * A[5]) * id;
    X[6] = c02 * id; X[7] = (A[1] * A[6] - A[0] * A[7]) * id; X[8] = (A[0] * A[4] - A[1] * A[3]) * id;
    return 1;
}

/* core/lie.py:228-242 so3_project: the nearest rotation in Frobenius norm (= the orthogonal polar factor when
 * det > 0), restated as Newton's polar iteration X <- (X + X^-T)/2 instead of an SVD. */
void o_so3_project(const double *R, double *out) {
    double X[9], Xi[9];
    memcpy(X, R, sizeof(X));
    for (int it = 0; it < 30; it++) {
        if (!inv3(X, Xi)) break;
        double d = 0.0;
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) {
                double nv = 0.5 * (X[3 * i + j] + Xi[3 * j + i]);
                d = fmax(d, fabs(nv - X[3 * i + j]));
                X[3 * i + j] = nv;
            }
        if (d < 1e-16) break;
    }
    memcpy(out, X, sizeof(X));
}

/* ---------------------------------------------------------------------------------------- the visual factor -- */
/* vio/window.py:110-122: Rwc = Rwb Rbc, pwc = pwb + Rwb pbc */
static void cam_poses(const ow_problem *pb, const double *Rwb, const double *pwb, double *Rwc, double *pwc) {
    for (int s = 0; s < pb->P; s++) {
        for (int i = 0; i < 3; i++) {
            for (int j = 0; j < 3; j++) {
                double acc = 0.0;
                for (int k = 0; k < 3; k++) acc += Rwb[9 * s + 3 * i + k] * pb->Rbc[3 * k + j];
                Rwc[9 * s + 3 * i + j] = acc;
            }
            double acc = pwb[3 * s + i];
            for (int k = 0; k < 3; k++) acc += Rwb[9 * s + 3 * i + k] * pb->pbc[k];
            pwc[3 * s + i] = acc;
        }
    }
}

/* vio/window.py:125-146; returns nd, writes q2 */
static int obs_residual(const double *Rwc, const double *pwc, const double *lm, int h, int t, int cam, const double *z,
                        double fx, double fy, double cx, double cy, double B, double *e, double *q2o) {
    double a = lm[0], b = lm[1], rho = lm[2];
    const double *Rh = Rwc + 9 * h, *Rt = Rwc + 9 * t;
    double Y0 = Rh[0] * a + Rh[1] * b + Rh[2] + rho * (pwc[3 * h] - pwc[3 * t]);
    double Y1 = Rh[3] * a + Rh[4] * b + Rh[5] + rho * (pwc[3 * h + 1] - pwc[3 * t + 1]);
    double Y2 = Rh[6] * a + Rh[7] * b + Rh[8] + rho * (pwc[3 * h + 2] - pwc[3 * t + 2]);
    double q0 = Rt[0] * Y0 + Rt[3] * Y1 + Rt[6] * Y2;
    double q1 = Rt[1] * Y0 + Rt[4] * Y1 + Rt[7] * Y2;
    double q2 = Rt[2] * Y0 + Rt[5] * Y1 + Rt[8] * Y2;
    *q2o = q2;
    if (cam == 0) {
        if (q2 <= 1e-9) return 2;
        e[0] = fx * q0 / q2 + cx - z[0];
        e[1] = fy * q1 / q2 + cy - z[1];
        return 2;
    }
    if (q2 <= 1e-9) return 1;
    e[0] = fx * (q0 - rho * B) / q2 + cx - z[0];
    return 1;
}

/* vio/window.py:164-263: Jacobians of the unwhitened residual wrt (a, b, rho), the target and the host body
 * pose blocks [dphi, dp].  Rcb = Rbc^T, Px = [pbc]x. */
static int obs_jacobians(const double *Rwb, const double *Rwc, const double *pwc, const double *Rcb, const double *Px,
                         const double *lm, int h, int t, int cam, double fx, double fy, double B, double Jl[2][3],
                         double Jt[2][6], double Jh[2][6]) {
    double a = lm[0], b = lm[1], rho = lm[2];
    const double *Rh = Rwc + 9 * h, *Rt = Rwc + 9 * t;
    double d0 = pwc[3 * h] - pwc[3 * t], d1 = pwc[3 * h + 1] - pwc[3 * t + 1], d2 = pwc[3 * h + 2] - pwc[3 * t + 2];
    double Y0 = Rh[0] * a + Rh[1] * b + Rh[2] + rho * d0;
    double Y1 = Rh[3] * a + Rh[4] * b + Rh[5] + rho * d1;
    double Y2 = Rh[6] * a + Rh[7] * b + Rh[8] + rho * d2;
    double q0 = Rt[0] * Y0 + Rt[3] * Y1 + Rt[6] * Y2;
    double q1 = Rt[1] * Y0 + Rt[4] * Y1 + Rt[7] * Y2;
    double q2 = Rt[2] * Y0 + Rt[5] * Y1 + Rt[8] * Y2;
    int nd = cam == 0 ? 2 : 1;
    double iq = 1.0 / q2, Dp[2][3];
    if (cam == 0) {
        Dp[0][0] = fx * iq; Dp[0][1] = 0.0; Dp[0][2] = -fx * q0 * iq * iq;
        Dp[1][0] = 0.0; Dp[1][1] = fy * iq; Dp[1][2] = -fy * q1 * iq * iq;
    } else {
        Dp[0][0] = fx * iq; Dp[0][1] = 0.0; Dp[0][2] = -fx * (q0 - rho * B) * iq * iq;
    }
    double RcR[9];
    mtm3(Rt, Rh, RcR);
    double r0 = Rt[0] * d0 + Rt[3] * d1 + Rt[6] * d2;
    double r1 = Rt[1] * d0 + Rt[4] * d1 + Rt[7] * d2;
    double r2 = Rt[2] * d0 + Rt[5] * d1 + Rt[8] * d2;
    if (cam == 1) r0 -= B;
    for (int r = 0; r < nd; r++) {
        Jl[r][0] = Dp[r][0] * RcR[0] + Dp[r][1] * RcR[3] + Dp[r][2] * RcR[6];
        Jl[r][1] = Dp[r][0] * RcR[1] + Dp[r][1] * RcR[4] + Dp[r][2] * RcR[7];
        Jl[r][2] = Dp[r][0] * r0 + Dp[r][1] * r1 + Dp[r][2] * r2;
    }
    if (t == h) {
        for (int r = 0; r < nd; r++)
            for (int j = 0; j < 6; j++) Jt[r][j] = Jh[r][j] = 0.0;
        return nd;
    }
    double Qx[9] = {0.0, -q2, q1, q2, 0.0, -q0, -q1, q0, 0.0};
    double Mx[9] = {0.0, -1.0, b, 1.0, 0.0, -a, -b, a, 0.0};
    double Ah[9], Rct[9], Wt[9], Wh[9], T1[9], T2[9], T3[9];
    mm3(RcR, Mx, Ah);
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) Rct[3 * i + j] = Rt[3 * j + i];
    mm3(Rwb + 9 * t, Px, Wt);
    mm3(Rwb + 9 * h, Px, Wh);
    mm3(Qx, Rcb, T1);
    mm3(Rct, Wt, T2);
    for (int i = 0; i < 9; i++) T1[i] = T1[i] + rho * T2[i];
    mm3(Ah, Rcb, T2);
    mm3(Rct, Wh, T3);
    for (int i = 0; i < 9; i++) T2[i] = -T2[i] - rho * T3[i];
    for (int r = 0; r < nd; r++)
        for (int j = 0; j < 3; j++) {
            Jt[r][j] = Dp[r][0] * T1[j] + Dp[r][1] * T1[3 + j] + Dp[r][2] * T1[6 + j];
            Jh[r][j] = Dp[r][0] * T2[j] + Dp[r][1] * T2[3 + j] + Dp[r][2] * T2[6 + j];
            double v = Dp[r][0] * Rct[j] + Dp[r][1] * Rct[3 + j] + Dp[r][2] * Rct[6 + j];
            Jt[r][3 + j] = -rho * v;
            Jh[r][3 + j] = rho * v;
        }
    return nd;
}

/* exported for the finite-difference tests */
int o_obs_residual_jac(const ow_problem *pb, int l, int k, double *e, double *Jl, double *Jt, double *Jh) {
    int P = pb->P;
    double *Rwc = malloc(9 * P * sizeof(double)), *pwc = malloc(3 * P * sizeof(double));
    cam_poses(pb, pb->Rwb, pb->pwb, Rwc, pwc);
    double Rcb[9], Px[9] = {0, -pb->pbc[2], pb->pbc[1], pb->pbc[2], 0, -pb->pbc[0], -pb->pbc[1], pb->pbc[0], 0};
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) Rcb[3 * i + j] = pb->Rbc[3 * j + i];
    double q2, Jl_[2][3] = {{0}}, Jt_[2][6] = {{0}}, Jh_[2][6] = {{0}};
    int nd = obs_residual(Rwc, pwc, pb->lmp + 3 * l, pb->lm_host[l], pb->ob_pose[k], pb->ob_cam[k], pb->ob_z + 2 * k,
                          pb->fx, pb->fy, pb->cx, pb->cy, pb->B, e, &q2);
    obs_jacobians(pb->Rwb, Rwc, pwc, Rcb, Px, pb->lmp + 3 * l, pb->lm_host[l], pb->ob_pose[k], pb->ob_cam[k], pb->fx,
                  pb->fy, pb->B, Jl_, Jt_, Jh_);
    for (int r = 0; r < 2; r++) {
        for (int j = 0; j < 3; j++) Jl[3 * r + j] = Jl_[r][j];
        for (int j = 0; j < 6; j++) { Jt[6 * r + j] = Jt_[r][j]; Jh[6 * r + j] = Jh_[r][j]; }
    }
    free(Rwc);
    free(pwc);
    return nd;
}

/* vio/window.py:284-288 */
static void huber(double en, double k, double *w, double *c) {
    if (en <= k) { *w = 1.0; *c = 0.5 * en * en; }
    else { *w = k / en; *c = k * (en - 0.5 * k); }
}

static void make_px(const ow_problem *pb, double *Rcb, double *Px) {
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) Rcb[3 * i + j] = pb->Rbc[3 * j + i];
    const double *p = pb->pbc;
    double P_[9] = {0, -p[2], p[1], p[2], 0, -p[0], -p[1], p[0], 0};
    memcpy(Px, P_, sizeof(P_));
}

/* vio/window.py:291-463 build_system.  H (6P x 6P), bvec (6P), Hll_inv (L,9), bl_out (L,3), Hlp (L,P,18),
 * lm_ok (L), ob_chi2 (M).  Returns the robust cost. */
double ow_build(const ow_problem *pb, const double *Rwb, const double *pwb, const double *lmp, double lam, double *H,
                double *bvec, double *Hll_inv, double *bl_out, double *Hlp, uint8_t *lm_ok, double *ob_chi2) {
    int P = pb->P, L = pb->L, n6 = 6 * P;
    double *Rwc = malloc(9 * P * sizeof(double)), *pwc = malloc(3 * P * sizeof(double));
    uint8_t *touched = malloc(P);
    cam_poses(pb, Rwb, pwb, Rwc, pwc);
    memset(H, 0, (size_t)n6 * n6 * sizeof(double));
    memset(bvec, 0, n6 * sizeof(double));
    double Rcb[9], Px[9];
    make_px(pb, Rcb, Px);
    double cost = 0.0;
    for (int l = 0; l < L; l++) {
        lm_ok[l] = 0;
        if (!pb->lm_mask[l]) continue;
        int h = pb->lm_host[l];
        double Hll[9] = {0}, bl[3] = {0};
        double *Hl = Hlp + (size_t)l * P * 18;
        for (int s = 0; s < P; s++) {
            touched[s] = 0;
            for (int i = 0; i < 18; i++) Hl[18 * s + i] = 0.0;
        }
        int nobs = 0;
        for (int k = pb->ob_start[l]; k < pb->ob_start[l + 1]; k++) {
            ob_chi2[k] = INFINITY;
            if (!pb->ob_valid[k]) continue;
            int t = pb->ob_pose[k], cam = pb->ob_cam[k];
            double e[2], q2, Jl[2][3], Jt[2][6], Jh[2][6];
            int nd = obs_residual(Rwc, pwc, lmp + 3 * l, h, t, cam, pb->ob_z + 2 * k, pb->fx, pb->fy, pb->cx, pb->cy,
                                  pb->B, e, &q2);
            if (q2 <= 1e-9) continue;
            double isg = pb->ob_isig[k], en2 = 0.0;
            for (int i = 0; i < nd; i++) {
                e[i] *= isg;
                en2 += e[i] * e[i];
            }
            ob_chi2[k] = en2;
            double w, c;
            huber(sqrt(en2), pb->huber_k, &w, &c);
            if (pb->ob_w) {
                w *= pb->ob_w[k];
                c *= pb->ob_w[k];
            }
            cost += c;
            obs_jacobians(Rwb, Rwc, pwc, Rcb, Px, lmp + 3 * l, h, t, cam, pb->fx, pb->fy, pb->B, Jl, Jt, Jh);
            for (int i = 0; i < nd; i++) {
                for (int j = 0; j < 3; j++) Jl[i][j] *= isg;
                for (int j = 0; j < 6; j++) { Jt[i][j] *= isg; Jh[i][j] *= isg; }
            }
            nobs++;
            for (int i = 0; i < 3; i++) {
                for (int j = 0; j < 3; j++) {
                    double acc = 0.0;
                    for (int r = 0; r < nd; r++) acc += Jl[r][i] * Jl[r][j];
                    Hll[3 * i + j] += w * acc;
                }
                double acc = 0.0;
                for (int r = 0; r < nd; r++) acc += Jl[r][i] * e[r];
                bl[i] += w * acc;
            }
            if (t == h) continue;
            touched[t] = 1;
            touched[h] = 1;
            for (int i = 0; i < 3; i++)
                for (int j = 0; j < 6; j++) {
                    double at = 0.0, ah = 0.0;
                    for (int r = 0; r < nd; r++) {
                        at += Jl[r][i] * Jt[r][j];
                        ah += Jl[r][i] * Jh[r][j];
                    }
                    Hl[18 * t + 6 * i + j] += w * at;
                    Hl[18 * h + 6 * i + j] += w * ah;
                }
            for (int i = 0; i < 6; i++) {
                double bt = 0.0, bh = 0.0;
                for (int r = 0; r < nd; r++) {
                    bt += Jt[r][i] * e[r];
                    bh += Jh[r][i] * e[r];
                }
                bvec[6 * t + i] += w * bt;
                bvec[6 * h + i] += w * bh;
                for (int j = 0; j < 6; j++) {
                    double tt = 0.0, hh = 0.0, th = 0.0;
                    for (int r = 0; r < nd; r++) {
                        tt += Jt[r][i] * Jt[r][j];
                        hh += Jh[r][i] * Jh[r][j];
                        th += Jt[r][i] * Jh[r][j];
                    }
                    H[(size_t)(6 * t + i) * n6 + 6 * t + j] += w * tt;
                    H[(size_t)(6 * h + i) * n6 + 6 * h + j] += w * hh;
                    H[(size_t)(6 * t + i) * n6 + 6 * h + j] += w * th;
                    H[(size_t)(6 * h + j) * n6 + 6 * t + i] += w * th;
                }
            }
        }
        if (pb->rho_isig[l] > 0.0) {
            double r0 = (lmp[3 * l + 2] - pb->rho_meas[l]) * pb->rho_isig[l], w, c;
            huber(fabs(r0), pb->huber_k, &w, &c);
            cost += c;
            Hll[8] += w * pb->rho_isig[l] * pb->rho_isig[l];
            bl[2] += w * pb->rho_isig[l] * r0;
        }
        for (int i = 0; i < 3; i++) Hll[4 * i] += lam * Hll[4 * i] + 1e-12;
        double det = (Hll[0] * (Hll[4] * Hll[8] - Hll[5] * Hll[7]) - Hll[1] * (Hll[3] * Hll[8] - Hll[5] * Hll[6]) +
                      Hll[2] * (Hll[3] * Hll[7] - Hll[4] * Hll[6]));
        double tr = Hll[0] + Hll[4] + Hll[8];
        if (nobs == 0 || !(det > 1e-18 * tr * tr * tr) || !isfinite(det)) continue;
        double idet = 1.0 / det, *Hi = Hll_inv + 9 * (size_t)l;
        Hi[0] = (Hll[4] * Hll[8] - Hll[5] * Hll[7]) * idet;
        Hi[1] = (Hll[2] * Hll[7] - Hll[1] * Hll[8]) * idet;
        Hi[2] = (Hll[1] * Hll[5] - Hll[2] * Hll[4]) * idet;
        Hi[3] = (Hll[5] * Hll[6] - Hll[3] * Hll[8]) * idet;
        Hi[4] = (Hll[0] * Hll[8] - Hll[2] * Hll[6]) * idet;
        Hi[5] = (Hll[2] * Hll[3] - Hll[0] * Hll[5]) * idet;
        Hi[6] = (Hll[3] * Hll[7] - Hll[4] * Hll[6]) * idet;
        Hi[7] = (Hll[1] * Hll[6] - Hll[0] * Hll[7]) * idet;
        Hi[8] = (Hll[0] * Hll[4] - Hll[1] * Hll[3]) * idet;
        lm_ok[l] = 1;
        for (int i = 0; i < 3; i++) bl_out[3 * l + i] = bl[i];
        double HiB[3];
        for (int i = 0; i < 3; i++) HiB[i] = Hi[3 * i] * bl[0] + Hi[3 * i + 1] * bl[1] + Hi[3 * i + 2] * bl[2];
        for (int s1 = 0; s1 < P; s1++) {
            if (!touched[s1]) continue;
            const double *A1 = Hl + 18 * s1;
            double T1[18];
            for (int i = 0; i < 3; i++)
                for (int j = 0; j < 6; j++)
                    T1[6 * i + j] = Hi[3 * i] * A1[j] + Hi[3 * i + 1] * A1[6 + j] + Hi[3 * i + 2] * A1[12 + j];
            for (int i = 0; i < 6; i++) {
                double acc = 0.0;
                for (int r = 0; r < 3; r++) acc += A1[6 * r + i] * HiB[r];
                bvec[6 * s1 + i] -= acc;
            }
            for (int s2 = 0; s2 < P; s2++) {
                if (!touched[s2]) continue;
                const double *A2 = Hl + 18 * s2;
                for (int i = 0; i < 6; i++)
                    for (int j = 0; j < 6; j++) {
                        double acc = 0.0;
                        for (int r = 0; r < 3; r++) acc += A2[6 * r + i] * T1[6 * r + j];
                        H[(size_t)(6 * s2 + i) * n6 + 6 * s1 + j] -= acc;
                    }
            }
        }
    }
    free(Rwc);
    free(pwc);
    free(touched);
    return cost;
}

/* vio/window.py:466-493 */
double ow_visual_cost(const ow_problem *pb, const double *Rwb, const double *pwb, const double *lmp) {
    int P = pb->P;
    double *Rwc = malloc(9 * P * sizeof(double)), *pwc = malloc(3 * P * sizeof(double));
    cam_poses(pb, Rwb, pwb, Rwc, pwc);
    double cost = 0.0;
    for (int l = 0; l < pb->L; l++) {
        if (!pb->lm_mask[l]) continue;
        int h = pb->lm_host[l];
        for (int k = pb->ob_start[l]; k < pb->ob_start[l + 1]; k++) {
            if (!pb->ob_valid[k]) continue;
            double e[2], q2;
            int nd = obs_residual(Rwc, pwc, lmp + 3 * l, h, pb->ob_pose[k], pb->ob_cam[k], pb->ob_z + 2 * k, pb->fx,
                                  pb->fy, pb->cx, pb->cy, pb->B, e, &q2);
            if (q2 <= 1e-9) {
                cost += 1e6;
                continue;
            }
            double en2 = 0.0, w, c;
            for (int i = 0; i < nd; i++) en2 += (e[i] * pb->ob_isig[k]) * (e[i] * pb->ob_isig[k]);
            huber(sqrt(en2), pb->huber_k, &w, &c);
            cost += pb->ob_w ? c * pb->ob_w[k] : c;
        }
        if (pb->rho_isig[l] > 0.0) {
            double r0 = (lmp[3 * l + 2] - pb->rho_meas[l]) * pb->rho_isig[l], w, c;
            huber(fabs(r0), pb->huber_k, &w, &c);
            cost += c;
        }
    }
    free(Rwc);
    free(pwc);
    return cost;
}

/* vio/window.py:496-519 */
static void back_substitute(int L, int P, const double *dp, const double *Hll_inv, const double *bl, const double *Hlp,
                            const uint8_t *lm_ok, double *dl) {
    for (int l = 0; l < L; l++) {
        dl[3 * l] = dl[3 * l + 1] = dl[3 * l + 2] = 0.0;
        if (!lm_ok[l]) continue;
        double v[3] = {bl[3 * l], bl[3 * l + 1], bl[3 * l + 2]};
        const double *Hl = Hlp + (size_t)l * P * 18;
        for (int s = 0; s < P; s++)
            for (int i = 0; i < 3; i++) {
                double acc = 0.0;
                for (int j = 0; j < 6; j++) acc += Hl[18 * s + 6 * i + j] * dp[6 * s + j];
                v[i] += acc;
            }
        const double *Hi = Hll_inv + 9 * (size_t)l;
        for (int i = 0; i < 3; i++) {
            double acc = 0.0;
            for (int j = 0; j < 3; j++) acc += Hi[3 * i + j] * v[j];
            dl[3 * l + i] = -acc;
        }
    }
}

/* vio/window.py:522-540 */
static void apply_update(const ow_problem *pb, const double *dp, const double *dl, const uint8_t *lm_ok, double *R2,
                         double *p2, double *l2) {
    for (int s = 0; s < pb->P; s++) {
        double w[3] = {dp[6 * s], dp[6 * s + 1], dp[6 * s + 2]}, E[9];
        o_so3_exp_nb(w, E);
        mm3(pb->Rwb + 9 * s, E, R2 + 9 * s);
        for (int i = 0; i < 3; i++) p2[3 * s + i] = pb->pwb[3 * s + i] + dp[6 * s + 3 + i];
    }
    for (int l = 0; l < pb->L; l++) {
        for (int i = 0; i < 3; i++) l2[3 * l + i] = pb->lmp[3 * l + i];
        if (lm_ok[l]) {
            double r = pb->lmp[3 * l + 2] + dl[3 * l + 2];
            if (r > 1e-9 && isfinite(r)) {
                l2[3 * l] = pb->lmp[3 * l] + dl[3 * l];
                l2[3 * l + 1] = pb->lmp[3 * l + 1] + dl[3 * l + 1];
                l2[3 * l + 2] = r;
            }
        }
    }
}

/* ------------------------------------------------------------------------------------ pose-free landmarks -- */
/* vio/window.py:550-672 _pf_block_fast (the host-left observation + inverse-depth prior, t == h) */
static int pf_block(const double *Rwc, const double *pwc, const double *RcR, int h, double a, double b, double rho,
                    double z0, double z1, double isg, double ris, double rmeas, double fx, double fy, double cx,
                    double cy, double huber_k, double lam, double *Hll, double *bl, double *Hi) {
    const double *Rh = Rwc + 9 * h;
    double d0 = pwc[3 * h] - pwc[3 * h], d1 = pwc[3 * h + 1] - pwc[3 * h + 1], d2 = pwc[3 * h + 2] - pwc[3 * h + 2];
    double Y0 = Rh[0] * a + Rh[1] * b + Rh[2] + rho * d0;
    double Y1 = Rh[3] * a + Rh[4] * b + Rh[5] + rho * d1;
    double Y2 = Rh[6] * a + Rh[7] * b + Rh[8] + rho * d2;
    double q0 = Rh[0] * Y0 + Rh[3] * Y1 + Rh[6] * Y2;
    double q1 = Rh[1] * Y0 + Rh[4] * Y1 + Rh[7] * Y2;
    double q2 = Rh[2] * Y0 + Rh[5] * Y1 + Rh[8] * Y2;
    for (int i = 0; i < 9; i++) Hll[i] = 0.0;
    bl[0] = bl[1] = bl[2] = 0.0;
    int nobs = 0;
    if (q2 > 1e-9) {
        double e0 = fx * q0 / q2 + cx - z0, e1 = fy * q1 / q2 + cy - z1, en2 = 0.0;
        e0 *= isg;
        en2 += e0 * e0;
        e1 *= isg;
        en2 += e1 * e1;
        double en = sqrt(en2), w = (en <= huber_k) ? 1.0 : huber_k / en;
        double iq = 1.0 / q2;
        double D00 = fx * iq, D01 = 0.0, D02 = -fx * q0 * iq * iq, D10 = 0.0, D11 = fy * iq, D12 = -fy * q1 * iq * iq;
        double r0 = Rh[0] * d0 + Rh[3] * d1 + Rh[6] * d2;
        double r1 = Rh[1] * d0 + Rh[4] * d1 + Rh[7] * d2;
        double r2 = Rh[2] * d0 + Rh[5] * d1 + Rh[8] * d2;
        double J[2][3];
        J[0][0] = D00 * RcR[0] + D01 * RcR[3] + D02 * RcR[6];
        J[0][1] = D00 * RcR[1] + D01 * RcR[4] + D02 * RcR[7];
        J[0][2] = D00 * r0 + D01 * r1 + D02 * r2;
        J[1][0] = D10 * RcR[0] + D11 * RcR[3] + D12 * RcR[6];
        J[1][1] = D10 * RcR[1] + D11 * RcR[4] + D12 * RcR[7];
        J[1][2] = D10 * r0 + D11 * r1 + D12 * r2;
        for (int r = 0; r < 2; r++)
            for (int j = 0; j < 3; j++) J[r][j] *= isg;
        nobs++;
        double e[2] = {e0, e1};
        for (int i = 0; i < 3; i++) {
            for (int j = 0; j < 3; j++) {
                double acc = 0.0;
                acc += J[0][i] * J[0][j];
                acc += J[1][i] * J[1][j];
                Hll[3 * i + j] += w * acc;
            }
            double acc = 0.0;
            acc += J[0][i] * e[0];
            acc += J[1][i] * e[1];
            bl[i] += w * acc;
        }
    }
    if (ris > 0.0) {
        double rr = (rho - rmeas) * ris, ar = fabs(rr), wp = (ar <= huber_k) ? 1.0 : huber_k / ar;
        Hll[8] += wp * ris * ris;
        bl[2] += wp * ris * rr;
    }
    for (int i = 0; i < 3; i++) Hll[4 * i] += lam * Hll[4 * i] + 1e-12;
    double det = (Hll[0] * (Hll[4] * Hll[8] - Hll[5] * Hll[7]) - Hll[1] * (Hll[3] * Hll[8] - Hll[5] * Hll[6]) +
                  Hll[2] * (Hll[3] * Hll[7] - Hll[4] * Hll[6]));
    double tr = Hll[0] + Hll[4] + Hll[8];
    if (nobs == 0 || !(det > 1e-18 * tr * tr * tr) || !isfinite(det)) return 0;
    double idet = 1.0 / det;
    Hi[0] = (Hll[4] * Hll[8] - Hll[5] * Hll[7]) * idet;
    Hi[1] = (Hll[2] * Hll[7] - Hll[1] * Hll[8]) * idet;
    Hi[2] = (Hll[1] * Hll[5] - Hll[2] * Hll[4]) * idet;
    Hi[3] = (Hll[5] * Hll[6] - Hll[3] * Hll[8]) * idet;
    Hi[4] = (Hll[0] * Hll[8] - Hll[2] * Hll[6]) * idet;
    Hi[5] = (Hll[2] * Hll[3] - Hll[0] * Hll[5]) * idet;
    Hi[6] = (Hll[3] * Hll[7] - Hll[4] * Hll[6]) * idet;
    Hi[7] = (Hll[1] * Hll[6] - Hll[0] * Hll[7]) * idet;
    Hi[8] = (Hll[0] * Hll[4] - Hll[1] * Hll[3]) * idet;
    return 1;
}

/* vio/window.py:675-700 posefree_cost */
static double pf_cost(const ow_problem *pb, const double *Rwb, const double *pwb, const double *lmp, int behind) {
    int P = pb->P;
    double *Rwc = malloc(9 * P * sizeof(double)), *pwc = malloc(3 * P * sizeof(double));
    cam_poses(pb, Rwb, pwb, Rwc, pwc);
    double cost = 0.0;
    for (int l = 0; l < pb->n_pf; l++) {
        int h = pb->pf_host[l];
        double e[2], q2, w, c;
        obs_residual(Rwc, pwc, lmp + 3 * l, h, h, 0, pb->pf_z + 2 * l, pb->fx, pb->fy, pb->cx, pb->cy, pb->B, e, &q2);
        if (q2 <= 1e-9) {
            if (behind) cost += 1e6;
        } else {
            double en2 = 0.0, e0 = e[0] * pb->pf_isg, e1 = e[1] * pb->pf_isg;
            en2 += e0 * e0;
            en2 += e1 * e1;
            huber(sqrt(en2), pb->huber_k, &w, &c);
            cost += c;
        }
        if (pb->pf_ris[l] > 0.0) {
            double r0 = (lmp[3 * l + 2] - pb->pf_rmeas[l]) * pb->pf_ris[l];
            huber(fabs(r0), pb->huber_k, &w, &c);
            cost += c;
        }
    }
    free(Rwc);
    free(pwc);
    return cost;
}

/* vio/window.py:703-765 posefree_step */
static void pf_step(const ow_problem *pb, const double *lmp_in, double lam, double *lmp_out, double *Hinv_out,
                    uint8_t *ok_out) {
    int P = pb->P;
    double *Rwc = malloc(9 * P * sizeof(double)), *pwc = malloc(3 * P * sizeof(double)),
           *RcR = malloc(9 * P * sizeof(double));
    cam_poses(pb, pb->Rwb, pb->pwb, Rwc, pwc);
    for (int s = 0; s < P; s++) mtm3(Rwc + 9 * s, Rwc + 9 * s, RcR + 9 * s);
    for (int l = 0; l < pb->n_pf; l++) {
        int h = pb->pf_host[l];
        double Hll[9], bl[3], Hi[9];
        for (int i = 0; i < 3; i++) lmp_out[3 * l + i] = lmp_in[3 * l + i];
        int ok = pf_block(Rwc, pwc, RcR + 9 * h, h, lmp_in[3 * l], lmp_in[3 * l + 1], lmp_in[3 * l + 2],
                          pb->pf_z[2 * l], pb->pf_z[2 * l + 1], pb->pf_isg, pb->pf_ris[l], pb->pf_rmeas[l], pb->fx,
                          pb->fy, pb->cx, pb->cy, pb->huber_k, lam, Hll, bl, Hi);
        ok_out[l] = (uint8_t)ok;
        if (!ok) continue;
        if (Hinv_out)
            for (int i = 0; i < 9; i++) Hinv_out[9 * l + i] = Hi[i];
        double v[3] = {bl[0] + 0.0, bl[1] + 0.0, bl[2] + 0.0}, d[3];
        for (int i = 0; i < 3; i++) {
            double acc = 0.0;
            acc += Hi[3 * i] * v[0];
            acc += Hi[3 * i + 1] * v[1];
            acc += Hi[3 * i + 2] * v[2];
            d[i] = -acc;
        }
        double r = lmp_in[3 * l + 2] + d[2];
        if (r > 1e-9 && isfinite(r)) {
            lmp_out[3 * l] = lmp_in[3 * l] + d[0];
            lmp_out[3 * l + 1] = lmp_in[3 * l + 1] + d[1];
            lmp_out[3 * l + 2] = r;
        }
    }
    free(Rwc);
    free(pwc);
    free(RcR);
}

/* ------------------------------------------------------------------------------------------- IMU factor -- */
/* vio/imu_factor.py:771-875 factor_core (whitened by the lower-triangular W when whiten) */
void o_imu_factor(const double *Ri, const double *pi, const double *vi, const double *bgi, const double *bai,
                  const double *Rj, const double *pj, const double *vj, const double *bgj, const double *baj,
                  const double *g, const double *rec, int whiten, double *r, double *Ji, double *Jj, double *Jg) {
    const double *dR = rec, *dv = rec + 9, *dp = rec + 12, dt = rec[15], *J = rec + 16, *bg_lin = rec + 70,
                 *ba_lin = rec + 73, *W = rec + 76;
    double dbg[3], dba[3], phib[3], rR[3], uv[3], up[3], Dv[3], Dp[3];
    double Eb[9], Jrb[9], dRc[9], Mij[9], E[9], Jri[9], T1[9], T2[9], T3[9], T4[9];
    for (int c = 0; c < 3; c++) {
        dbg[c] = bgi[c] - bg_lin[c];
        dba[c] = bai[c] - ba_lin[c];
    }
    for (int i = 0; i < 3; i++) phib[i] = J[6 * i] * dbg[0] + J[6 * i + 1] * dbg[1] + J[6 * i + 2] * dbg[2];
    o_exp_jr(phib, Eb, Jrb);
    mm3(dR, Eb, dRc);
    mtm3(Ri, Rj, Mij);
    mtm3(dRc, Mij, E);
    o_log_quat(E, rR);
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
    o_jr_inv(rR, Jri);
    mmt3(Jri, Mij, T1);
    mmt3(Jri, E, T2);
    mm3(T2, Jrb, T3);
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) T4[3 * i + j] = T3[3 * i] * J[j] + T3[3 * i + 1] * J[6 + j] + T3[3 * i + 2] * J[12 + j];
    memset(Ji, 0, 225 * sizeof(double));
    memset(Jj, 0, 225 * sizeof(double));
    memset(Jg, 0, 45 * sizeof(double));
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
            Jg[3 * (3 + i) + j] = -RiT * dt;
            Jg[3 * (6 + i) + j] = -0.5 * RiT * dt2;
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
    if (whiten) { /* imu_factor.py:749-768 _apply_lower*: X <- W X, bottom-up */
        for (int i = 14; i >= 0; i--) {
            double s = 0.0;
            for (int k = 0; k <= i; k++) s += W[15 * i + k] * r[k];
            r[i] = s;
        }
        double *Xs[3] = {Ji, Jj, Jg};
        int ms[3] = {15, 15, 3};
        for (int q = 0; q < 3; q++) {
            double *X = Xs[q];
            int m = ms[q];
            for (int i = 14; i >= 0; i--)
                for (int c = 0; c < m; c++) {
                    double s = 0.0;
                    for (int k = 0; k <= i; k++) s += W[15 * i + k] * X[m * k + c];
                    X[m * i + c] = s;
                }
        }
    }
}

static void slot_offsets(const ow_problem *pb, int *off) {
    off[0] = 0;
    for (int s = 0; s < pb->P; s++) off[s + 1] = off[s] + (pb->dims ? pb->dims[s] : 6);
}

/* vio/window.py:831-862 _imu_terms at state (R, p, v, bg, ba); H, b may be NULL (cost only).  chi2 (n_imu) opt. */
static double imu_terms(const ow_problem *pb, const double *R, const double *p, const double *v, const double *bg,
                        const double *ba, const int *off, int n, double *H, double *b, double *chi2) {
    double cost = 0.0, kd = pb->imu_k;
    if (pb->n_imu == 0) return 0.0;
    double r[15], Ji[225], Jj[225], Jg[45];
    for (int f = 0; f < pb->n_imu; f++) {
        int i = pb->imu_ij[2 * f], j = pb->imu_ij[2 * f + 1];
        o_imu_factor(R + 9 * i, p + 3 * i, v + 3 * i, bg + 3 * i, ba + 3 * i, R + 9 * j, p + 3 * j, v + 3 * j,
                     bg + 3 * j, ba + 3 * j, pb->g_w, pb->imu_rec + (size_t)IMU_REC * f, 1, r, Ji, Jj, Jg);
        double e2 = 0.0;
        for (int k = 0; k < 15; k++) e2 += r[k] * r[k];
        if (chi2) chi2[f] = e2;
        double e = sqrt(e2), w = (e <= kd) ? 1.0 : kd / e;
        cost += (e <= kd) ? 0.5 * e2 : kd * (e - 0.5 * kd);
        if (!H) continue;
        int ci = off[i], cj = off[j];
        for (int a = 0; a < 15; a++) {
            for (int c = 0; c < 15; c++) {
                double ii = 0.0, jj = 0.0, ij = 0.0;
                for (int k = 0; k < 15; k++) {
                    ii += Ji[15 * k + a] * Ji[15 * k + c];
                    jj += Jj[15 * k + a] * Jj[15 * k + c];
                    ij += Ji[15 * k + a] * Jj[15 * k + c];
                }
                H[(size_t)(ci + a) * n + ci + c] += w * ii;
                H[(size_t)(cj + a) * n + cj + c] += w * jj;
                H[(size_t)(ci + a) * n + cj + c] += w * ij;
                H[(size_t)(cj + c) * n + ci + a] += w * ij;
            }
            double bi = 0.0, bj = 0.0;
            for (int k = 0; k < 15; k++) {
                bi += Ji[15 * k + a] * r[k];
                bj += Jj[15 * k + a] * r[k];
            }
            b[ci + a] += w * bi;
            b[cj + a] += w * bj;
        }
    }
    return cost;
}

/* vio/marg.py:47-57 _boxminus and :102-127 Prior._d/terms/cost (H, b may be NULL) */
static double prior_terms(const ow_problem *pb, const double *R, const double *p, const double *v, const double *bg,
                          const double *ba, const int *off, int n, double *H, double *b) {
    if (pb->prior_nb == 0) return 0.0;
    int np_ = 0;
    for (int k = 0; k < pb->prior_nb; k++) np_ += pb->prior_dim[k];
    double *d = calloc(np_, sizeof(double)), *res = calloc(pb->prior_m, sizeof(double));
    int *cols = malloc(np_ * sizeof(int)), o = 0;
    for (int k = 0; k < pb->prior_nb; k++) {
        int s = pb->prior_slot[k], dim = pb->prior_dim[k];
        const double *x0 = pb->prior_x0 + 21 * k;
        double M[9];
        mtm3(x0, R + 9 * s, M);
        o_so3_log_nb(M, d + o);
        for (int i = 0; i < 3; i++) d[o + 3 + i] = p[3 * s + i] - x0[9 + i];
        if (dim == 15)
            for (int i = 0; i < 3; i++) {
                d[o + 6 + i] = v[3 * s + i] - x0[12 + i];
                d[o + 9 + i] = bg[3 * s + i] - x0[15 + i];
                d[o + 12 + i] = ba[3 * s + i] - x0[18 + i];
            }
        for (int i = 0; i < dim; i++) cols[o + i] = off[s] + i;
        o += dim;
    }
    double cost = 0.0;
    for (int m = 0; m < pb->prior_m; m++) {
        double acc = 0.0;
        for (int c = 0; c < np_; c++) acc += pb->prior_S[(size_t)m * np_ + c] * d[c];
        res[m] = acc + pb->prior_r[m];
        cost += res[m] * res[m];
    }
    if (H) {
        for (int a = 0; a < np_; a++) {
            for (int c = 0; c < np_; c++) {
                double acc = 0.0;
                for (int m = 0; m < pb->prior_m; m++)
                    acc += pb->prior_S[(size_t)m * np_ + a] * pb->prior_S[(size_t)m * np_ + c];
                H[(size_t)cols[a] * n + cols[c]] += acc;
            }
            double acc = 0.0;
            for (int m = 0; m < pb->prior_m; m++) acc += pb->prior_S[(size_t)m * np_ + a] * res[m];
            b[cols[a]] += acc;
        }
    }
    free(d);
    free(res);
    free(cols);
    return 0.5 * cost;
}

/* vio/window.py:865-886 _full_system: scatter H6 into n, add the prior and the IMU terms.  Returns c_rest. */
double ow_full_system(const ow_problem *pb, const double *H6, const double *b6, double *H, double *b) {
    int P = pb->P, *off = malloc((P + 1) * sizeof(int));
    slot_offsets(pb, off);
    int n = off[P], n6 = 6 * P;
    memset(H, 0, (size_t)n * n * sizeof(double));
    memset(b, 0, n * sizeof(double));
    for (int s2 = 0; s2 < P; s2++)
        for (int i = 0; i < 6; i++) {
            for (int s1 = 0; s1 < P; s1++)
                for (int j = 0; j < 6; j++)
                    H[(size_t)(off[s2] + i) * n + off[s1] + j] = H6[(size_t)(6 * s2 + i) * n6 + 6 * s1 + j];
            b[off[s2] + i] = b6[6 * s2 + i];
        }
    double cost = prior_terms(pb, pb->Rwb, pb->pwb, pb->vel, pb->bgs, pb->bas, off, n, H, b);
    if (pb->n_imu && pb->vel) cost += imu_terms(pb, pb->Rwb, pb->pwb, pb->vel, pb->bgs, pb->bas, off, n, H, b, NULL);
    free(off);
    return cost;
}

/* Cholesky (lower, row-major, in place); 0 when not positive definite (numpy raises LinAlgError) */
static int chol(double *A, int n) {
    for (int j = 0; j < n; j++) {
        double s = A[(size_t)j * n + j];
        for (int k = 0; k < j; k++) s -= A[(size_t)j * n + k] * A[(size_t)j * n + k];
        if (!(s > 0.0) || !isfinite(s)) return 0;
        double d = sqrt(s);
        A[(size_t)j * n + j] = d;
        for (int i = j + 1; i < n; i++) {
            double t = A[(size_t)i * n + j];
            for (int k = 0; k < j; k++) t -= A[(size_t)i * n + k] * A[(size_t)j * n + k];
            A[(size_t)i * n + j] = t / d;
        }
    }
    return 1;
}

static void chol_solve(const double *Lc, int n, const double *b, double *x) { /* x = (L L^T)^-1 b */
    for (int i = 0; i < n; i++) {
        double s = b[i];
        for (int k = 0; k < i; k++) s -= Lc[(size_t)i * n + k] * x[k];
        x[i] = s / Lc[(size_t)i * n + i];
    }
    for (int i = n - 1; i >= 0; i--) {
        double s = x[i];
        for (int k = i + 1; k < n; k++) s -= Lc[(size_t)k * n + i] * x[k];
        x[i] = s / Lc[(size_t)i * n + i];
    }
}

int o_chol_solve(double *A, int n, const double *b, double *x) {
    if (!chol(A, n)) return 0;
    chol_solve(A, n, b, x);
    return 1;
}

/* vio/window.py:889-982 solve.  stats: cost0, cost, accepted, iters, lam.  dx_hist (iters, n) optional: the full
 * reduced update of every iteration (NaN when the Cholesky failed); acc_hist (iters) optional. */
int ow_solve(ow_problem *pb, int iters, double lam0, double *stats, double *dx_hist, uint8_t *acc_hist) {
    int P = pb->P, L = pb->L, M = pb->M, n6 = 6 * P;
    int *off = malloc((P + 1) * sizeof(int));
    slot_offsets(pb, off);
    int n = off[P];
    int vi = pb->vel != NULL;
    double *H6 = malloc((size_t)n6 * n6 * sizeof(double)), *b6 = malloc(n6 * sizeof(double));
    double *Hll_inv = calloc((size_t)L * 9, sizeof(double)), *bl = calloc((size_t)L * 3, sizeof(double));
    double *Hlp = calloc((size_t)L * P * 18, sizeof(double));
    uint8_t *lm_ok = calloc(L ? L : 1, 1);
    double *ob_chi2 = malloc((M ? M : 1) * sizeof(double)), *dl = malloc((L ? L : 1) * 3 * sizeof(double));
    double *R2 = malloc(9 * P * sizeof(double)), *p2 = malloc(3 * P * sizeof(double)),
           *l2 = malloc((L ? L : 1) * 3 * sizeof(double));
    double *H = malloc((size_t)n * n * sizeof(double)), *bv = malloc(n * sizeof(double)),
           *dx = malloc(n * sizeof(double)), *dp6 = malloc(n6 * sizeof(double));
    double *v2 = vi ? malloc(3 * P * sizeof(double)) : NULL, *g2 = vi ? malloc(3 * P * sizeof(double)) : NULL,
           *a2 = vi ? malloc(3 * P * sizeof(double)) : NULL;
    int npf = pb->n_pf;
    double *pf_trial = malloc((npf ? npf : 1) * 3 * sizeof(double));
    uint8_t *pf_ok = malloc(npf ? npf : 1);
    double lam = lam0, cost0 = NAN, cur = NAN;
    int accepted = 0, have0 = 0;
    for (int it = 0; it < iters; it++) {
        if (acc_hist) acc_hist[it] = 0;
        if (dx_hist)
            for (int i = 0; i < n; i++) dx_hist[(size_t)it * n + i] = NAN;
        double c_vis = ow_build(pb, pb->Rwb, pb->pwb, pb->lmp, lam, H6, b6, Hll_inv, bl, Hlp, lm_ok, ob_chi2);
        if (npf) {
            c_vis += pf_cost(pb, pb->Rwb, pb->pwb, pb->pf_lmp, 0);
            pf_step(pb, pb->pf_lmp, lam, pf_trial, NULL, pf_ok);
        }
        double c_rest = ow_full_system(pb, H6, b6, H, bv);
        double c_tot = c_vis + c_rest;
        if (!have0) { cost0 = c_tot; have0 = 1; }
        cur = c_tot;
        double *dg = malloc(n * sizeof(double));
        for (int i = 0; i < n; i++) dg[i] = H[(size_t)i * n + i];
        for (int i = 0; i < n; i++) H[(size_t)i * n + i] += lam * dg[i] + 1e-9;
        free(dg);
        if (!chol(H, n)) {
            lam *= 10.0;
            continue;
        }
        chol_solve(H, n, bv, dx);
        for (int i = 0; i < n; i++) dx[i] = -dx[i];
        if (dx_hist)
            for (int i = 0; i < n; i++) dx_hist[(size_t)it * n + i] = dx[i];
        for (int s = 0; s < P; s++)
            for (int i = 0; i < 6; i++) dp6[6 * s + i] = dx[off[s] + i];
        back_substitute(L, P, dp6, Hll_inv, bl, Hlp, lm_ok, dl);
        apply_update(pb, dp6, dl, lm_ok, R2, p2, l2);
        if (vi) {
            memcpy(v2, pb->vel, 3 * P * sizeof(double));
            memcpy(g2, pb->bgs, 3 * P * sizeof(double));
            memcpy(a2, pb->bas, 3 * P * sizeof(double));
            for (int s = 0; s < P; s++)
                if (pb->dims && pb->dims[s] == 15)
                    for (int i = 0; i < 3; i++) {
                        v2[3 * s + i] += dx[off[s] + 6 + i];
                        g2[3 * s + i] += dx[off[s] + 9 + i];
                        a2[3 * s + i] += dx[off[s] + 12 + i];
                    }
        }
        double c_new = ow_visual_cost(pb, R2, p2, l2);
        if (npf) c_new += pf_cost(pb, R2, p2, pf_trial, 1);
        c_new += prior_terms(pb, R2, p2, v2, g2, a2, off, n, NULL, NULL);
        if (vi && pb->n_imu) c_new += imu_terms(pb, R2, p2, v2, g2, a2, off, n, NULL, NULL, NULL);
        if (c_new < c_tot) {
            memcpy(pb->Rwb, R2, 9 * P * sizeof(double));
            memcpy(pb->pwb, p2, 3 * P * sizeof(double));
            memcpy(pb->lmp, l2, 3 * L * sizeof(double));
            if (npf) memcpy(pb->pf_lmp, pf_trial, 3 * npf * sizeof(double));
            if (vi) {
                memcpy(pb->vel, v2, 3 * P * sizeof(double));
                memcpy(pb->bgs, g2, 3 * P * sizeof(double));
                memcpy(pb->bas, a2, 3 * P * sizeof(double));
            }
            lam = fmax(lam / 10.0, 1e-12);
            accepted++;
            cur = c_new;
            if (acc_hist) acc_hist[it] = 1;
        } else {
            lam *= 10.0;
        }
    }
    stats[0] = cost0;
    stats[1] = cur;
    stats[2] = accepted;
    stats[3] = iters;
    stats[4] = lam;
    free(off); free(H6); free(b6); free(Hll_inv); free(bl); free(Hlp); free(lm_ok); free(ob_chi2); free(dl);
    free(R2); free(p2); free(l2); free(H); free(bv); free(dx); free(dp6); free(v2); free(g2); free(a2);
    free(pf_trial); free(pf_ok);
    return 0;
}

/* vio/window.py:985-1008 final_linearisation (lam = 0): H (n,n), b (n), ob_chi2, lm_ok, Hll_inv, imu_chi2, and the
 * pose-free marginals into pf_Hinv / pf_ok when given. */
int ow_final_linearisation(const ow_problem *pb, double *H, double *b, double *ob_chi2, uint8_t *lm_ok,
                           double *Hll_inv, double *imu_chi2, double *pf_Hinv, uint8_t *pf_ok) {
    int P = pb->P, L = pb->L, n6 = 6 * P;
    double *H6 = malloc((size_t)n6 * n6 * sizeof(double)), *b6 = malloc(n6 * sizeof(double));
    double *bl = calloc((size_t)(L ? L : 1) * 3, sizeof(double)), *Hlp = calloc((size_t)(L ? L : 1) * P * 18, sizeof(double));
    ow_build(pb, pb->Rwb, pb->pwb, pb->lmp, 0.0, H6, b6, Hll_inv, bl, Hlp, lm_ok, ob_chi2);
    if (pb->n_pf && pf_Hinv) {
        double *tmp = malloc(3 * pb->n_pf * sizeof(double));
        pf_step(pb, pb->pf_lmp, 0.0, tmp, pf_Hinv, pf_ok);
        free(tmp);
    }
    ow_full_system(pb, H6, b6, H, b);
    if (pb->n_imu && pb->vel && imu_chi2) {
        int *off = malloc((P + 1) * sizeof(int));
        slot_offsets(pb, off);
        imu_terms(pb, pb->Rwb, pb->pwb, pb->vel, pb->bgs, pb->bas, off, off[P], NULL, NULL, imu_chi2);
        free(off);
    }
    free(H6); free(b6); free(bl); free(Hlp);
    return 0;
}

/* ================================================================ backend/ba.py _solve_np (Euclidean, T_cw) == */
typedef struct {
    int32_t nk, nl, N;
    double *poses;              /* (nk,4,4) T_wc in, T_wc out */
    double *pts;                /* (nl,3) in/out */
    const int32_t *okf, *olm;   /* (N) */
    const double *ouv;          /* (N,2) */
    double fx, cx, cy, huber_k;
    const uint8_t *free_mask;   /* (nk) */
    /* optional inverse-range prior (ba.py:774-794) */
    const double *pr_c;         /* (nl,3) or NULL */
    const double *pr_rho0;      /* (nl) */
    const double *pr_invsig;    /* (nl) */
    const double *pr_valid;     /* (nl) */
    /* optional gyro rotation factors (ba.py:644-697) */
    int32_t n_rot;
    const int32_t *rot_ab;      /* (n_rot,2) */
    const double *rot_M;        /* (n_rot,3,3) */
    const double *rot_sig;      /* (n_rot) */
    /* LM policy: 0 _solve_np (trial cost, accept iff the unrobust chi2 does not increase); 1 the GPU drivers
     * _solve_gpu / _solve_gpu_id (ba.py:1015-1076, 1436-1497): a step is always applied, the next iteration rejects it
     * when the Huber-weighted chi2 grew (restore the last accepted state, lam x4), a non-finite step sets lam x10,
     * the rotation factors' Hessian enters, SO(3) is projected once at the end, Cholesky instead of LU */
    int32_t lm_policy;
} onp_problem;

static double np_residuals(const onp_problem *q, const double *Tcw, const double *pts, double *r, double *Pc) {
    double s = 0.0;
    for (int k = 0; k < q->N; k++) {
        const double *T = Tcw + 16 * q->okf[k], *X = pts + 3 * q->olm[k];
        double P3[3];
        for (int i = 0; i < 3; i++) P3[i] = (T[4 * i] * X[0] + T[4 * i + 1] * X[1] + T[4 * i + 2] * X[2]) + T[4 * i + 3];
        double z = fmax(P3[2], 1e-6);
        r[2 * k] = q->fx * P3[0] / z + q->cx - q->ouv[2 * k];
        r[2 * k + 1] = q->fx * P3[1] / z + q->cy - q->ouv[2 * k + 1];
        if (Pc) memcpy(Pc + 3 * k, P3, sizeof(P3));
        s += r[2 * k] * r[2 * k] + r[2 * k + 1] * r[2 * k + 1];
    }
    return s;
}

/* LU with partial pivoting (numpy.linalg.solve / LAPACK gesv); 0 when exactly singular */
static int lu_solve(double *A, int n, double *b) {
    for (int c = 0; c < n; c++) {
        int p = c;
        for (int i = c + 1; i < n; i++)
            if (fabs(A[(size_t)i * n + c]) > fabs(A[(size_t)p * n + c])) p = i;
        if (A[(size_t)p * n + c] == 0.0) return 0;
        if (p != c) {
            for (int j = 0; j < n; j++) {
                double t = A[(size_t)c * n + j];
                A[(size_t)c * n + j] = A[(size_t)p * n + j];
                A[(size_t)p * n + j] = t;
            }
            double t = b[c];
            b[c] = b[p];
            b[p] = t;
        }
        for (int i = c + 1; i < n; i++) {
            double f = A[(size_t)i * n + c] / A[(size_t)c * n + c];
            if (f == 0.0) continue;
            for (int j = c + 1; j < n; j++) A[(size_t)i * n + j] -= f * A[(size_t)c * n + j];
            b[i] -= f * b[c];
        }
    }
    for (int i = n - 1; i >= 0; i--) {
        double s = b[i];
        for (int j = i + 1; j < n; j++) s -= A[(size_t)i * n + j] * b[j];
        b[i] = s / A[(size_t)i * n + i];
    }
    return 1;
}

/* backend/ba.py:644-697 with the Hessian blocks (H may be NULL: gradient only); defined below */
static void rot_blocks(int n_rot, const int32_t *rot_ab, const double *rot_M, const double *rot_sig, const double *Tcw,
                       int n, double *H, double *g);

/* backend/ba.py:644-697: only the gradient is used by _solve_np (ba.py:825-828) */
static void rot_factor_grad(const onp_problem *q, const double *Tcw, double *g) {
    for (int f = 0; f < q->n_rot; f++) {
        int a = q->rot_ab[2 * f], b = q->rot_ab[2 * f + 1];
        double w = 1.0 / fmax(q->rot_sig[f], 1e-9), Ri[9], Rj[9], A[9], E[9], e[3], Jr[9], Ja[9], Jb[9], r[3];
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) {
                Ri[3 * i + j] = Tcw[16 * a + 4 * i + j];
                Rj[3 * i + j] = Tcw[16 * b + 4 * i + j];
            }
        mmt3(Ri, Rj, A);
        mtm3(q->rot_M + 9 * f, A, E);
        o_log_quat(E, e);
        o_jr_inv(e, Jr);
        mmt3(Jr, A, Ja);
        for (int i = 0; i < 9; i++) {
            Ja[i] *= w;
            Jb[i] = -w * Jr[i];
        }
        for (int i = 0; i < 3; i++) r[i] = w * e[i];
        for (int i = 0; i < 3; i++) {
            double ga = 0.0, gb = 0.0;
            for (int j = 0; j < 3; j++) {
                ga += Ja[3 * j + i] * r[j];
                gb += Jb[3 * j + i] * r[j];
            }
            g[6 * a + 3 + i] += -ga;
            g[6 * b + 3 + i] += -gb;
        }
    }
}

/* backend/ba.py:713-869.  stats: chi2_0, chi2_n, rms_0, rms_n.  dx_hist (iters, 6 nk) optional. */
int onp_solve(onp_problem *q, int iters, double lam0, double *stats, double *dx_hist, uint8_t *acc_hist) {
    int nk = q->nk, nl = q->nl, N = q->N, n = 6 * nk;
    double fx = q->fx, K = q->huber_k, lam = lam0;
    const int pol = q->lm_policy;
    double chi_prev = -1.0, best_chi = -1.0, rms0_w = NAN;  /* the GPU drivers' policy state */
    double *Tcw = malloc(16 * nk * sizeof(double)), *Tn = malloc(16 * nk * sizeof(double));
    for (int i = 0; i < nk; i++) se3_inv(q->poses + 16 * i, Tcw + 16 * i);
    double *r = malloc(2 * (N ? N : 1) * sizeof(double)), *Pc = malloc(3 * (N ? N : 1) * sizeof(double));
    double *r2 = malloc(2 * (N ? N : 1) * sizeof(double));
    double *pts = q->pts, *pn = malloc(3 * (nl ? nl : 1) * sizeof(double));
    double chi0 = np_residuals(q, Tcw, pts, r, NULL);
    double rms0 = N ? sqrt(chi0 / N) : NAN;
    double *Hpp = malloc((size_t)n * n * sizeof(double)), *Hred = malloc((size_t)n * n * sizeof(double));
    double *bp = malloc(n * sizeof(double)), *Hll = malloc(9 * (nl ? nl : 1) * sizeof(double));
    double *Hlli = malloc(9 * (nl ? nl : 1) * sizeof(double)), *bl = malloc(3 * (nl ? nl : 1) * sizeof(double));
    double *Hpl = malloc(18 * (N ? N : 1) * sizeof(double)), *T3 = malloc(18 * (N ? N : 1) * sizeof(double));
    double *dxl = malloc(3 * (nl ? nl : 1) * sizeof(double));
    int *order = malloc((N ? N : 1) * sizeof(int)), *bounds = malloc((nl + 1) * sizeof(int));
    double *Tbest = malloc(16 * nk * sizeof(double)), *pbest = malloc(3 * (nl ? nl : 1) * sizeof(double));
    /* stable argsort by landmark: counting sort */
    memset(bounds, 0, (nl + 1) * sizeof(int));
    for (int k = 0; k < N; k++) bounds[q->olm[k] + 1]++;
    for (int j = 0; j < nl; j++) bounds[j + 1] += bounds[j];
    {
        int *fill = malloc((size_t)(nl > 0 ? nl : 1) * sizeof(int));
        for (int j = 0; j < nl; j++) fill[j] = bounds[j];
        for (int k = 0; k < N; k++) order[fill[q->olm[k]]++] = k;
        free(fill);
    }
    for (int it = 0; it < iters; it++) {
        if (acc_hist) acc_hist[it] = 0;
        if (dx_hist)
            for (int i = 0; i < n; i++) dx_hist[(size_t)it * n + i] = NAN;
        double chi_cur = np_residuals(q, Tcw, pts, r, Pc);
        memset(Hpp, 0, (size_t)n * n * sizeof(double));
        memset(bp, 0, n * sizeof(double));
        memset(Hll, 0, 9 * nl * sizeof(double));
        memset(bl, 0, 3 * nl * sizeof(double));
        double chiw = 0.0;
        for (int k = 0; k < N; k++) {
            double e = sqrt(r[2 * k] * r[2 * k] + r[2 * k + 1] * r[2 * k + 1]);
            double w = (e <= K) ? 1.0 : K / fmax(e, 1e-9), sw = sqrt(w);
            double x = Pc[3 * k], y = Pc[3 * k + 1], zz = fmax(Pc[3 * k + 2], 1e-6), iz = 1.0 / zz, iz2 = iz * iz;
            double Jp3[2][3] = {{fx * iz, 0.0, -fx * x * iz2}, {0.0, fx * iz, -fx * y * iz2}};
            double dP[3][6] = {{1, 0, 0, 0, zz, -y}, {0, 1, 0, -zz, 0, x}, {0, 0, 1, y, -x, 0}};
            const double *T = Tcw + 16 * q->okf[k];
            double Jpose[2][6], Jlm[2][3], rw[2];
            for (int a = 0; a < 2; a++) {
                for (int c = 0; c < 6; c++)
                    Jpose[a][c] = (Jp3[a][0] * dP[0][c] + Jp3[a][1] * dP[1][c] + Jp3[a][2] * dP[2][c]) * sw;
                for (int c = 0; c < 3; c++)
                    Jlm[a][c] = (Jp3[a][0] * T[c] + Jp3[a][1] * T[4 + c] + Jp3[a][2] * T[8 + c]) * sw;
                rw[a] = r[2 * k + a] * sw;
            }
            chiw += rw[0] * rw[0] + rw[1] * rw[1];
            int l = q->olm[k], kf = q->okf[k];
            for (int i = 0; i < 3; i++) {
                for (int j = 0; j < 3; j++) Hll[9 * l + 3 * i + j] += Jlm[0][i] * Jlm[0][j] + Jlm[1][i] * Jlm[1][j];
                bl[3 * l + i] += -(Jlm[0][i] * rw[0] + Jlm[1][i] * rw[1]);
            }
            for (int i = 0; i < 6; i++)
                for (int j = 0; j < 3; j++) Hpl[18 * k + 3 * i + j] = Jpose[0][i] * Jlm[0][j] + Jpose[1][i] * Jlm[1][j];
            for (int a = 0; a < 6; a++) {
                bp[kf * 6 + a] += -(Jpose[0][a] * rw[0] + Jpose[1][a] * rw[1]);
                for (int b = 0; b < 6; b++)
                    Hpp[(size_t)(kf * 6 + a) * n + kf * 6 + b] += Jpose[0][a] * Jpose[0][b] + Jpose[1][a] * Jpose[1][b];
            }
        }
        if (pol) { /* ba.py:1015-1026: reject the previous step when the weighted chi2 grew */
            if (isnan(rms0_w)) rms0_w = sqrt(chiw / N);
            if (chi_prev >= 0.0 && chiw > chi_prev) {
                memcpy(Tcw, Tbest, 16 * nk * sizeof(double));
                memcpy(pts, pbest, 3 * nl * sizeof(double));
                lam = fmin(lam * 4.0, 1e2);
                continue;
            }
            chi_prev = chiw;
            best_chi = chiw;
            memcpy(Tbest, Tcw, 16 * nk * sizeof(double));
            memcpy(pbest, pts, 3 * nl * sizeof(double));
        }
        if (q->pr_c) { /* ba.py:774-794 */
            int any = 0;
            for (int j = 0; j < nl; j++) {
                double dP[3] = {pts[3 * j] - q->pr_c[3 * j], pts[3 * j + 1] - q->pr_c[3 * j + 1],
                                pts[3 * j + 2] - q->pr_c[3 * j + 2]};
                double dist = sqrt(dP[0] * dP[0] + dP[1] * dP[1] + dP[2] * dP[2]);
                if (q->pr_valid[j] > 0.5 && dist > 1e-6) any = 1;
            }
            if (any)
                for (int j = 0; j < nl; j++) {
                    double dP[3] = {pts[3 * j] - q->pr_c[3 * j], pts[3 * j + 1] - q->pr_c[3 * j + 1],
                                    pts[3 * j + 2] - q->pr_c[3 * j + 2]};
                    double dist = sqrt(dP[0] * dP[0] + dP[1] * dP[1] + dP[2] * dP[2]);
                    int ok = q->pr_valid[j] > 0.5 && dist > 1e-6;
                    double rho = ok ? 1.0 / dist : 0.0, rp = (rho - q->pr_rho0[j]) * q->pr_invsig[j];
                    double wp = fabs(rp) <= K ? 1.0 : K / fmax(fabs(rp), 1e-9), swp = sqrt(wp);
                    double gg = ok ? -rho * rho / dist * q->pr_invsig[j] : 0.0, Jp_[3], rpw = ok ? rp * swp : 0.0;
                    for (int i = 0; i < 3; i++) Jp_[i] = ok ? dP[i] * (gg * swp) : 0.0;
                    for (int i = 0; i < 3; i++) {
                        for (int c = 0; c < 3; c++) Hll[9 * j + 3 * i + c] += Jp_[i] * Jp_[c];
                        bl[3 * j + i] -= Jp_[i] * rpw;
                    }
                }
        }
        int singular = 0;
        for (int j = 0; j < nl; j++) {
            double tr = Hll[9 * j] + Hll[9 * j + 4] + Hll[9 * j + 8], d = lam * fmax(tr / 3.0, 1e-9);
            for (int i = 0; i < 3; i++) Hll[9 * j + 4 * i] += d;
            if (!inv3(Hll + 9 * j, Hlli + 9 * j)) singular = 1;
        }
        if (singular) break; /* np.linalg.inv raises LinAlgError out of the solver */
        for (int k = 0; k < N; k++) {
            const double *Hi = Hlli + 9 * q->olm[k];
            for (int i = 0; i < 6; i++)
                for (int j = 0; j < 3; j++)
                    T3[18 * k + 3 * i + j] = Hpl[18 * k + 3 * i] * Hi[j] + Hpl[18 * k + 3 * i + 1] * Hi[3 + j] +
                                             Hpl[18 * k + 3 * i + 2] * Hi[6 + j];
        }
        memcpy(Hred, Hpp, (size_t)n * n * sizeof(double));
        for (int j = 0; j < nl; j++) {
            int a0 = bounds[j], a1 = bounds[j + 1];
            for (int ai = a0; ai < a1; ai++) {
                int ka = order[ai], ki = q->okf[ka] * 6;
                const double *Ta = T3 + 18 * ka;
                for (int i = 0; i < 6; i++)
                    bp[ki + i] -= Ta[3 * i] * bl[3 * j] + Ta[3 * i + 1] * bl[3 * j + 1] + Ta[3 * i + 2] * bl[3 * j + 2];
                for (int bi = a0; bi < a1; bi++) {
                    int kb = order[bi], kk = q->okf[kb] * 6;
                    const double *Hb = Hpl + 18 * kb;
                    for (int i = 0; i < 6; i++)
                        for (int c = 0; c < 6; c++)
                            Hred[(size_t)(ki + i) * n + kk + c] -=
                                Ta[3 * i] * Hb[3 * c] + Ta[3 * i + 1] * Hb[3 * c + 1] + Ta[3 * i + 2] * Hb[3 * c + 2];
                }
            }
        }
        if (q->n_rot) {
            if (pol)
                rot_blocks(q->n_rot, q->rot_ab, q->rot_M, q->rot_sig, Tcw, n, Hred, bp); /* ba.py:1041-1047 */
            else
                rot_factor_grad(q, Tcw, bp);
        }
        for (int i = 0; i < nk; i++) {
            if (q->free_mask[i]) continue;
            int f0 = 6 * i;
            for (int a = 0; a < 6; a++) {
                for (int c = 0; c < n; c++) {
                    Hred[(size_t)(f0 + a) * n + c] = 0.0;
                    Hred[(size_t)c * n + f0 + a] = 0.0;
                }
            }
            for (int a = 0; a < 6; a++) {
                Hred[(size_t)(f0 + a) * n + f0 + a] = 1.0;
                bp[f0 + a] = 0.0;
            }
        }
        for (int i = 0; i < n; i++) Hred[(size_t)i * n + i] += lam * fmax(Hred[(size_t)i * n + i], 1e-9);
        int solved;
        if (pol) { /* the GPU Cholesky; its status is ignored there, a failure shows as a non-finite step */
            solved = chol(Hred, n);
            if (solved) {
                double *x = malloc(n * sizeof(double));
                chol_solve(Hred, n, bp, x);
                memcpy(bp, x, n * sizeof(double));
                free(x);
            }
        } else {
            solved = lu_solve(Hred, n, bp);
            if (!solved) break;
        }
        double *dxp = bp;
        if (dx_hist)
            for (int i = 0; i < n; i++) dx_hist[(size_t)it * n + i] = dxp[i];
        for (int j = 0; j < nl; j++) {
            double rhs[3] = {bl[3 * j], bl[3 * j + 1], bl[3 * j + 2]};
            for (int ai = bounds[j]; ai < bounds[j + 1]; ai++) {
                int k = order[ai];
                const double *Hk = Hpl + 18 * k, *dk = dxp + 6 * q->okf[k];
                for (int c = 0; c < 3; c++) {
                    double s = 0.0;
                    for (int i = 0; i < 6; i++) s += Hk[3 * i + c] * dk[i];
                    rhs[c] += -s;
                }
            }
            const double *Hi = Hlli + 9 * j;
            for (int i = 0; i < 3; i++) dxl[3 * j + i] = Hi[3 * i] * rhs[0] + Hi[3 * i + 1] * rhs[1] + Hi[3 * i + 2] * rhs[2];
        }
        if (pol) { /* ba.py:1058-1073: non-finite -> lam x10; else apply (no projection), lam / 2 */
            int fin = solved;
            for (int i = 0; i < n && fin; i++) fin = isfinite(dxp[i]);
            for (int i = 0; i < 3 * nl && fin; i++) fin = isfinite(dxl[i]);
            if (!fin) {
                if (dx_hist)
                    for (int i = 0; i < n; i++) dx_hist[(size_t)it * n + i] = NAN;
                lam = fmin(lam * 10.0, 1e3);
                continue;
            }
            for (int i = 0; i < nk; i++) {
                if (!q->free_mask[i]) continue;
                double E[16], T[16];
                o_se3_exp(dxp + 6 * i, E);
                for (int a = 0; a < 4; a++)
                    for (int c = 0; c < 4; c++) {
                        double sum = 0.0;
                        for (int m = 0; m < 4; m++) sum += E[4 * a + m] * Tcw[16 * i + 4 * m + c];
                        T[4 * a + c] = sum;
                    }
                memcpy(Tcw + 16 * i, T, sizeof(T));
            }
            for (int j = 0; j < 3 * nl; j++) pts[j] += dxl[j];
            lam = fmax(lam * 0.5, 1e-7);
            if (acc_hist) acc_hist[it] = 1;
            continue;
        }
        memcpy(Tn, Tcw, 16 * nk * sizeof(double));
        for (int i = 0; i < nk; i++) {
            if (!q->free_mask[i]) continue;
            double E[16], R[9], Rp[9];
            o_se3_exp(dxp + 6 * i, E);
            for (int a = 0; a < 4; a++)
                for (int c = 0; c < 4; c++) {
                    double s = 0.0;
                    for (int m = 0; m < 4; m++) s += E[4 * a + m] * Tcw[16 * i + 4 * m + c];
                    Tn[16 * i + 4 * a + c] = s;
                }
            for (int a = 0; a < 3; a++)
                for (int c = 0; c < 3; c++) R[3 * a + c] = Tn[16 * i + 4 * a + c];
            o_so3_project(R, Rp);
            for (int a = 0; a < 3; a++)
                for (int c = 0; c < 3; c++) Tn[16 * i + 4 * a + c] = Rp[3 * a + c];
        }
        for (int j = 0; j < 3 * nl; j++) pn[j] = pts[j] + dxl[j];
        double chi_new = np_residuals(q, Tn, pn, r2, NULL);
        if (chi_new <= chi_cur) {
            memcpy(Tcw, Tn, 16 * nk * sizeof(double));
            memcpy(pts, pn, 3 * nl * sizeof(double));
            lam = fmax(lam * 0.5, 1e-7);
            if (acc_hist) acc_hist[it] = 1;
        } else {
            lam = fmin(lam * 4.0, 1e2);
        }
    }
    double chin = np_residuals(q, Tcw, pts, r, NULL);
    stats[0] = chi0;
    stats[1] = chin;
    stats[2] = rms0;
    stats[3] = N ? sqrt(chin / N) : NAN;
    if (pol) { /* ba.py:1075-1089: one SO(3) projection, then the Huber-weighted chi2 of the result */
        for (int i = 0; i < nk; i++) {
            double R[9], Rp[9];
            for (int a = 0; a < 3; a++)
                for (int c = 0; c < 3; c++) R[3 * a + c] = Tcw[16 * i + 4 * a + c];
            o_so3_project(R, Rp);
            for (int a = 0; a < 3; a++)
                for (int c = 0; c < 3; c++) Tcw[16 * i + 4 * a + c] = Rp[3 * a + c];
        }
        np_residuals(q, Tcw, pts, r, NULL);
        double cw = 0.0;
        for (int k = 0; k < N; k++) {
            double e2 = r[2 * k] * r[2 * k] + r[2 * k + 1] * r[2 * k + 1], e = sqrt(e2);
            double w = (e <= K) ? 1.0 : K / fmax(e, 1e-9);
            cw += w * e2;
        }
        stats[0] = best_chi > 0.0 ? best_chi : cw;
        stats[1] = cw;
        stats[2] = rms0_w;
        stats[3] = sqrt(cw / N);
    }
    for (int i = 0; i < nk; i++) se3_inv(Tcw + 16 * i, q->poses + 16 * i);
    free(Tcw); free(Tn); free(r); free(Pc); free(r2); free(pn); free(Hpp); free(Hred); free(bp); free(Hll);
    free(Hlli); free(bl); free(Hpl); free(T3); free(dxl); free(order); free(bounds); free(Tbest); free(pbest);
    return 0;
}

/* ======================================================= backend/ba.py _solve_np_id (inverse depth, T_cw) == */
typedef struct {
    int32_t nk, nl, N;
    double *poses;              /* (nk,4,4) T_wc in, T_wc out */
    double *rho;                /* (nl) in/out */
    double *mbear;              /* (nl,3) host-frame bearing (x_n, y_n, 1), in/out */
    const int32_t *hidx;        /* (nl) host keyframe, or -1 (host pose from hpose_ext) */
    const double *hpose_ext;    /* (nl,12) T_wc rows [R | t] of hosts outside the window */
    const int32_t *okf, *olm;   /* (N) */
    const double *ouv;          /* (N,2) */
    double fx, cx, cy, huber_k, rho_floor;
    const uint8_t *free_mask;   /* (nk) */
    /* optional prior (ba.py:1250-1264): m0 (nl,3), rho0, wbear, invsig, valid (nl) */
    const double *pr_m0, *pr_rho0, *pr_wbear, *pr_invsig, *pr_valid;
    int32_t n_rot;
    const int32_t *rot_ab;
    const double *rot_M, *rot_sig;
    int32_t lm_policy;          /* as onp_problem.lm_policy */
} onpid_problem;

/* backend/ba.py:1116-1135 _host_frames: (R_wc, t_wc) of landmark l's host */
static void id_host(const onpid_problem *q, const double *Twc, int l, double *R, double *t) {
    int h = q->hidx[l];
    if (h >= 0) {
        const double *T = Twc + 16 * h;
        for (int i = 0; i < 3; i++) {
            for (int j = 0; j < 3; j++) R[3 * i + j] = T[4 * i + j];
            t[i] = T[4 * i + 3];
        }
    } else {
        const double *E = q->hpose_ext + 12 * l;
        for (int i = 0; i < 3; i++) {
            for (int j = 0; j < 3; j++) R[3 * i + j] = E[4 * i + j];
            t[i] = E[4 * i + 3];
        }
    }
}

/* backend/ba.py:1226-1234 geom: residuals r (N,2), camera points Pc (N,3), host points Ph (N,3); returns sum r^2 */
static double id_residuals(const onpid_problem *q, const double *Twc, const double *Tcw, const double *rho,
                           const double *mb, double *r, double *Pc, double *Ph) {
    double s = 0.0;
    for (int k = 0; k < q->N; k++) {
        int l = q->olm[k];
        double R[9], t[3], P[3], Pw[3], Pcc[3];
        id_host(q, Twc, l, R, t);
        double ir = fmax(rho[l], 1e-12);
        for (int i = 0; i < 3; i++) P[i] = mb[3 * l + i] / ir;
        for (int i = 0; i < 3; i++) Pw[i] = (R[3 * i] * P[0] + R[3 * i + 1] * P[1] + R[3 * i + 2] * P[2]) + t[i];
        const double *T = Tcw + 16 * q->okf[k];
        for (int i = 0; i < 3; i++) Pcc[i] = (T[4 * i] * Pw[0] + T[4 * i + 1] * Pw[1] + T[4 * i + 2] * Pw[2]) + T[4 * i + 3];
        double z = fmax(Pcc[2], 1e-6);
        r[2 * k] = q->fx * Pcc[0] / z + q->cx - q->ouv[2 * k];
        r[2 * k + 1] = q->fx * Pcc[1] / z + q->cy - q->ouv[2 * k + 1];
        if (Pc) memcpy(Pc + 3 * k, Pcc, sizeof(Pcc));
        if (Ph) memcpy(Ph + 3 * k, P, sizeof(P));
        s += r[2 * k] * r[2 * k] + r[2 * k + 1] * r[2 * k + 1];
    }
    return s;
}

/* backend/ba.py:1138-1209 _analytic_blocks_id for one observation: J (32) = Jobs (2x6), Jhost (2x6), Jlm (2x3),
 * all times sw, and rw = r sw (2).  Jhost = 0 when the host is outside the window. */
void o_id_blocks(const double *Tcw_k, const double *Rh, double rho, const double *Pc, const double *Ph, int host_in,
                 double fx, const double *r, double sw, double *J) {
    double x = Pc[0], y = Pc[1], zz = fmax(Pc[2], 1e-6), iz = 1.0 / zz;
    double A[2][3] = {{fx * iz, 0.0, -fx * x * iz * iz}, {0.0, fx * iz, -fx * y * iz * iz}};
    double dPo[3][6] = {{1, 0, 0, 0, zz, -y}, {0, 1, 0, -zz, 0, x}, {0, 0, 1, y, -x, 0}};
    double Rcw[9], Rch[9], ARch[2][3];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) Rcw[3 * i + j] = Tcw_k[4 * i + j];
    mm3(Rcw, Rh, Rch);
    for (int a = 0; a < 2; a++)
        for (int c = 0; c < 3; c++) ARch[a][c] = A[a][0] * Rch[c] + A[a][1] * Rch[3 + c] + A[a][2] * Rch[6 + c];
    double dPh[3][6] = {{-1, 0, 0, 0, -Ph[2], Ph[1]}, {0, -1, 0, Ph[2], 0, -Ph[0]}, {0, 0, -1, -Ph[1], Ph[0], 0}};
    double irho = 1.0 / fmax(rho, 1e-12);
    for (int a = 0; a < 2; a++) {
        for (int c = 0; c < 6; c++) {
            J[6 * a + c] = (A[a][0] * dPo[0][c] + A[a][1] * dPo[1][c] + A[a][2] * dPo[2][c]) * sw;
            J[12 + 6 * a + c] =
                host_in ? (ARch[a][0] * dPh[0][c] + ARch[a][1] * dPh[1][c] + ARch[a][2] * dPh[2][c]) * sw : 0.0;
        }
        J[24 + 3 * a] = ARch[a][0] * irho * sw;
        J[24 + 3 * a + 1] = ARch[a][1] * irho * sw;
        J[24 + 3 * a + 2] = -(ARch[a][0] * Ph[0] + ARch[a][1] * Ph[1] + ARch[a][2] * Ph[2]) * irho * sw;
        J[30 + a] = r[a] * sw;
    }
}

/* backend/ba.py:644-697 with the Hessian blocks: _solve_np_id adds both to Hpp and grad (ba.py:1293-1296) */
static void rot_blocks(int n_rot, const int32_t *rot_ab, const double *rot_M, const double *rot_sig, const double *Tcw,
                       int n, double *H, double *g) {
    for (int f = 0; f < n_rot; f++) {
        int a = rot_ab[2 * f], b = rot_ab[2 * f + 1];
        double w = 1.0 / fmax(rot_sig[f], 1e-9), Ri[9], Rj[9], A[9], E[9], e[3], Jr[9], Ja[9], Jb[9], r[3];
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) {
                Ri[3 * i + j] = Tcw[16 * a + 4 * i + j];
                Rj[3 * i + j] = Tcw[16 * b + 4 * i + j];
            }
        mmt3(Ri, Rj, A);
        mtm3(rot_M + 9 * f, A, E);
        o_log_quat(E, e);
        o_jr_inv(e, Jr);
        mmt3(Jr, A, Ja);
        for (int i = 0; i < 9; i++) {
            Ja[i] *= w;
            Jb[i] = -w * Jr[i];
        }
        for (int i = 0; i < 3; i++) r[i] = w * e[i];
        int ia = 6 * a + 3, ib = 6 * b + 3;
        for (int i = 0; i < 3 && H; i++) {
            for (int j = 0; j < 3; j++) {
                double haa = 0.0, hab = 0.0, hbb = 0.0;
                for (int m = 0; m < 3; m++) {
                    haa += Ja[3 * m + i] * Ja[3 * m + j];
                    hab += Ja[3 * m + i] * Jb[3 * m + j];
                    hbb += Jb[3 * m + i] * Jb[3 * m + j];
                }
                H[(size_t)(ia + i) * n + ia + j] += haa;
                H[(size_t)(ia + i) * n + ib + j] += hab;
                H[(size_t)(ib + j) * n + ia + i] += hab;
                H[(size_t)(ib + i) * n + ib + j] += hbb;
            }
        }
        for (int i = 0; i < 3; i++) {
            double ga = 0.0, gb = 0.0;
            for (int j = 0; j < 3; j++) {
                ga += Ja[3 * j + i] * r[j];
                gb += Jb[3 * j + i] * r[j];
            }
            g[ia + i] += -ga;
            g[ib + i] += -gb;
        }
    }
}

/* backend/ba.py:1212-1342 _solve_np_id.  stats: chi2_0, chi2_n, rms_0, rms_n.  dx_hist (iters, 6 nk) optional. */
int onpid_solve(onpid_problem *q, int iters, double lam0, double *stats, double *dx_hist, uint8_t *acc_hist) {
    int nk = q->nk, nl = q->nl, N = q->N, n = 6 * nk;
    double K = q->huber_k, lam = lam0;
    const int pol = q->lm_policy;
    double chi_prev = -1.0, best_chi = -1.0, rms0_w = NAN;  /* the GPU driver's policy state */
    size_t Nn = N ? N : 1, nln = nl ? nl : 1;
    double *Tcw = malloc(16 * nk * sizeof(double)), *Tn = malloc(16 * nk * sizeof(double));
    double *Twc = malloc(16 * nk * sizeof(double)), *Twn = malloc(16 * nk * sizeof(double));
    memcpy(Twc, q->poses, 16 * nk * sizeof(double));
    for (int i = 0; i < nk; i++) se3_inv(Twc + 16 * i, Tcw + 16 * i);
    double *r = malloc(2 * Nn * sizeof(double)), *Pc = malloc(3 * Nn * sizeof(double));
    double *Ph = malloc(3 * Nn * sizeof(double)), *r2 = malloc(2 * Nn * sizeof(double));
    double *J = malloc(32 * Nn * sizeof(double)), *Hpl = malloc(36 * Nn * sizeof(double));
    double *rho = q->rho, *mb = q->mbear;
    double *rhon = malloc(nln * sizeof(double)), *mbn = malloc(3 * nln * sizeof(double));
    double *Hpp = malloc((size_t)n * n * sizeof(double)), *Hred = malloc((size_t)n * n * sizeof(double));
    double *bp = malloc(n * sizeof(double)), *Hll = malloc(9 * nln * sizeof(double));
    double *Hlli = malloc(9 * nln * sizeof(double)), *bl = malloc(3 * nln * sizeof(double));
    double *dxl = malloc(3 * nln * sizeof(double));
    int *order = malloc(Nn * sizeof(int)), *bounds = malloc((nl + 1) * sizeof(int));
    double *Tbest = malloc(16 * nk * sizeof(double)), *Wbest = malloc(16 * nk * sizeof(double));
    double *rbest = malloc(nln * sizeof(double)), *mbest = malloc(3 * nln * sizeof(double));
    memset(bounds, 0, (nl + 1) * sizeof(int));
    for (int k = 0; k < N; k++) bounds[q->olm[k] + 1]++;
    for (int j = 0; j < nl; j++) bounds[j + 1] += bounds[j];
    {
        int *fill = malloc((size_t)(nl > 0 ? nl : 1) * sizeof(int));
        for (int j = 0; j < nl; j++) fill[j] = bounds[j];
        for (int k = 0; k < N; k++) order[fill[q->olm[k]]++] = k;
        free(fill);
    }
    double chi0 = id_residuals(q, Twc, Tcw, rho, mb, r, NULL, NULL);
    double rms0 = N ? sqrt(chi0 / N) : NAN;
    for (int it = 0; it < iters; it++) {
        if (acc_hist) acc_hist[it] = 0;
        if (dx_hist)
            for (int i = 0; i < n; i++) dx_hist[(size_t)it * n + i] = NAN;
        double chi_cur = id_residuals(q, Twc, Tcw, rho, mb, r, Pc, Ph);
        memset(Hpp, 0, (size_t)n * n * sizeof(double));
        memset(bp, 0, n * sizeof(double));
        memset(Hll, 0, 9 * nln * sizeof(double));
        memset(bl, 0, 3 * nln * sizeof(double));
        double chiw = 0.0;
        for (int k = 0; k < N; k++) {
            int l = q->olm[k], ko = q->okf[k], kh = q->hidx[l];
            double e = sqrt(r[2 * k] * r[2 * k] + r[2 * k + 1] * r[2 * k + 1]);
            double w = (e <= K) ? 1.0 : K / fmax(e, 1e-9), sw = sqrt(w), Rh[9], th[3];
            id_host(q, Twc, l, Rh, th);
            double *Jk = J + 32 * k;
            o_id_blocks(Tcw + 16 * ko, Rh, rho[l], Pc + 3 * k, Ph + 3 * k, kh >= 0, q->fx, r + 2 * k, sw, Jk);
            const double *Jl = Jk + 24, *rw = Jk + 30;
            chiw += rw[0] * rw[0] + rw[1] * rw[1];
            for (int i = 0; i < 3; i++) {
                for (int j = 0; j < 3; j++) Hll[9 * l + 3 * i + j] += Jl[i] * Jl[j] + Jl[3 + i] * Jl[3 + j];
                bl[3 * l + i] += -(Jl[i] * rw[0] + Jl[3 + i] * rw[1]);
            }
            for (int sa = 0; sa < 2; sa++) {
                const double *Ja = Jk + 12 * sa;
                for (int i = 0; i < 6; i++)
                    for (int c = 0; c < 3; c++) Hpl[36 * k + 18 * sa + 3 * i + c] = Ja[i] * Jl[c] + Ja[6 + i] * Jl[3 + c];
            }
            int slot[2] = {ko, kh};
            for (int sa = 0; sa < 2; sa++) {
                if (slot[sa] < 0) continue;
                const double *Ja = Jk + 12 * sa;
                for (int a = 0; a < 6; a++) bp[6 * slot[sa] + a] += -(Ja[a] * rw[0] + Ja[6 + a] * rw[1]);
                for (int sb = 0; sb < 2; sb++) {
                    if (slot[sb] < 0) continue;
                    const double *Jb = Jk + 12 * sb;
                    for (int a = 0; a < 6; a++)
                        for (int b = 0; b < 6; b++)
                            Hpp[(size_t)(6 * slot[sa] + a) * n + 6 * slot[sb] + b] += Ja[a] * Jb[b] + Ja[6 + a] * Jb[6 + b];
                }
            }
        }
        if (pol) { /* ba.py:1436-1445 */
            if (isnan(rms0_w)) rms0_w = sqrt(chiw / (N > 1 ? N : 1));
            if (chi_prev >= 0.0 && chiw > chi_prev) {
                memcpy(Tcw, Tbest, 16 * nk * sizeof(double));
                memcpy(Twc, Wbest, 16 * nk * sizeof(double));
                memcpy(rho, rbest, nl * sizeof(double));
                memcpy(mb, mbest, 3 * nl * sizeof(double));
                lam = fmin(lam * 4.0, 1e2);
                continue;
            }
            chi_prev = chiw;
            best_chi = chiw;
            memcpy(Tbest, Tcw, 16 * nk * sizeof(double));
            memcpy(Wbest, Twc, 16 * nk * sizeof(double));
            memcpy(rbest, rho, nl * sizeof(double));
            memcpy(mbest, mb, 3 * nl * sizeof(double));
        }
        if (q->pr_m0) { /* ba.py:1250-1264: bearing prior unrobust, rho prior Huber */
            for (int j = 0; j < nl; j++) {
                if (!(q->pr_valid[j] > 0.5)) continue;
                double rp[3] = {(mb[3 * j] - q->pr_m0[3 * j]) * q->pr_wbear[j],
                                (mb[3 * j + 1] - q->pr_m0[3 * j + 1]) * q->pr_wbear[j], (rho[j] - q->pr_rho0[j]) * q->pr_invsig[j]};
                double J0[3] = {q->pr_wbear[j], q->pr_wbear[j], q->pr_invsig[j]};
                for (int c = 0; c < 3; c++) {
                    double wp = fabs(rp[c]) <= K ? 1.0 : K / fmax(fabs(rp[c]), 1e-9);
                    double swp = c < 2 ? 1.0 : sqrt(wp), Jpr = J0[c] * swp;
                    Hll[9 * j + 4 * c] += Jpr * Jpr;
                    bl[3 * j + c] -= Jpr * rp[c] * swp;
                }
            }
        }
        if (q->n_rot) rot_blocks(q->n_rot, q->rot_ab, q->rot_M, q->rot_sig, Tcw, n, Hpp, bp);
        int singular = 0;
        for (int j = 0; j < nl; j++) {
            double tr = Hll[9 * j] + Hll[9 * j + 4] + Hll[9 * j + 8], d = lam * fmax(tr / 3.0, 1e-9);
            for (int i = 0; i < 3; i++) Hll[9 * j + 4 * i] += d;
            if (!inv3(Hll + 9 * j, Hlli + 9 * j)) singular = 1;
        }
        if (singular) break; /* np.linalg.inv raises LinAlgError out of the solver */
        memcpy(Hred, Hpp, (size_t)n * n * sizeof(double));
        for (int j = 0; j < nl; j++) {
            const double *Hi = Hlli + 9 * j;
            for (int ai = bounds[j]; ai < bounds[j + 1]; ai++) {
                int ka = order[ai], sla[2] = {q->okf[ka], q->hidx[j]};
                for (int sa = 0; sa < 2; sa++) {
                    if (sla[sa] < 0) continue;
                    const double *Ha = Hpl + 36 * ka + 18 * sa;
                    double T3[18];
                    for (int i = 0; i < 6; i++)
                        for (int c = 0; c < 3; c++) T3[3 * i + c] = Ha[3 * i] * Hi[c] + Ha[3 * i + 1] * Hi[3 + c] + Ha[3 * i + 2] * Hi[6 + c];
                    for (int i = 0; i < 6; i++)
                        bp[6 * sla[sa] + i] -= T3[3 * i] * bl[3 * j] + T3[3 * i + 1] * bl[3 * j + 1] + T3[3 * i + 2] * bl[3 * j + 2];
                    for (int bi = bounds[j]; bi < bounds[j + 1]; bi++) {
                        int kb = order[bi], slb[2] = {q->okf[kb], q->hidx[j]};
                        for (int sb = 0; sb < 2; sb++) {
                            if (slb[sb] < 0) continue;
                            const double *Hb = Hpl + 36 * kb + 18 * sb;
                            for (int i = 0; i < 6; i++)
                                for (int c = 0; c < 6; c++)
                                    Hred[(size_t)(6 * sla[sa] + i) * n + 6 * slb[sb] + c] -=
                                        T3[3 * i] * Hb[3 * c] + T3[3 * i + 1] * Hb[3 * c + 1] + T3[3 * i + 2] * Hb[3 * c + 2];
                        }
                    }
                }
            }
        }
        for (int i = 0; i < nk; i++) {
            if (q->free_mask[i]) continue;
            int f0 = 6 * i;
            for (int a = 0; a < 6; a++)
                for (int c = 0; c < n; c++) {
                    Hred[(size_t)(f0 + a) * n + c] = 0.0;
                    Hred[(size_t)c * n + f0 + a] = 0.0;
                }
            for (int a = 0; a < 6; a++) {
                Hred[(size_t)(f0 + a) * n + f0 + a] = 1.0;
                bp[f0 + a] = 0.0;
            }
        }
        for (int i = 0; i < n; i++) Hred[(size_t)i * n + i] += lam * fmax(Hred[(size_t)i * n + i], 1e-9);
        int solved;
        if (pol) { /* the GPU Cholesky; its status is ignored there, a failure shows as a non-finite step */
            solved = chol(Hred, n);
            if (solved) {
                double *x = malloc(n * sizeof(double));
                chol_solve(Hred, n, bp, x);
                memcpy(bp, x, n * sizeof(double));
                free(x);
            }
        } else {
            solved = lu_solve(Hred, n, bp);
            if (!solved) break;
        }
        double *dxp = bp;
        if (dx_hist)
            for (int i = 0; i < n; i++) dx_hist[(size_t)it * n + i] = dxp[i];
        for (int j = 0; j < nl; j++) {
            double rhs[3] = {bl[3 * j], bl[3 * j + 1], bl[3 * j + 2]};
            for (int ai = bounds[j]; ai < bounds[j + 1]; ai++) {
                int k = order[ai], sl[2] = {q->okf[k], q->hidx[j]};
                for (int sa = 0; sa < 2; sa++) {
                    if (sl[sa] < 0) continue;
                    const double *Hk = Hpl + 36 * k + 18 * sa, *dk = dxp + 6 * sl[sa];
                    for (int c = 0; c < 3; c++) {
                        double s = 0.0;
                        for (int i = 0; i < 6; i++) s += Hk[3 * i + c] * dk[i];
                        rhs[c] += -s;
                    }
                }
            }
            const double *Hi = Hlli + 9 * j;
            for (int i = 0; i < 3; i++) dxl[3 * j + i] = Hi[3 * i] * rhs[0] + Hi[3 * i + 1] * rhs[1] + Hi[3 * i + 2] * rhs[2];
        }
        if (pol) { /* ba.py:1468-1478: non-finite -> lam x10; else apply (no projection), lam / 2 */
            int fin = solved;
            for (int i = 0; i < n && fin; i++) fin = isfinite(dxp[i]);
            for (int i = 0; i < 3 * nl && fin; i++) fin = isfinite(dxl[i]);
            if (!fin) {
                if (dx_hist)
                    for (int i = 0; i < n; i++) dx_hist[(size_t)it * n + i] = NAN;
                lam = fmin(lam * 10.0, 1e3);
                continue;
            }
            for (int i = 0; i < nk; i++) {
                if (!q->free_mask[i]) continue;
                double E[16], T[16];
                o_se3_exp(dxp + 6 * i, E);
                for (int a = 0; a < 4; a++)
                    for (int c = 0; c < 4; c++) {
                        double sum = 0.0;
                        for (int m = 0; m < 4; m++) sum += E[4 * a + m] * Tcw[16 * i + 4 * m + c];
                        T[4 * a + c] = sum;
                    }
                memcpy(Tcw + 16 * i, T, sizeof(T));
                se3_inv(Tcw + 16 * i, Twc + 16 * i);
            }
            for (int j = 0; j < nl; j++) {
                mb[3 * j] += dxl[3 * j];
                mb[3 * j + 1] += dxl[3 * j + 1];
                rho[j] = fmax(rho[j] + dxl[3 * j + 2], q->rho_floor);
            }
            lam = fmax(lam * 0.5, 1e-7);
            if (acc_hist) acc_hist[it] = 1;
            continue;
        }
        memcpy(Tn, Tcw, 16 * nk * sizeof(double));
        for (int i = 0; i < nk; i++) {
            if (!q->free_mask[i]) continue;
            double E[16], R[9], Rp[9];
            o_se3_exp(dxp + 6 * i, E);
            for (int a = 0; a < 4; a++)
                for (int c = 0; c < 4; c++) {
                    double s = 0.0;
                    for (int m = 0; m < 4; m++) s += E[4 * a + m] * Tcw[16 * i + 4 * m + c];
                    Tn[16 * i + 4 * a + c] = s;
                }
            for (int a = 0; a < 3; a++)
                for (int c = 0; c < 3; c++) R[3 * a + c] = Tn[16 * i + 4 * a + c];
            o_so3_project(R, Rp);
            for (int a = 0; a < 3; a++)
                for (int c = 0; c < 3; c++) Tn[16 * i + 4 * a + c] = Rp[3 * a + c];
        }
        for (int i = 0; i < nk; i++) se3_inv(Tn + 16 * i, Twn + 16 * i);
        for (int j = 0; j < nl; j++) {
            mbn[3 * j] = mb[3 * j] + dxl[3 * j];
            mbn[3 * j + 1] = mb[3 * j + 1] + dxl[3 * j + 1];
            mbn[3 * j + 2] = mb[3 * j + 2];
            rhon[j] = fmax(rho[j] + dxl[3 * j + 2], q->rho_floor);
        }
        double chi_new = id_residuals(q, Twn, Tn, rhon, mbn, r2, NULL, NULL);
        if (chi_new <= chi_cur) {
            memcpy(Tcw, Tn, 16 * nk * sizeof(double));
            memcpy(Twc, Twn, 16 * nk * sizeof(double));
            memcpy(rho, rhon, nl * sizeof(double));
            memcpy(mb, mbn, 3 * nl * sizeof(double));
            lam = fmax(lam * 0.5, 1e-7);
            if (acc_hist) acc_hist[it] = 1;
        } else {
            lam = fmin(lam * 4.0, 1e2);
        }
    }
    double chin = id_residuals(q, Twc, Tcw, rho, mb, r, NULL, NULL);
    stats[0] = chi0;
    stats[1] = chin;
    stats[2] = rms0;
    stats[3] = N ? sqrt(chin / N) : NAN;
    if (pol) { /* ba.py:1480-1497: one SO(3) projection, then the Huber-weighted chi2 of the result */
        for (int i = 0; i < nk; i++) {
            double R[9], Rp[9];
            for (int a = 0; a < 3; a++)
                for (int c = 0; c < 3; c++) R[3 * a + c] = Tcw[16 * i + 4 * a + c];
            o_so3_project(R, Rp);
            for (int a = 0; a < 3; a++)
                for (int c = 0; c < 3; c++) Tcw[16 * i + 4 * a + c] = Rp[3 * a + c];
            se3_inv(Tcw + 16 * i, Twc + 16 * i);
        }
        id_residuals(q, Twc, Tcw, rho, mb, r, NULL, NULL);
        double cw = 0.0;
        for (int k = 0; k < N; k++) {
            double e2 = r[2 * k] * r[2 * k] + r[2 * k + 1] * r[2 * k + 1], e = sqrt(e2);
            double w = (e <= K) ? 1.0 : K / fmax(e, 1e-9);
            cw += w * e2;
        }
        stats[0] = best_chi > 0.0 ? best_chi : cw;
        stats[1] = cw;
        stats[2] = rms0_w;
        stats[3] = sqrt(cw / (N > 1 ? N : 1));
    }
    memcpy(q->poses, Twc, 16 * nk * sizeof(double));
    free(Tcw); free(Tn); free(Twc); free(Twn); free(r); free(Pc); free(Ph); free(r2); free(J); free(Hpl);
    free(rhon); free(mbn); free(Hpp); free(Hred); free(bp); free(Hll); free(Hlli); free(bl); free(dxl);
    free(order); free(bounds); free(Tbest); free(Wbest); free(rbest); free(mbest);
    return 0;
}

/* ================================================================ IMU preintegration (vio/imu_factor.py) ===== */
/* info vector and status codes of vio/imu_factor.py:360-372 */
enum { PI_STATUS, PI_NSUB, PI_NINNER, PI_EXTRAP, PI_TEND, PI_GAPMAX, PI_NREL, PI_TNOM, PI_SMAX, PI_NINFO };
enum { PI_OK = 0, PI_BAD_INTERVAL = 1, PI_NO_DATA = 2, PI_BEFORE_FIRST = 3, PI_NONFINITE = 4, PI_NO_RATE = 5 };

/* Y = F X for the sub-interval transition F (vio/imu_factor.py:366-389), X and Y 15x15 row-major, order
   [phi, v, p, eps_g, eps_a]; B[0..7] = Rt, F1, F2, Bg0, Bg1, Bg2, Ba1, Ba2 (3x3 row-major each). */
static void pre_fleft(double *const B[8], double h, const double *X, double *Y) {
    const double *Rt = B[0], *F1 = B[1], *F2 = B[2], *G0 = B[3], *G1 = B[4], *G2 = B[5], *A1 = B[6], *A2 = B[7];
    for (int c = 0; c < 15; c++)
        for (int i = 0; i < 3; i++) {
            double s0 = 0.0, s1 = 0.0, s2 = 0.0;
            for (int k = 0; k < 3; k++) {
                double xph = X[k * 15 + c], xg = X[(9 + k) * 15 + c], xa = X[(12 + k) * 15 + c];
                s0 += Rt[3 * i + k] * xph + G0[3 * i + k] * xg;
                s1 += F1[3 * i + k] * xph + G1[3 * i + k] * xg + A1[3 * i + k] * xa;
                s2 += F2[3 * i + k] * xph + G2[3 * i + k] * xg + A2[3 * i + k] * xa;
            }
            Y[i * 15 + c] = s0;
            Y[(3 + i) * 15 + c] = s1 + X[(3 + i) * 15 + c];
            Y[(6 + i) * 15 + c] = s2 + h * X[(3 + i) * 15 + c] + X[(6 + i) * 15 + c];
            Y[(9 + i) * 15 + c] = X[(9 + i) * 15 + c];
            Y[(12 + i) * 15 + c] = X[(12 + i) * 15 + c];
        }
}

/* vio/imu_factor.py:392-542 _step: advance (dR, dv, dp), the bias Jacobians J (9x6 row-major: rows phi v p, cols
   bg ba) and the internal covariance P (15x15, errors computed minus true, eps = bias deviation) over one
   sub-interval of length h with the bias-corrected rates w, a and the (already scaled) white densities. */
static void pre_step(double h, const double *w, const double *a, double sg2, double sa2, double swg2, double swa2,
                     double *dR, double *dv, double *dp, double *J, double *P) {
    double wh[3] = {w[0] * h, w[1] * h, w[2] * h}, wm[3] = {0.5 * wh[0], 0.5 * wh[1], 0.5 * wh[2]};
    double dRk[9], Jrf[9], Rh[9], Jrh[9], dRm[9], am[3], M[9], N[9], T[9], MMt[9];
    o_exp_jr(wh, dRk, Jrf);
    o_exp_jr(wm, Rh, Jrh);
    mm3(dR, Rh, dRm); /* the midpoint attitude */
    for (int i = 0; i < 3; i++) {
        const double *r = dRm + 3 * i;
        am[i] = r[0] * a[0] + r[1] * a[1] + r[2] * a[2];
        M[3 * i + 0] = r[1] * a[2] - r[2] * a[1]; /* M = dRm [a]x */
        M[3 * i + 1] = -r[0] * a[2] + r[2] * a[0];
        M[3 * i + 2] = r[0] * a[1] - r[1] * a[0];
    }
    mmt3(M, Rh, N);
    mm3(M, Jrh, T);
    double h2 = h * h, h3 = h2 * h, h4 = h3 * h, h5 = h4 * h, h6 = h5 * h, h7 = h6 * h;
    double Rt[9], F1[9], F2[9], G0[9], G1[9], G2[9], A1[9], A2[9];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            int e = 3 * i + j;
            Rt[e] = dRk[3 * j + i];
            F1[e] = -N[e] * h;
            F2[e] = -0.5 * N[e] * h2;
            G0[e] = Jrf[e] * h;
            G1[e] = -T[e] * (0.5 * h2);
            G2[e] = -T[e] * (0.25 * h3);
            A1[e] = dRm[e] * h;
            A2[e] = 0.5 * dRm[e] * h2;
        }
    double *const B[8] = {Rt, F1, F2, G0, G1, G2, A1, A2};
    double Y[225], Z[225];
    pre_fleft(B, h, P, Y); /* F P */
    for (int i = 0; i < 15; i++)
        for (int j = 0; j < 15; j++) Z[i * 15 + j] = Y[j * 15 + i];
    pre_fleft(B, h, Z, Y); /* F P F^T */
    for (int i = 0; i < 15; i++) {
        P[i * 15 + i] = Y[i * 15 + i];
        for (int j = i + 1; j < 15; j++) P[i * 15 + j] = P[j * 15 + i] = 0.5 * (Y[i * 15 + j] + Y[j * 15 + i]);
    }
    mmt3(M, M, MMt);
    /* continuous white noise and bias drift through the sub-interval, closed forms (imu_factor.py:459-500): every
       entry of P receives exactly one term */
    double qpp = sg2 * h + swg2 * h3 / 3.0, qvp1 = sg2 * h2 / 2.0 + swg2 * h4 / 8.0;
    double qpp1 = sg2 * h3 / 6.0 + swg2 * h5 / 30.0;
    double qvv_i = sa2 * h + swa2 * h3 / 3.0, qvv_m = sg2 * h3 / 3.0 + swg2 * h5 / 20.0;
    double qpv_i = sa2 * h2 / 2.0 + swa2 * h4 / 8.0, qpv_m = sg2 * h4 / 8.0 + swg2 * h6 / 72.0;
    double qpp_i = sa2 * h3 / 3.0 + swa2 * h5 / 20.0, qpp_m = sg2 * h5 / 20.0 + swg2 * h7 / 252.0;
#define PADD(r, c, v) P[(r) * 15 + (c)] += (v)
#define PADD2(r, c, v) do { double q_ = (v); PADD(r, c, q_); PADD(c, r, q_); } while (0)
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            double d = i == j ? 1.0 : 0.0, mm = MMt[3 * i + j], mij = M[3 * i + j], rij = dRm[3 * i + j];
            PADD(i, j, qpp * d);
            PADD2(3 + i, j, -mij * qvp1);
            PADD2(6 + i, j, -mij * qpp1);
            PADD(3 + i, 3 + j, qvv_i * d + qvv_m * mm);
            PADD2(6 + i, 3 + j, qpv_i * d + qpv_m * mm);
            PADD(6 + i, 6 + j, qpp_i * d + qpp_m * mm);
            PADD2(i, 9 + j, swg2 * h2 / 2.0 * d);
            PADD2(3 + i, 9 + j, -mij * swg2 * h3 / 6.0);
            PADD2(6 + i, 9 + j, -mij * swg2 * h4 / 24.0);
            PADD2(3 + i, 12 + j, rij * swa2 * h2 / 2.0);
            PADD2(6 + i, 12 + j, rij * swa2 * h3 / 6.0);
            PADD(9 + i, 9 + j, swg2 * h * d);
            PADD(12 + i, 12 + j, swa2 * h * d);
        }
#undef PADD
#undef PADD2
    /* bias Jacobians, the exact derivatives of the discrete scheme: J <- A9 J - B (imu_factor.py:502-520) */
    double Jn[54];
    for (int c = 0; c < 6; c++)
        for (int i = 0; i < 3; i++) {
            double s0 = 0.0, s1 = 0.0, s2 = 0.0;
            for (int k = 0; k < 3; k++) {
                double jp = J[k * 6 + c];
                s0 += Rt[3 * i + k] * jp;
                s1 += F1[3 * i + k] * jp;
                s2 += F2[3 * i + k] * jp;
            }
            Jn[i * 6 + c] = s0;
            Jn[(3 + i) * 6 + c] = s1 + J[(3 + i) * 6 + c];
            Jn[(6 + i) * 6 + c] = s2 + h * J[(3 + i) * 6 + c] + J[(6 + i) * 6 + c];
        }
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            Jn[i * 6 + j] -= G0[3 * i + j];
            Jn[(3 + i) * 6 + j] -= G1[3 * i + j];
            Jn[(6 + i) * 6 + j] -= G2[3 * i + j];
            Jn[(3 + i) * 6 + 3 + j] -= A1[3 * i + j];
            Jn[(6 + i) * 6 + 3 + j] -= A2[3 * i + j];
        }
    memcpy(J, Jn, sizeof(Jn));
    /* the mean */
    for (int i = 0; i < 3; i++) {
        dp[i] = dp[i] + dv[i] * h + 0.5 * am[i] * h2;
        dv[i] = dv[i] + am[i] * h;
    }
    double Rn[9];
    mm3(dR, dRk, Rn);
    memcpy(dR, Rn, sizeof(Rn));
}

static int finite6(const double *w, const double *a) {
    return isfinite(w[0]) && isfinite(w[1]) && isfinite(w[2]) && isfinite(a[0]) && isfinite(a[1]) && isfinite(a[2]);
}

/* vio/imu_factor.py:545-704 preint_core: preintegrate [t0, t1] from the samples with t <= horizon (stamps t (n,),
   gyr / acc (n,3) row-major; bg, ba the linearisation biases; squared densities; tnom the nominal sample period).
   Outputs dR (9), dv, dp (3), J (9x6), C (15x15 residual order [rR rv rp rbg rba]) and info (PI_NINFO); nothing but
   info is meaningful unless info[PI_STATUS] == PI_OK. */
void o_preint(const double *t, const double *gyr, const double *acc, int n, double t0, double t1, double horizon,
              const double *bg, const double *ba, double sg2, double sa2, double swg2, double swa2, double tnom,
              double *dR, double *dv, double *dp, double *J, double *C, double *info) {
    for (int i = 0; i < 9; i++) dR[i] = (i % 4 == 0) ? 1.0 : 0.0;
    for (int i = 0; i < 3; i++) dv[i] = dp[i] = 0.0;
    memset(J, 0, 54 * sizeof(double));
    memset(C, 0, 225 * sizeof(double));
    memset(info, 0, PI_NINFO * sizeof(double));
    int lo = 0, hi = n;
    while (lo < hi) { /* released count, from the stamps alone */
        int mid = (lo + hi) / 2;
        if (t[mid] <= horizon) lo = mid + 1; else hi = mid;
    }
    const int nrel = lo;
    info[PI_NREL] = nrel;
    if (!(t1 > t0) || !isfinite(t0) || !isfinite(t1)) { info[PI_STATUS] = PI_BAD_INTERVAL; return; }
    if (nrel == 0) { info[PI_STATUS] = PI_NO_DATA; return; }
    if (t0 < t[0]) { info[PI_STATUS] = PI_BEFORE_FIRST; return; }
    const int rate_ok = tnom > 0.0 && isfinite(tnom);
    if ((sg2 > 0.0 || sa2 > 0.0) && !rate_ok) { info[PI_STATUS] = PI_NO_RATE; return; }
    info[PI_TNOM] = rate_ok ? tnom : 0.0;
    const double te = t[nrel - 1];
    info[PI_TEND] = te;
    double P[225], w[3], a[3], smax = 1.0, gapmax = 0.0;
    memset(P, 0, sizeof(P));
    int nsub = 0, ninner = 0;
    if (te > t0) {
        const double tie = t1 < te ? t1 : te;
        lo = 0;
        hi = nrel;
        while (lo < hi) { /* first released stamp > t0 */
            int mid = (lo + hi) / 2;
            if (t[mid] <= t0) lo = mid + 1; else hi = mid;
        }
        int k = lo, left = k - 1;
        double e0 = t0;
        for (;;) {
            const int samp = k < nrel && t[k] < tie;
            const double e1 = samp ? t[k] : tie, h = e1 - e0;
            if (h > 0.0) {
                const double m = 0.5 * (e0 + e1);
                while (left + 2 < nrel && t[left + 1] <= m) left++; /* bracket, bounded at the last released pair */
                const double ta = t[left], gap = t[left + 1] - ta;
                if (gap > gapmax) gapmax = gap;
                for (int c = 0; c < 3; c++) {
                    double s = (gyr[3 * (left + 1) + c] - gyr[3 * left + c]) / gap;
                    w[c] = s * (m - ta) + gyr[3 * left + c] - bg[c];
                    s = (acc[3 * (left + 1) + c] - acc[3 * left + c]) / gap;
                    a[c] = s * (m - ta) + acc[3 * left + c] - ba[c];
                }
                if (!finite6(w, a)) { info[PI_STATUS] = PI_NONFINITE; return; }
                double sw = 1.0; /* sample-support rule: max(1, nu h / T_nom) */
                if (rate_ok) {
                    const double u = (m - ta) / gap;
                    sw = 2.0 * ((1.0 - u) * (1.0 - u) + u * u) * h / tnom;
                    if (sw < 1.0) sw = 1.0;
                    if (sw > smax) smax = sw;
                }
                pre_step(h, w, a, sg2 * sw, sa2 * sw, swg2, swa2, dR, dv, dp, J, P);
                nsub++;
            }
            if (!samp) break;
            ninner++;
            e0 = e1;
            k++;
        }
    }
    if (t1 > te) { /* hold the last released sample over (t_end, t1] */
        const double ts = t0 > te ? t0 : te, h = t1 - ts;
        for (int c = 0; c < 3; c++) {
            w[c] = gyr[3 * (nrel - 1) + c] - bg[c];
            a[c] = acc[3 * (nrel - 1) + c] - ba[c];
        }
        if (!finite6(w, a)) { info[PI_STATUS] = PI_NONFINITE; return; }
        double sw = 1.0;
        if (rate_ok) {
            sw = (te > t0 ? 2.0 : 1.0) * h / tnom;
            if (sw < 1.0) sw = 1.0;
            if (sw > smax) smax = sw;
        }
        pre_step(h, w, a, sg2 * sw, sa2 * sw, swg2, swa2, dR, dv, dp, J, P);
        nsub++;
        if (te > t0) ninner++;
        info[PI_EXTRAP] = h;
    }
    double Tm[9], T2[9], Rn[9]; /* one Newton-Schulz polar step */
    mtm3(dR, dR, Tm);
    for (int i = 0; i < 9; i++) T2[i] = (i % 4 == 0 ? 1.5 : 0.0) - 0.5 * Tm[i];
    mm3(dR, T2, Rn);
    memcpy(dR, Rn, sizeof(Rn));
    for (int i = 0; i < 15; i++) /* residual order: r9 = -(computed - true), r_b = +eps: the cross block flips */
        for (int j = 0; j < 15; j++) C[i * 15 + j] = ((i < 9) != (j < 9)) ? -P[i * 15 + j] : P[i * 15 + j];
    info[PI_NSUB] = nsub;
    info[PI_NINNER] = ninner;
    info[PI_GAPMAX] = gapmax;
    info[PI_SMAX] = smax;
    info[PI_STATUS] = PI_OK;
}

/* vio/imu_factor.py:707-746 sqrt_info_core: W = L^-1 (lower), L L^T = kappa C + J_bg sbg2 J_bg^T + J_ba sba2 J_ba^T
   on the 9 preintegration rows.  Returns 0 (W undefined) when the matrix is not positive definite. */
int o_sqrt_info(const double *C, double kappa, double sbg2, double sba2, const double *J, double *W) {
    double A[225], L[225];
    memset(L, 0, sizeof(L));
    for (int i = 0; i < 225; i++) A[i] = kappa * C[i];
    if (sbg2 > 0.0 || sba2 > 0.0)
        for (int i = 0; i < 9; i++)
            for (int j = 0; j < 9; j++) {
                double s = 0.0;
                for (int k = 0; k < 3; k++) s += sbg2 * J[i * 6 + k] * J[j * 6 + k] + sba2 * J[i * 6 + 3 + k] * J[j * 6 + 3 + k];
                A[i * 15 + j] += s;
            }
    for (int j = 0; j < 15; j++) {
        double s = A[j * 15 + j];
        for (int k = 0; k < j; k++) s -= L[j * 15 + k] * L[j * 15 + k];
        if (!(s > 0.0)) return 0;
        const double dj = sqrt(s);
        L[j * 15 + j] = dj;
        for (int i = j + 1; i < 15; i++) {
            s = A[i * 15 + j];
            for (int k = 0; k < j; k++) s -= L[i * 15 + k] * L[j * 15 + k];
            L[i * 15 + j] = s / dj;
        }
    }
    memset(W, 0, 225 * sizeof(double));
    for (int j = 0; j < 15; j++) {
        W[j * 15 + j] = 1.0 / L[j * 15 + j];
        for (int i = j + 1; i < 15; i++) {
            double s = 0.0;
            for (int k = j; k < i; k++) s += L[i * 15 + k] * W[k * 15 + j];
            W[i * 15 + j] = -s / L[i * 15 + i];
        }
    }
    return 1;
}
