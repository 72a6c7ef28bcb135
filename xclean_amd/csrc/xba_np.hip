// The local-BA drop-in: backend/ba.py:713-869 `_solve_np` (Euclidean world points, poses T_cw perturbed on the left
// T <- exp(xi) T with xi = [rho, phi], Huber IRLS on the pixel norm, Marquardt damping of the 3x3 landmark blocks by
// lam * max(tr/3, 1e-9), Schur complement, held slots, damping of the reduced diagonal by lam * max(d, 1e-9), SE(3)
// update + projection onto SO(3), accept iff the unrobust chi2 did not increase: lam = max(lam/2, 1e-7), else
// lam = min(4 lam, 1e2)), in FP64 on the device.  The accept/reject decision and lambda live on the device; the whole
// fixed-iteration loop is one stream of kernels with no host round trip.
#include "xba_internal.h"

namespace xba {

struct NpArgs {
    int32_t nk, nl, N, n_rot, has_prior;
    double fx, cx, cy, K;
    double *Tcw, *Tnew;            // (nk,16) row-major
    double *pts, *pnew;            // (nl,3)
    const int32_t *okf, *olm;      // (N)
    const double *ouv;             // (N,2)
    const uint8_t *free_mask;      // (nk)
    const int32_t *order, *lstart; // observations sorted by landmark (stable), CSR offsets (nl+1)
    const double *pr_c, *pr_rho0, *pr_invsig, *pr_valid;
    const int32_t *rot_ab;
    const double *rot_M, *rot_sig;
    double *J;                     // (N,20): Jpose 12, Jlm 6, rw 2 (sqrt(w)-scaled)
    double *r2;                    // (N) squared residual (unrobust)
    double *Hlli, *bl;             // (nl,9), (nl,3)
    double *H, *bp, *dx;           // (n,n), (n), (n)
    double *st;                    // device scalars: [0] lam [1] chi_cur [2] chi_new [3] accept [4] ok [5] chi0 [6] chi_n
    double *dx_hist;               // (iters, n) or null
    // the inverse-depth solve (backend/ba.py:1212-1342); J is then (N,32): Jobs 12, Jhost 12, Jlm 6, rw 2
    int32_t rot_h;                 // 1: the rotation factors' Hessian blocks enter H as well as their gradient
    const int32_t *hidx;           // (nl) host keyframe, or -1 (host pose from hext)
    const double *hext;            // (nl,12) host T_wc rows [R | t] when hidx < 0
    double *Twc, *Twn;             // (nk,16) T_wc of the current and the trial poses
    double *rho, *mb, *rhon, *mbn; // (nl), (nl,3) inverse depth and host-frame bearing, current and trial
    double rho_floor;
    const double *pr_m0, *pr_wbear; // bearing prior: (nl,3) mean, (nl) weight
    // LM policy: 0 the numpy solves (trial cost, accept iff the unrobust chi2 did not grow); 1 the GPU drivers
    // _solve_gpu / _solve_gpu_id (ba.py:1015-1076, 1436-1483): the step is applied at once, the next iteration rejects
    // it when the Huber-weighted chi2 grew (restoring the last accepted state, lam x4), a non-finite step sets lam
    // x10, SO(3) is projected once at the end
    int32_t policy, jstride, rwoff;  // J row stride and offset of the weighted residual (the chi2 of the gate)
    double *Tbest, *pbest, *Wbest, *rbest, *mbest;  // the last accepted state (policy 1)
};

// device scalars beyond [0] lam [1] chi_cur [2] chi_new [3] accept [4] ok [5] chi0 [6] chi_n
enum { ST_SKIP = 8, ST_NONFIN = 9, ST_CHIPREV = 10, ST_BESTCHI = 11, ST_RMS0 = 12, NST = 16 };
__device__ inline bool np_skip(const NpArgs &A) { return A.policy && A.st[ST_SKIP] != 0.0; }

__device__ inline void np_residual(const NpArgs &A, const double *T, const double *X, const double *uv, double *P,
                                   double *r) {
    for (int i = 0; i < 3; i++) P[i] = (T[4 * i] * X[0] + T[4 * i + 1] * X[1] + T[4 * i + 2] * X[2]) + T[4 * i + 3];
    const double z = fmax(P[2], 1e-6);
    r[0] = A.fx * P[0] / z + A.cx - uv[0];
    r[1] = A.fx * P[1] / z + A.cy - uv[1];
}

// Residual, Huber weight and sqrt(w)-scaled Jacobians of every observation (ba.py:723-757).
__global__ void k_np_obs(NpArgs A) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= A.N) return;
    const double *T = A.Tcw + 16 * (size_t)A.okf[k], *X = A.pts + 3 * (size_t)A.olm[k];
    double P[3], r[2];
    np_residual(A, T, X, A.ouv + 2 * (size_t)k, P, r);
    A.r2[k] = r[0] * r[0] + r[1] * r[1];
    const double e = sqrt(r[0] * r[0] + r[1] * r[1]);
    const double w = (e <= A.K) ? 1.0 : A.K / fmax(e, 1e-9), sw = sqrt(w);
    const double x = P[0], y = P[1], zz = fmax(P[2], 1e-6), iz = 1.0 / zz, iz2 = iz * iz, fx = A.fx;
    const double Jp3[2][3] = {{fx * iz, 0.0, -fx * x * iz2}, {0.0, fx * iz, -fx * y * iz2}};
    const double dP[3][6] = {{1, 0, 0, 0, zz, -y}, {0, 1, 0, -zz, 0, x}, {0, 0, 1, y, -x, 0}};
    double *J = A.J + 20 * (size_t)k;
    for (int a = 0; a < 2; a++) {
        for (int c = 0; c < 6; c++) J[6 * a + c] = (Jp3[a][0] * dP[0][c] + Jp3[a][1] * dP[1][c] + Jp3[a][2] * dP[2][c]) * sw;
        for (int c = 0; c < 3; c++) J[12 + 3 * a + c] = (Jp3[a][0] * T[c] + Jp3[a][1] * T[4 + c] + Jp3[a][2] * T[8 + c]) * sw;
        J[18 + a] = r[a] * sw;
    }
}

// Per landmark (observations in original order): Hll, bl, the inverse-range prior, damping, inverse (ba.py:759-797).
__global__ void k_np_lm(NpArgs A) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= A.nl || np_skip(A)) return;
    const double lam = A.st[0];
    double Hll[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, bl[3] = {0, 0, 0};
    for (int q = A.lstart[j]; q < A.lstart[j + 1]; q++) {
        const double *J = A.J + 20 * (size_t)A.order[q];
        for (int i = 0; i < 3; i++) {
            for (int c = 0; c < 3; c++) Hll[3 * i + c] += J[12 + i] * J[12 + c] + J[15 + i] * J[15 + c];
            bl[i] += -(J[12 + i] * J[18] + J[15 + i] * J[19]);
        }
    }
    if (A.has_prior) {  // ba.py:774-794 (the reference applies it when any landmark is valid; rows of invalid ones are 0)
        const double *X = A.pts + 3 * (size_t)j, *c = A.pr_c + 3 * (size_t)j;
        const double dP[3] = {X[0] - c[0], X[1] - c[1], X[2] - c[2]};
        const double dist = sqrt(dP[0] * dP[0] + dP[1] * dP[1] + dP[2] * dP[2]);
        const bool ok = A.pr_valid[j] > 0.5 && dist > 1e-6;
        if (ok) {
            const double rho = 1.0 / dist, rp = (rho - A.pr_rho0[j]) * A.pr_invsig[j];
            const double wp = fabs(rp) <= A.K ? 1.0 : A.K / fmax(fabs(rp), 1e-9), swp = sqrt(wp);
            const double gg = -rho * rho / dist * A.pr_invsig[j], rpw = rp * swp;
            double Jp_[3];
            for (int i = 0; i < 3; i++) Jp_[i] = dP[i] * (gg * swp);
            for (int i = 0; i < 3; i++) {
                for (int c2 = 0; c2 < 3; c2++) Hll[3 * i + c2] += Jp_[i] * Jp_[c2];
                bl[i] -= Jp_[i] * rpw;
            }
        }
    }
    const double tr = Hll[0] + Hll[4] + Hll[8], d = lam * fmax(tr / 3.0, 1e-9);
    for (int i = 0; i < 3; i++) Hll[4 * i] += d;
    const double c00 = Hll[4] * Hll[8] - Hll[5] * Hll[7], c01 = Hll[5] * Hll[6] - Hll[3] * Hll[8],
                 c02 = Hll[3] * Hll[7] - Hll[4] * Hll[6];
    const double det = Hll[0] * c00 + Hll[1] * c01 + Hll[2] * c02, id = 1.0 / det;
    double *Hi = A.Hlli + 9 * (size_t)j;
    Hi[0] = c00 * id; Hi[1] = (Hll[2] * Hll[7] - Hll[1] * Hll[8]) * id; Hi[2] = (Hll[1] * Hll[5] - Hll[2] * Hll[4]) * id;
    Hi[3] = c01 * id; Hi[4] = (Hll[0] * Hll[8] - Hll[2] * Hll[6]) * id; Hi[5] = (Hll[2] * Hll[3] - Hll[0] * Hll[5]) * id;
    Hi[6] = c02 * id; Hi[7] = (Hll[1] * Hll[6] - Hll[0] * Hll[7]) * id; Hi[8] = (Hll[0] * Hll[4] - Hll[1] * Hll[3]) * id;
    for (int i = 0; i < 3; i++) A.bl[3 * (size_t)j + i] = bl[i];
}

__device__ inline void hpl_of(const double *J, double *Hpl) {  // (6x3) Jpose^T Jlm
    for (int i = 0; i < 6; i++)
        for (int c = 0; c < 3; c++) Hpl[3 * i + c] = J[i] * J[12 + c] + J[6 + i] * J[15 + c];
}

__global__ void k_np_zero(NpArgs A, int n) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n * n; i += gridDim.x * blockDim.x) A.H[i] = 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) A.bp[i] = 0.0;
}

// so3 log via the quaternion and Jr^-1 (for the rotation factors) come from xba_dense.hip.
// One workgroup: chi2, rotation-factor gradient, held slots, damping, Cholesky, solve, trial poses (ba.py:825-856).
// 512 threads.  The damped matrix is factored in LDS (rows padded to a multiple of 4 entries) when it fits
// `lds_cap` bytes of dynamic LDS, else in place in H.
__global__ __launch_bounds__(512) void k_np_solve(NpArgs A, int n, int it, int lds_cap) {
    extern __shared__ __align__(16) double sF[];
    __shared__ double sred[512];
    __shared__ double sxv[XBA_MAX_STATE];
    __shared__ int sflag;
    __shared__ double sinv[CH_SCRATCH];
    const int tid = threadIdx.x, nt = blockDim.x;
    if (np_skip(A)) return;  // (uniform: no kernel of this launch writes the flag)
    // chi2 of the current residuals, fixed-order reduction
    double s = 0.0;
    for (int k = tid; k < A.N; k += nt) s += A.r2[k];
    sred[tid] = s;
    __syncthreads();
    for (int d = nt / 2; d > 0; d >>= 1) {
        if (tid < d) sred[tid] += sred[tid + d];
        __syncthreads();
    }
    if (tid == 0) A.st[1] = sred[0];
    // gyro rotation factors: only their gradient enters (ba.py:644-697, 825-828)
    if (tid == 0)
        for (int f = 0; f < A.n_rot; f++) {
            const int a = A.rot_ab[2 * f], b = A.rot_ab[2 * f + 1];
            const double w = 1.0 / fmax(A.rot_sig[f], 1e-9);
            double Ri[9], Rj[9], Am[9], E[9], e[3], Jr[9], Ja[9], Jb[9], r[3];
            for (int i = 0; i < 3; i++)
                for (int c = 0; c < 3; c++) {
                    Ri[3 * i + c] = A.Tcw[16 * a + 4 * i + c];
                    Rj[3 * i + c] = A.Tcw[16 * b + 4 * i + c];
                }
            mmt3(Ri, Rj, Am);
            mtm3(A.rot_M + 9 * f, Am, E);
            log_quat(E, e);
            jr_inv(e, Jr);
            mmt3(Jr, Am, Ja);
            for (int i = 0; i < 9; i++) {
                Ja[i] *= w;
                Jb[i] = -w * Jr[i];
            }
            for (int i = 0; i < 3; i++) r[i] = w * e[i];
            for (int i = 0; i < 3; i++) {
                double ga = 0.0, gb = 0.0;
                for (int c = 0; c < 3; c++) {
                    ga += Ja[3 * c + i] * r[c];
                    gb += Jb[3 * c + i] * r[c];
                }
                A.bp[6 * a + 3 + i] += -ga;
                A.bp[6 * b + 3 + i] += -gb;
            }
            if (A.rot_h) {  // the inverse-depth solve adds the Hessian blocks too (ba.py:1293-1296)
                const int ia = 6 * a + 3, ib = 6 * b + 3;
                for (int i = 0; i < 3; i++)
                    for (int c = 0; c < 3; c++) {
                        double haa = 0.0, hab = 0.0, hbb = 0.0;
                        for (int m = 0; m < 3; m++) {
                            haa += Ja[3 * m + i] * Ja[3 * m + c];
                            hab += Ja[3 * m + i] * Jb[3 * m + c];
                            hbb += Jb[3 * m + i] * Jb[3 * m + c];
                        }
                        A.H[(size_t)(ia + i) * n + ia + c] += haa;
                        A.H[(size_t)(ia + i) * n + ib + c] += hab;
                        A.H[(size_t)(ib + c) * n + ia + i] += hab;
                        A.H[(size_t)(ib + i) * n + ib + c] += hbb;
                    }
            }
        }
    __syncthreads();
    // held slots: rows and columns leave the solve (ba.py:829-836); damping Hred_ii += lam max(Hred_ii, 1e-9)
    // (ba.py:837-838); the result in the factor's storage
    const int ldp = (n + 3) & ~3;
    const bool in_lds = (size_t)n * ldp * sizeof(double) <= (size_t)lds_cap;
    const int ld = in_lds ? ldp : n;
    double *F = in_lds ? sF : A.H;
    const double lam = A.st[0];
    for (int e = tid; e < n * n; e += nt) {
        const int i = e / n, c = e % n;
        const bool hi_ = !A.free_mask[i / 6], hc = !A.free_mask[c / 6];
        double h = A.H[e];
        if (hi_ || hc) h = (hi_ && hc && i / 6 == c / 6 && i == c) ? 1.0 : 0.0;
        if (i == c) h += lam * fmax(h, 1e-9);
        F[(size_t)i * ld + c] = h;
    }
    for (int i = tid; i < n; i += nt)
        if (!A.free_mask[i / 6]) A.bp[i] = 0.0;
    __syncthreads();
    // Cholesky (the damped reduced system is symmetric positive definite), then solve
    bool ok = in_lds ? chol_wg<double>(sF, n, &sflag, sinv, ld) : chol_wg<double>(A.H, n, &sflag, sinv, ld);
    if (ok) {
        for (int i = tid; i < n; i += nt) sxv[i] = A.bp[i];
        __syncthreads();
        if (in_lds)
            chol_solve_wg<double>(sF, n, sxv, ld);
        else
            chol_solve_wg<double>(A.H, n, sxv, ld);
        for (int i = tid; i < n; i += nt) A.dx[i] = sxv[i];
        __syncthreads();
        if (A.dx_hist)
            for (int i = tid; i < n; i += nt) A.dx_hist[(size_t)it * n + i] = A.dx[i];
        // trial poses: T_new = exp(dx) T (free slots), rotation projected onto SO(3) by polar iteration
        for (int k = tid; k < A.nk; k += nt) {
            double *Tn = A.Tnew + 16 * (size_t)k;
            const double *T = A.Tcw + 16 * (size_t)k;
            if (!A.free_mask[k]) {
                for (int i = 0; i < 16; i++) Tn[i] = T[i];
                continue;
            }
            const double *xi = A.dx + 6 * k;
            double a, b, c, K[9] = {0, -xi[5], xi[4], xi[5], 0, -xi[3], -xi[4], xi[3], 0}, KK[9], R[9], V[9];
            abc(xi[3] * xi[3] + xi[4] * xi[4] + xi[5] * xi[5], &a, &b, &c);
            mm3(K, K, KK);
            for (int i = 0; i < 9; i++) {
                const double e = (i % 4 == 0) ? 1.0 : 0.0;
                R[i] = e + a * K[i] + b * KK[i];
                V[i] = e + b * K[i] + c * KK[i];
            }
            double E[16] = {0};
            for (int i = 0; i < 3; i++) {
                for (int j = 0; j < 3; j++) E[4 * i + j] = R[3 * i + j];
                E[4 * i + 3] = V[3 * i] * xi[0] + V[3 * i + 1] * xi[1] + V[3 * i + 2] * xi[2];
            }
            E[15] = 1.0;
            for (int i = 0; i < 4; i++)
                for (int j = 0; j < 4; j++) {
                    double acc = 0.0;
                    for (int m = 0; m < 4; m++) acc += E[4 * i + m] * T[4 * m + j];
                    Tn[4 * i + j] = acc;
                }
            double X[9];
            for (int i = 0; i < 3; i++)
                for (int j = 0; j < 3; j++) X[3 * i + j] = Tn[4 * i + j];
            for (int itp = 0; itp < (A.policy ? 0 : 30); itp++) {  // the GPU drivers project once, at the end  // X <- (X + X^-T) / 2 converges to the orthogonal polar factor
                double c00 = X[4] * X[8] - X[5] * X[7], c01 = X[5] * X[6] - X[3] * X[8], c02 = X[3] * X[7] - X[4] * X[6];
                double det = X[0] * c00 + X[1] * c01 + X[2] * c02, id = 1.0 / det;
                double Xi[9] = {c00 * id, (X[2] * X[7] - X[1] * X[8]) * id, (X[1] * X[5] - X[2] * X[4]) * id,
                                c01 * id, (X[0] * X[8] - X[2] * X[6]) * id, (X[2] * X[3] - X[0] * X[5]) * id,
                                c02 * id, (X[1] * X[6] - X[0] * X[7]) * id, (X[0] * X[4] - X[1] * X[3]) * id};
                double dmax = 0.0;
                for (int i = 0; i < 3; i++)
                    for (int j = 0; j < 3; j++) {
                        double nv = 0.5 * (X[3 * i + j] + Xi[3 * j + i]);
                        dmax = fmax(dmax, fabs(nv - X[3 * i + j]));
                        X[3 * i + j] = nv;
                    }
                if (dmax < 1e-16) break;
            }
            for (int i = 0; i < 3; i++)
                for (int j = 0; j < 3; j++) Tn[4 * i + j] = X[3 * i + j];
        }
    }
    if (tid == 0) A.st[4] = ok ? 1.0 : 0.0;
}

// Landmark back-substitution dxl = Hlli (bl - sum Hpl^T dxp_k) and the trial points (ba.py:844-856).
__global__ void k_np_backsub(NpArgs A) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= A.nl || A.st[4] == 0.0 || np_skip(A)) return;
    double rhs[3] = {A.bl[3 * (size_t)j], A.bl[3 * (size_t)j + 1], A.bl[3 * (size_t)j + 2]};
    for (int q = A.lstart[j]; q < A.lstart[j + 1]; q++) {
        const int k = A.order[q];
        double Hp[18];
        hpl_of(A.J + 20 * (size_t)k, Hp);
        const double *d = A.dx + 6 * A.okf[k];
        for (int c = 0; c < 3; c++) {
            double s = 0.0;
            for (int i = 0; i < 6; i++) s += Hp[3 * i + c] * d[i];
            rhs[c] += -s;
        }
    }
    const double *Hi = A.Hlli + 9 * (size_t)j;
    for (int i = 0; i < 3; i++) {
        const double dl = Hi[3 * i] * rhs[0] + Hi[3 * i + 1] * rhs[1] + Hi[3 * i + 2] * rhs[2];
        if (!isfinite(dl)) A.st[ST_NONFIN] = 1.0;
        A.pnew[3 * (size_t)j + i] = A.pts[3 * (size_t)j + i] + dl;
    }
}

__global__ void k_np_trial_r2(NpArgs A) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= A.N || A.st[4] == 0.0) return;
    double P[3], r[2];
    np_residual(A, A.Tnew + 16 * (size_t)A.okf[k], A.pnew + 3 * (size_t)A.olm[k], A.ouv + 2 * (size_t)k, P, r);
    A.r2[k] = r[0] * r[0] + r[1] * r[1];
}

// chi_new, accept (chi_new <= chi_cur), lambda rule, commit of the poses (ba.py:858-863).
__global__ __launch_bounds__(256) void k_np_accept(NpArgs A) {
    __shared__ double sred[256];
    __shared__ int sacc;
    const int tid = threadIdx.x, nt = blockDim.x;
    if (A.st[4] == 0.0) {  // a singular reduced system stops the loop (LinAlgError): keep the state, freeze lambda
        if (tid == 0) A.st[3] = 0.0;
        return;
    }
    double s = 0.0;
    for (int k = tid; k < A.N; k += nt) s += A.r2[k];
    sred[tid] = s;
    __syncthreads();
    for (int d = nt / 2; d > 0; d >>= 1) {
        if (tid < d) sred[tid] += sred[tid + d];
        __syncthreads();
    }
    if (tid == 0) {
        const double cn = sred[0];
        A.st[2] = cn;
        const bool acc = cn <= A.st[1];
        A.st[3] = acc ? 1.0 : 0.0;
        A.st[0] = acc ? fmax(A.st[0] * 0.5, 1e-7) : fmin(A.st[0] * 4.0, 1e2);
        sacc = acc;
    }
    __syncthreads();
    if (sacc)
        for (int i = tid; i < 16 * A.nk; i += nt) A.Tcw[i] = A.Tnew[i];
}

__global__ void k_np_commit_pts(NpArgs A) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < 3 * A.nl && A.st[3] != 0.0) A.pts[i] = A.pnew[i];
}

__global__ void k_np_inv_poses(int nk, const double *in, double *out) {  // T_wc <-> T_cw
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nk) return;
    const double *T = in + 16 * (size_t)k;
    double *o = out + 16 * (size_t)k;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) o[4 * i + j] = T[4 * j + i];
    for (int i = 0; i < 3; i++) o[4 * i + 3] = -(T[i] * T[3] + T[4 + i] * T[7] + T[8 + i] * T[11]);
    o[12] = o[13] = o[14] = 0.0;
    o[15] = 1.0;
}

// Fixed-order sum of r2 into st[slot] (slot 5: chi2 of the initial state; 6: chi2 of the returned state).
__global__ __launch_bounds__(256) void k_np_chi(NpArgs A, int slot) {
    __shared__ double sred[256];
    const int tid = threadIdx.x, nt = blockDim.x;
    double s = 0.0;
    for (int k = tid; k < A.N; k += nt) s += A.r2[k];
    sred[tid] = s;
    __syncthreads();
    for (int d = nt / 2; d > 0; d >>= 1) {
        if (tid < d) sred[tid] += sred[tid + d];
        __syncthreads();
    }
    if (tid == 0) A.st[slot] = sred[0];
}

__global__ void k_np_r2_only(NpArgs A) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= A.N) return;
    double P[3], r[2];
    np_residual(A, A.Tcw + 16 * (size_t)A.okf[k], A.pts + 3 * (size_t)A.olm[k], A.ouv + 2 * (size_t)k, P, r);
    A.r2[k] = r[0] * r[0] + r[1] * r[1];
}

// stats (device, 4): chi2_0, chi2_n, rms_0, rms_n (ba.py:720-722, 865-869).
__global__ void k_np_stats(const double *st, int N, double *stats) {
    if (threadIdx.x) return;
    stats[0] = st[5];
    stats[1] = st[6];
    stats[2] = N ? sqrt(st[5] / N) : NAN;
    stats[3] = N ? sqrt(st[6] / N) : NAN;
}


// ---- device-side planning of the Euclidean solve (the allocation-free entry point): observations grouped by
// landmark in index order, as the host plan of xba_np_solve groups them, and the index checks.  err[0] != 0 after
// k_np_plan_count when an index is out of range; the later kernels then write nothing.
__global__ void k_np_fill_i32(int32_t *a, int n, int32_t v) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) a[i] = v;
}
__global__ void k_np_fill_f64(double *a, int64_t n, double v) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        a[i] = v;
}

// counts per landmark into lstart[l + 1] (zeroed), the range checks of okf / olm and of the rotation factors' slots
__global__ void k_np_plan_count(int N, int nk, int nl, const int32_t *okf, const int32_t *olm, int nrot2,
                                const int32_t *rot_ab, int32_t *lstart, int32_t *err) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < N) {
        const int f = okf[k], l = olm[k];
        if (f < 0 || f >= nk || l < 0 || l >= nl)
            atomicOr(err, 1);
        else
            atomicAdd(lstart + l + 1, 1);
    }
    if (k < nrot2 && (rot_ab[k] < 0 || rot_ab[k] >= nk)) atomicOr(err, 1);
}

// one workgroup: lstart <- exclusive offsets (lstart[0] = 0), fill <- a copy of the first nl offsets
__global__ __launch_bounds__(1024) void k_np_plan_scan(int nl, int32_t *lstart, int32_t *fill, const int32_t *err) {
    __shared__ int32_t part[1024];
    __shared__ int32_t carry;
    if (*err) return;
    const int tid = threadIdx.x, nt = blockDim.x;
    if (tid == 0) carry = 0;
    __syncthreads();
    for (int base = 1; base <= nl; base += nt) {
        const int i = base + tid;
        int32_t v = i <= nl ? lstart[i] : 0;
        part[tid] = v;
        __syncthreads();
        for (int d = 1; d < nt; d <<= 1) {  // inclusive Hillis-Steele scan of the block
            const int32_t add = tid >= d ? part[tid - d] : 0;
            __syncthreads();
            part[tid] += add;
            __syncthreads();
        }
        const int32_t c0 = carry;
        if (i <= nl) lstart[i] = c0 + part[tid];
        __syncthreads();
        if (tid == nt - 1) carry = c0 + part[tid];
        __syncthreads();
    }
    for (int l = tid; l < nl; l += nt) fill[l] = lstart[l];
}

// observation k to its landmark's segment (arbitrary order inside the segment; k_np_plan_sort orders it)
__global__ void k_np_plan_scatter(int N, int nl, const int32_t *olm, int32_t *fill, int32_t *order,
                                  const int32_t *err) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= N || *err) return;
    const int l = olm[k];
    if (l < 0 || l >= nl) return;
    order[atomicAdd(fill + l, 1)] = k;
}

// one thread per landmark: its segment in ascending observation index (insertion sort; segments are short)
__global__ void k_np_plan_sort(int nl, const int32_t *lstart, int32_t *order, const int32_t *err) {
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= nl || *err) return;
    const int a = lstart[l], b = lstart[l + 1];
    for (int i = a + 1; i < b; i++) {
        const int32_t v = order[i];
        int j = i - 1;
        while (j >= a && order[j] > v) {
            order[j + 1] = order[j];
            j--;
        }
        order[j + 1] = v;
    }
}

// ============================================================= the inverse-depth local BA (ba.py:1212-1342) ==
// Host frame (R_wc, t_wc) of landmark l: the window pose Twc[hidx] or the constant hext row (ba.py:1116-1135).
__device__ inline void id_host(const NpArgs &A, const double *Twc, int l, double *R, double *t) {
    const int h = A.hidx[l];
    const double *T = h >= 0 ? Twc + 16 * (size_t)h : nullptr, *E = A.hext + 12 * (size_t)l;
    for (int i = 0; i < 3; i++) {
        for (int j = 0; j < 3; j++) R[3 * i + j] = T ? T[4 * i + j] : E[4 * i + j];
        t[i] = T ? T[4 * i + 3] : E[4 * i + 3];
    }
}

// ba.py:1226-1234 geom for observation k: host point Ph, camera point Pc, residual r.
__device__ inline void id_residual(const NpArgs &A, int k, const double *Twc, const double *Tcw, const double *rho,
                                   const double *mb, double *R, double *Ph, double *Pc, double *r) {
    const int l = A.olm[k];
    double t[3], Pw[3];
    id_host(A, Twc, l, R, t);
    const double ir = fmax(rho[l], 1e-12);
    for (int i = 0; i < 3; i++) Ph[i] = mb[3 * (size_t)l + i] / ir;
    for (int i = 0; i < 3; i++) Pw[i] = (R[3 * i] * Ph[0] + R[3 * i + 1] * Ph[1] + R[3 * i + 2] * Ph[2]) + t[i];
    const double *T = Tcw + 16 * (size_t)A.okf[k];
    for (int i = 0; i < 3; i++) Pc[i] = (T[4 * i] * Pw[0] + T[4 * i + 1] * Pw[1] + T[4 * i + 2] * Pw[2]) + T[4 * i + 3];
    const double z = fmax(Pc[2], 1e-6);
    r[0] = A.fx * Pc[0] / z + A.cx - A.ouv[2 * (size_t)k];
    r[1] = A.fx * Pc[1] / z + A.cy - A.ouv[2 * (size_t)k + 1];
}

// Residual, Huber weight and the sqrt(w)-scaled blocks of _analytic_blocks_id (ba.py:1138-1209): observer
// [I | -hat(Pc)], host [-I | hat(Ph)] in the host frame (0 for hosts outside the window), landmark (x_n, y_n, rho).
__global__ void k_id_obs(NpArgs A) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= A.N) return;
    double R[9], Ph[3], Pc[3], r[2];
    id_residual(A, k, A.Twc, A.Tcw, A.rho, A.mb, R, Ph, Pc, r);
    A.r2[k] = r[0] * r[0] + r[1] * r[1];
    const double e = sqrt(r[0] * r[0] + r[1] * r[1]);
    const double w = (e <= A.K) ? 1.0 : A.K / fmax(e, 1e-9), sw = sqrt(w);
    const double x = Pc[0], y = Pc[1], zz = fmax(Pc[2], 1e-6), iz = 1.0 / zz, fx = A.fx;
    const double Am[2][3] = {{fx * iz, 0.0, -fx * x * iz * iz}, {0.0, fx * iz, -fx * y * iz * iz}};
    const double dPo[3][6] = {{1, 0, 0, 0, zz, -y}, {0, 1, 0, -zz, 0, x}, {0, 0, 1, y, -x, 0}};
    const double dPh[3][6] = {{-1, 0, 0, 0, -Ph[2], Ph[1]}, {0, -1, 0, Ph[2], 0, -Ph[0]}, {0, 0, -1, -Ph[1], Ph[0], 0}};
    const double *T = A.Tcw + 16 * (size_t)A.okf[k];
    double Rcw[9], Rch[9], ARch[2][3];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) Rcw[3 * i + j] = T[4 * i + j];
    mm3(Rcw, R, Rch);
    for (int a = 0; a < 2; a++)
        for (int c = 0; c < 3; c++) ARch[a][c] = Am[a][0] * Rch[c] + Am[a][1] * Rch[3 + c] + Am[a][2] * Rch[6 + c];
    const bool host_in = A.hidx[A.olm[k]] >= 0;
    const double irho = 1.0 / fmax(A.rho[A.olm[k]], 1e-12);
    double *J = A.J + 32 * (size_t)k;
    for (int a = 0; a < 2; a++) {
        for (int c = 0; c < 6; c++) {
            J[6 * a + c] = (Am[a][0] * dPo[0][c] + Am[a][1] * dPo[1][c] + Am[a][2] * dPo[2][c]) * sw;
            J[12 + 6 * a + c] =
                host_in ? (ARch[a][0] * dPh[0][c] + ARch[a][1] * dPh[1][c] + ARch[a][2] * dPh[2][c]) * sw : 0.0;
        }
        J[24 + 3 * a] = ARch[a][0] * irho * sw;
        J[24 + 3 * a + 1] = ARch[a][1] * irho * sw;
        J[24 + 3 * a + 2] = -(ARch[a][0] * Ph[0] + ARch[a][1] * Ph[1] + ARch[a][2] * Ph[2]) * irho * sw;
        J[30 + a] = r[a] * sw;
    }
}

// Per landmark: Hll, bl, the bearing prior (unrobust) and inverse-range prior (Huber), damping by
// lam * max(tr/3, 1e-9), inverse (ba.py:1244-1264, 1298-1301).
__global__ void k_id_lm(NpArgs A) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= A.nl || np_skip(A)) return;
    const double lam = A.st[0];
    double Hll[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, bl[3] = {0, 0, 0};
    for (int q = A.lstart[j]; q < A.lstart[j + 1]; q++) {
        const double *J = A.J + 32 * (size_t)A.order[q], *Jl = J + 24, *rw = J + 30;
        for (int i = 0; i < 3; i++) {
            for (int c = 0; c < 3; c++) Hll[3 * i + c] += Jl[i] * Jl[c] + Jl[3 + i] * Jl[3 + c];
            bl[i] += -(Jl[i] * rw[0] + Jl[3 + i] * rw[1]);
        }
    }
    if (A.has_prior && A.pr_valid[j] > 0.5) {
        const double *mb = A.mb + 3 * (size_t)j, *m0 = A.pr_m0 + 3 * (size_t)j;
        const double rp[3] = {(mb[0] - m0[0]) * A.pr_wbear[j], (mb[1] - m0[1]) * A.pr_wbear[j],
                              (A.rho[j] - A.pr_rho0[j]) * A.pr_invsig[j]};
        const double J0[3] = {A.pr_wbear[j], A.pr_wbear[j], A.pr_invsig[j]};
        for (int c = 0; c < 3; c++) {
            const double wp = fabs(rp[c]) <= A.K ? 1.0 : A.K / fmax(fabs(rp[c]), 1e-9);
            const double swp = c < 2 ? 1.0 : sqrt(wp), Jpr = J0[c] * swp;
            Hll[4 * c] += Jpr * Jpr;
            bl[c] -= Jpr * rp[c] * swp;
        }
    }
    const double tr = Hll[0] + Hll[4] + Hll[8], d = lam * fmax(tr / 3.0, 1e-9);
    for (int i = 0; i < 3; i++) Hll[4 * i] += d;
    const double c00 = Hll[4] * Hll[8] - Hll[5] * Hll[7], c01 = Hll[5] * Hll[6] - Hll[3] * Hll[8],
                 c02 = Hll[3] * Hll[7] - Hll[4] * Hll[6];
    const double det = Hll[0] * c00 + Hll[1] * c01 + Hll[2] * c02, id = 1.0 / det;
    double *Hi = A.Hlli + 9 * (size_t)j;
    Hi[0] = c00 * id; Hi[1] = (Hll[2] * Hll[7] - Hll[1] * Hll[8]) * id; Hi[2] = (Hll[1] * Hll[5] - Hll[2] * Hll[4]) * id;
    Hi[3] = c01 * id; Hi[4] = (Hll[0] * Hll[8] - Hll[2] * Hll[6]) * id; Hi[5] = (Hll[2] * Hll[3] - Hll[0] * Hll[5]) * id;
    Hi[6] = c02 * id; Hi[7] = (Hll[1] * Hll[6] - Hll[0] * Hll[7]) * id; Hi[8] = (Hll[0] * Hll[4] - Hll[1] * Hll[3]) * id;
    for (int i = 0; i < 3; i++) A.bl[3 * (size_t)j + i] = bl[i];
}

__device__ inline void id_hpl(const double *J, int sa, double *Hpl) {  // (6x3) Jpose_sa^T Jlm
    const double *Jp = J + 12 * sa, *Jl = J + 24;
    for (int i = 0; i < 6; i++)
        for (int c = 0; c < 3; c++) Hpl[3 * i + c] = Jp[i] * Jl[c] + Jp[6 + i] * Jl[3 + c];
}

// ---- the reduced system of both local-BA solves, one thread per (observation block, observation block) pair of a
// landmark (ba.py:759-823 Euclidean: one pose block per observation; :1266-1305 inverse depth: the observer's and
// the in-window host's).  A thread's pair (a, sa) x (b, sb) adds -(Hpl_a Hlli) Hpl_b^T (plus Jp_a^T Jp_b when a == b,
// and the gradient of (a, sa) when also sa == sb) to the upper triangle of H only: pairs whose row block lies
// below their column block are the transposes of pairs that are kept, and the flush mirrors the triangle.  When the
// triangle and the gradient fit (INLDS), a workgroup accumulates into its LDS copy (LDS atomics) and adds it to H
// and bp once; otherwise straight into H and bp (global atomics).  Grid: workgroups of 256 threads over blocks of
// NP_LB landmarks; the pairs of a block are numbered by a prefix sum of (blocks per landmark)^2.
#define NP_LB 64
template <bool ID>
__device__ inline int np_block_slot(const NpArgs &A, int k, int sb, int j) {
    return ID ? (sb == 0 ? A.okf[k] : A.hidx[j]) : A.okf[k];
}

__device__ inline int np_tri(int r, int c, int n) { return r * n - r * (r - 1) / 2 + (c - r); }  // r <= c

template <bool ID, bool INLDS>
__global__ __launch_bounds__(256) void k_np_schur_pairs(NpArgs A, int n) {
    extern __shared__ double tri[];  // INLDS: n (n + 1) / 2 upper-triangle entries, then the n gradient entries
    __shared__ int spfx[NP_LB + 1];
    if (np_skip(A)) return;  // uniform
    constexpr int NB = ID ? 2 : 1, JS = ID ? 32 : 20;
    const int tid = threadIdx.x, ntri = n * (n + 1) / 2;
    double *Hacc = INLDS ? tri : A.H, *bacc = INLDS ? tri + ntri : A.bp;
    if (INLDS) {
        for (int e = tid; e < ntri + n; e += 256) tri[e] = 0.0;
        __syncthreads();
    }
    auto add = [&](double *p, double v) {
        if (INLDS)
            __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        else
            atomicAdd(p, v);
    };
    for (int l0 = blockIdx.x * NP_LB; l0 < A.nl; l0 += gridDim.x * NP_LB) {
        const int l = l0 + tid;
        const int m = (tid < NP_LB && l < A.nl) ? NB * (A.lstart[l + 1] - A.lstart[l]) : 0;  // its blocks
        if (tid < NP_LB) spfx[tid + 1] = m * m;
        if (tid == 0) spfx[0] = 0;
        __syncthreads();
        for (int d = 1; d < NP_LB; d <<= 1) {  // inclusive scan of spfx[1..NP_LB]
            const int v = (tid < NP_LB && tid + 1 > d) ? spfx[tid + 1 - d] : 0;
            __syncthreads();
            if (tid < NP_LB) spfx[tid + 1] += v;
            __syncthreads();
        }
        const int tot = spfx[NP_LB];
        for (int p = tid; p < tot; p += 256) {
            int lo = 0, hi = NP_LB - 1;  // the last landmark whose first pair is at or before p
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (spfx[mid] <= p) lo = mid; else hi = mid - 1;
            }
            const int j = l0 + lo, q0 = A.lstart[j], mj = NB * (A.lstart[j + 1] - q0), q = p - spfx[lo];
            const int ea = q / mj, eb = q % mj;  // observation blocks a and b of landmark j
            const int ka = A.order[q0 + ea / NB], kb = A.order[q0 + eb / NB], sa = ea % NB, sb = eb % NB;
            const int s1 = np_block_slot<ID>(A, ka, sa, j), s2 = np_block_slot<ID>(A, kb, sb, j);
            if (s1 < 0 || s2 < 0 || s1 > s2) continue;  // an out-of-window host, or a transposed pair
            const double *Ja = A.J + JS * (size_t)ka, *Jb = A.J + JS * (size_t)kb;
            const double *Pa = Ja + 12 * sa, *Pb = Jb + 12 * sb;  // the 2 x 6 pose blocks
            const double *Hi = A.Hlli + 9 * (size_t)j;
            double Hpa[18], Hpb[18], T3[18];
            if (ID) {
                id_hpl(Ja, sa, Hpa);
                id_hpl(Jb, sb, Hpb);
            } else {
                hpl_of(Ja, Hpa);
                hpl_of(Jb, Hpb);
            }
            for (int i = 0; i < 6; i++)
                for (int c = 0; c < 3; c++)
                    T3[3 * i + c] = Hpa[3 * i] * Hi[c] + Hpa[3 * i + 1] * Hi[3 + c] + Hpa[3 * i + 2] * Hi[6 + c];
            const bool same_obs = ka == kb, diag = s1 == s2;
            const int r0 = 6 * s1, c0 = 6 * s2;
            for (int i = 0; i < 6; i++)
                for (int c = diag ? i : 0; c < 6; c++) {
                    double v = -(T3[3 * i] * Hpb[3 * c] + T3[3 * i + 1] * Hpb[3 * c + 1] + T3[3 * i + 2] * Hpb[3 * c + 2]);
                    if (same_obs) v += Pa[i] * Pb[c] + Pa[6 + i] * Pb[6 + c];
                    add(INLDS ? Hacc + np_tri(r0 + i, c0 + c, n) : Hacc + (size_t)(r0 + i) * n + c0 + c, v);
                    if (!INLDS && (r0 + i) != (c0 + c)) add(Hacc + (size_t)(c0 + c) * n + r0 + i, v);
                }
            if (same_obs && sa == sb) {
                const double *bl = A.bl + 3 * (size_t)j, *rw = Ja + (ID ? 30 : 18);
                for (int i = 0; i < 6; i++)
                    add(bacc + r0 + i, -(Pa[i] * rw[0] + Pa[6 + i] * rw[1]) -
                                           (T3[3 * i] * bl[0] + T3[3 * i + 1] * bl[1] + T3[3 * i + 2] * bl[2]));
            }
        }
        __syncthreads();
    }
    if (INLDS) {  // the workgroup's triangle and gradient into H (both triangles) and bp
        __syncthreads();
        for (int e = tid; e < n * n; e += 256) {
            const int r = e / n, c = e % n;
            const double v = r <= c ? tri[np_tri(r, c, n)] : tri[np_tri(c, r, n)];
            if (v != 0.0) atomicAdd(A.H + e, v);
        }
        for (int i = tid; i < n; i += 256)
            if (tri[ntri + i] != 0.0) atomicAdd(A.bp + i, tri[ntri + i]);
    }
}

// dxl = Hlli (bl - sum Hpl^T dxp), trial bearing (x_n, y_n += dxl) and inverse depth max(rho + dxl, floor)
// (ba.py:1318-1330).
__global__ void k_id_backsub(NpArgs A) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= A.nl || A.st[4] == 0.0 || np_skip(A)) return;
    double rhs[3] = {A.bl[3 * (size_t)j], A.bl[3 * (size_t)j + 1], A.bl[3 * (size_t)j + 2]};
    const int hs = A.hidx[j];
    for (int q = A.lstart[j]; q < A.lstart[j + 1]; q++) {
        const int k = A.order[q];
        const int sl[2] = {A.okf[k], hs};
        for (int sa = 0; sa < 2; sa++) {
            if (sl[sa] < 0) continue;
            double Hp[18];
            id_hpl(A.J + 32 * (size_t)k, sa, Hp);
            const double *d = A.dx + 6 * sl[sa];
            for (int c = 0; c < 3; c++) {
                double s = 0.0;
                for (int i = 0; i < 6; i++) s += Hp[3 * i + c] * d[i];
                rhs[c] += -s;
            }
        }
    }
    const double *Hi = A.Hlli + 9 * (size_t)j;
    double dl[3];
    for (int i = 0; i < 3; i++) dl[i] = Hi[3 * i] * rhs[0] + Hi[3 * i + 1] * rhs[1] + Hi[3 * i + 2] * rhs[2];
    if (!(isfinite(dl[0]) && isfinite(dl[1]) && isfinite(dl[2]))) A.st[ST_NONFIN] = 1.0;
    A.mbn[3 * (size_t)j] = A.mb[3 * (size_t)j] + dl[0];
    A.mbn[3 * (size_t)j + 1] = A.mb[3 * (size_t)j + 1] + dl[1];
    A.mbn[3 * (size_t)j + 2] = A.mb[3 * (size_t)j + 2];
    A.rhon[j] = fmax(A.rho[j] + dl[2], A.rho_floor);
}

// r2 at the trial (trial = 1) or the current state.
__global__ void k_id_r2(NpArgs A, int trial) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= A.N || (trial && A.st[4] == 0.0)) return;
    double R[9], Ph[3], Pc[3], r[2];
    if (trial)
        id_residual(A, k, A.Twn, A.Tnew, A.rhon, A.mbn, R, Ph, Pc, r);
    else
        id_residual(A, k, A.Twc, A.Tcw, A.rho, A.mb, R, Ph, Pc, r);
    A.r2[k] = r[0] * r[0] + r[1] * r[1];
}

// On an accepted step: the host frames, inverse depths and bearings follow the poses k_np_accept committed.
__global__ void k_id_commit(NpArgs A) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (A.st[3] == 0.0) return;
    if (i < 16 * A.nk) A.Twc[i] = A.Twn[i];
    if (i < A.nl) A.rho[i] = A.rhon[i];
    if (i < 3 * A.nl) A.mb[i] = A.mbn[i];
}


// ================================================================== the GPU drivers' LM policy (policy 1) ==
// Fixed-order sum of the Huber-weighted squared residuals (rw of the J rows) into st[slot].
__device__ inline double np_wchi_sum(const NpArgs &A, double *sred) {
    const int tid = threadIdx.x, nt = blockDim.x;
    double s = 0.0;
    for (int k = tid; k < A.N; k += nt) {
        const double *rw = A.J + (size_t)A.jstride * k + A.rwoff;
        s += rw[0] * rw[0] + rw[1] * rw[1];
    }
    sred[tid] = s;
    __syncthreads();
    for (int d = nt / 2; d > 0; d >>= 1) {
        if (tid < d) sred[tid] += sred[tid + d];
        __syncthreads();
    }
    return sred[0];
}

// ba.py:1015-1026 / 1436-1445, after the linearisation: chi = weighted chi2 at the current state; if it grew past the
// last accepted one, restore that state, lam = min(4 lam, 1e2) and skip the rest of the iteration; else remember the
// state as the accepted one.
__global__ __launch_bounds__(256) void k_np_gate(NpArgs A, int id) {
    __shared__ double sred[256];
    __shared__ int srej;
    const int tid = threadIdx.x, nt = blockDim.x;
    const double chi = np_wchi_sum(A, sred);
    if (tid == 0) {
        if (A.st[ST_RMS0] < 0.0) A.st[ST_RMS0] = sqrt(chi / (A.N > 1 ? A.N : 1));
        srej = A.st[ST_CHIPREV] >= 0.0 && chi > A.st[ST_CHIPREV];
        A.st[ST_SKIP] = srej ? 1.0 : 0.0;
        A.st[ST_NONFIN] = 0.0;
        A.st[3] = 0.0;
        if (srej) {
            A.st[0] = fmin(A.st[0] * 4.0, 1e2);
        } else {
            A.st[ST_CHIPREV] = chi;
            A.st[ST_BESTCHI] = chi;
        }
    }
    __syncthreads();
    double *dT = srej ? A.Tcw : A.Tbest;
    const double *sT = srej ? A.Tbest : A.Tcw;
    for (int i = tid; i < 16 * A.nk; i += nt) dT[i] = sT[i];
    if (id) {
        double *dW = srej ? A.Twc : A.Wbest, *dr = srej ? A.rho : A.rbest, *dm = srej ? A.mb : A.mbest;
        const double *sW = srej ? A.Wbest : A.Twc, *sr = srej ? A.rbest : A.rho, *sm = srej ? A.mbest : A.mb;
        for (int i = tid; i < 16 * A.nk; i += nt) dW[i] = sW[i];
        for (int i = tid; i < A.nl; i += nt) dr[i] = sr[i];
        for (int i = tid; i < 3 * A.nl; i += nt) dm[i] = sm[i];
    } else {
        double *dp = srej ? A.pts : A.pbest;
        const double *sp = srej ? A.pbest : A.pts;
        for (int i = tid; i < 3 * A.nl; i += nt) dp[i] = sp[i];
    }
}

// ba.py:1058-1073 / 1468-1478: a failed or non-finite step sets lam = min(10 lam, 1e3); otherwise the step is
// applied (poses exp(dx) T for free keyframes, landmarks + dl, inverse depth floored) and lam = max(lam / 2, 1e-7).
__global__ __launch_bounds__(256) void k_np_apply(NpArgs A, int n, int id) {
    __shared__ int sfin;
    const int tid = threadIdx.x, nt = blockDim.x;
    if (np_skip(A)) return;
    if (tid == 0) sfin = A.st[4] != 0.0 && A.st[ST_NONFIN] == 0.0;
    __syncthreads();
    if (sfin) {
        int bad = 0;
        for (int i = tid; i < n; i += nt) bad |= !isfinite(A.dx[i]);
        if (bad) sfin = 0;  // benign race: every writer stores 0
    }
    __syncthreads();
    if (!sfin) {
        if (tid == 0) A.st[0] = fmin(A.st[0] * 10.0, 1e3);
        return;
    }
    for (int i = tid; i < 16 * A.nk; i += nt) A.Tcw[i] = A.Tnew[i];
    if (id) {
        for (int i = tid; i < 16 * A.nk; i += nt) A.Twc[i] = A.Twn[i];
        for (int i = tid; i < A.nl; i += nt) A.rho[i] = A.rhon[i];
        for (int i = tid; i < 3 * A.nl; i += nt) A.mb[i] = A.mbn[i];
    } else {
        for (int i = tid; i < 3 * A.nl; i += nt) A.pts[i] = A.pnew[i];
    }
    if (tid == 0) {
        A.st[0] = fmax(A.st[0] * 0.5, 1e-7);
        A.st[3] = 1.0;
    }
}

// The one SO(3) projection of the GPU drivers (ba.py:1076, 1480): the orthogonal polar factor of every rotation.
__global__ void k_np_project(int nk, double *Tcw) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nk) return;
    double *T = Tcw + 16 * (size_t)k, X[9];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) X[3 * i + j] = T[4 * i + j];
    for (int itp = 0; itp < 30; itp++) {  // X <- (X + X^-T) / 2
        double c00 = X[4] * X[8] - X[5] * X[7], c01 = X[5] * X[6] - X[3] * X[8], c02 = X[3] * X[7] - X[4] * X[6];
        double det = X[0] * c00 + X[1] * c01 + X[2] * c02, id = 1.0 / det;
        double Xi[9] = {c00 * id, (X[2] * X[7] - X[1] * X[8]) * id, (X[1] * X[5] - X[2] * X[4]) * id,
                        c01 * id, (X[0] * X[8] - X[2] * X[6]) * id, (X[2] * X[3] - X[0] * X[5]) * id,
                        c02 * id, (X[1] * X[6] - X[0] * X[7]) * id, (X[0] * X[4] - X[1] * X[3]) * id};
        double dmax = 0.0;
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) {
                double nv = 0.5 * (X[3 * i + j] + Xi[3 * j + i]);
                dmax = fmax(dmax, fabs(nv - X[3 * i + j]));
                X[3 * i + j] = nv;
            }
        if (dmax < 1e-16) break;
    }
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) T[4 * i + j] = X[3 * i + j];
}

// stats of the GPU drivers (ba.py:1086-1089, 1494-1497): chi2_0 is the weighted chi2 of the last accepted state (the
// final one when there is none), rms_0 that of the first iteration, chi2_n / rms_n the weighted chi2 of the result.
__global__ __launch_bounds__(256) void k_np_wstats(NpArgs A, double *stats) {
    __shared__ double sred[256];
    const double chi = np_wchi_sum(A, sred);
    if (threadIdx.x) return;
    stats[0] = A.st[ST_BESTCHI] > 0.0 ? A.st[ST_BESTCHI] : chi;
    stats[1] = chi;
    stats[2] = A.st[ST_RMS0] >= 0.0 ? A.st[ST_RMS0] : NAN;
    stats[3] = sqrt(chi / (A.N > 1 ? A.N : 1));
}

}  // namespace xba
