// The implicit-function adjoint of the window solve's Gauss-Newton step.
//
// At a fixed state the step is dx = -F^-1 b with F = H + 1e-9 I, H and b the undamped (lambda = 0) reduced system
// of vio/window.py:291-463 + :865-886 and the damping rule of vio/window.py:933-941 at lambda = 0.  Every visual
// observation row q adds om_q w_q J_q^T J_q to the full (pose + landmark) system and om_q w_q J_q^T r_q to its
// gradient, with r_q = isig_q (pred_q - z_q) the whitened residual, w_q the Huber weight of |r_q| and om_q the
// row's weight multiplier.  For L(dx) with g = dL/d(dx), y = F^-1 g (the same factor), extended to the landmarks
// by y_l = -Hll^-1 Hlp y_p, and dx extended by dx_l = -Hll^-1 (bl + Hlp dx_p), the implicit-function theorem gives
// per row, with u = J_q dx, v = J_q y and s = v . (u + r_q):
//     dL/dz_q  = om_q isig_q (dw/dr s + w_q v)          dL/dom_q = -w_q s
// (J does not depend on z; a landmark that failed the det guard is not eliminated, so its dx_l = y_l = 0.)
#include "xba_internal.h"

namespace xba {

// Per window: F = H + 1e-9 I factored in FP64 into fact64 (kept for the adjoint), dx = -F^-1 b.
__global__ __launch_bounds__(256) void k_gn_solve(DArgs D) {
    __shared__ int sflag;
    __shared__ double sinv[CH_SCRATCH];
    const int w = blockIdx.x, tid = threadIdx.x, nt = blockDim.x;
    const WinInfo W = D.wins[w];
    const int n = W.n;
    double *F = D.fact64 + W.h0, *x = D.dx + W.st0;
    const double *H = D.H + W.h0, *bv = D.b + W.st0;
    for (int e = tid; e < n * n; e += nt) F[e] = H[e] + ((e / n == e % n) ? 1e-9 : 0.0);
    __syncthreads();
    const bool ok = chol_wg<double>(F, n, &sflag, sinv, n);
    if (ok) {
        for (int i = tid; i < n; i += nt) x[i] = bv[i];
        __syncthreads();
        chol_solve_wg<double>(F, n, x, n);
        for (int i = tid; i < n; i += nt) x[i] = -x[i];
    } else {
        for (int i = tid; i < n; i += nt) x[i] = NAN;
    }
    if (tid == 0) D.wstate[(size_t)NWS * w + WS_CHOLOK] = ok ? 1.0 : 0.0;
}

// Per window: y = F^-1 g with the kept factor (0 where the factorisation failed).
__global__ __launch_bounds__(256) void k_adj_y(DArgs D, const double *g, double *y) {
    const int w = blockIdx.x, tid = threadIdx.x, nt = blockDim.x;
    const WinInfo W = D.wins[w];
    const int n = W.n;
    double *yw = y + W.st0;
    const bool ok = D.wstate[(size_t)NWS * w + WS_CHOLOK] != 0.0;
    for (int i = tid; i < n; i += nt) yw[i] = ok ? g[W.st0 + i] : 0.0;
    __syncthreads();
    if (ok) chol_solve_wg<double>(D.fact64 + W.h0, n, yw, n);
}

// Camera pose of one slot in the layouts obs_eval reads: c64 = [Rwc 9, pwc 3], c = [Rwc 9, pwc 3, Rwb 9].
__device__ inline void adj_slot(const double *Rwb, const double *pwb, const double *wp, int gs, double *c64,
                                double *c) {
    const double *Rbc = wp + 8, *pbc = wp + 17, *Rb = Rwb + 9 * (size_t)gs, *pb = pwb + 3 * (size_t)gs;
    for (int i = 0; i < 3; i++) {
        for (int j = 0; j < 3; j++) {
            double acc = 0.0;
            for (int k = 0; k < 3; k++) acc += Rb[3 * i + k] * Rbc[3 * k + j];
            c64[3 * i + j] = c[3 * i + j] = acc;
        }
        double acc = pb[i];
        for (int k = 0; k < 3; k++) acc += Rb[3 * i + k] * pbc[k];
        c64[9 + i] = c[9 + i] = acc;
    }
    for (int i = 0; i < 9; i++) c[12 + i] = Rb[i];
}

// One observation row at the current state: whitened residual, Jacobians, Huber weight (0: row inactive).
struct AdjRow {
    int nd, t;
    double e[2], Jl[2][3], Jt[2][6], Jh[2][6], wh, om, isg, en;
};

__device__ inline bool adj_row(const KArgs<double> &A, int k, int h, int slot0, const double *wp, const double *sh64,
                               const double *sh, const double *wc, const double *wd, const double *lm, AdjRow &R) {
    const int info = A.ob_info[k];
    R.t = info & 0xfff;
    const int cam = (info >> 12) & 1, valid = (info >> 13) & 1;
    R.nd = cam == 0 ? 2 : 1;
    if (!valid) return false;
    double st64[12], st[24];
    adj_slot(A.Rwb, A.pwb, wp, slot0 + R.t, st64, st);
    if (!obs_eval<double>(sh64, st64, sh, st, wc, wd, lm[0], lm[1], lm[2], cam, A.z[2 * (size_t)k],
                          A.z[2 * (size_t)k + 1], R.t == h, true, R.e, R.Jl, R.Jt, R.Jh))
        return false;
    R.isg = A.isig[k];
    double en2 = 0.0;
    for (int i = 0; i < R.nd; i++) {
        R.e[i] *= R.isg;
        en2 += R.e[i] * R.e[i];
        for (int j = 0; j < 3; j++) R.Jl[i][j] *= R.isg;
        for (int j = 0; j < 6; j++) {
            R.Jt[i][j] *= R.isg;
            R.Jh[i][j] *= R.isg;
        }
    }
    if (R.nd == 1) {
        R.e[1] = 0.0;
        for (int j = 0; j < 3; j++) R.Jl[1][j] = 0.0;
        for (int j = 0; j < 6; j++) R.Jt[1][j] = R.Jh[1][j] = 0.0;
    }
    R.en = sqrt(en2);
    double c;
    huber(R.en, wd[5], &R.wh, &c);
    R.om = A.ob_w ? A.ob_w[k] : 1.0;
    return true;
}

// One thread per landmark: dx_l, y_l from its rows, then dL/dz and dL/dom of every row.
__global__ void k_adj_obs(KArgs<double> A, const int32_t *lm_win, const double *y, double *gz, double *gw) {
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= A.nlm) return;
    const int o0 = A.ob_start[l], o1 = A.ob_start[l + 1];
    const int li = A.lm_info[l];
    if (!((li >> 12) & 1)) {
        for (int k = o0; k < o1; k++) {
            gz[2 * (size_t)k] = gz[2 * (size_t)k + 1] = 0.0;
            gw[k] = 0.0;
        }
        return;
    }
    const int w = lm_win[l], h = li & 0xfff;
    const WinInfo W = A.wins[w];
    const double *wp = A.win_par + (size_t)XBA_WPAR * w;
    double sh64[12], sh[24], wc[24], wd[6];
    adj_slot(A.Rwb, A.pwb, wp, W.slot0 + h, sh64, sh);
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) wc[3 * i + j] = wp[8 + 3 * j + i];  // Rcb = Rbc^T
    const double *pbc = wp + 17;
    const double Px[9] = {0, -pbc[2], pbc[1], pbc[2], 0, -pbc[0], -pbc[1], pbc[0], 0};
    for (int i = 0; i < 9; i++) wc[9 + i] = Px[i];
    for (int i = 0; i < 6; i++) wc[18 + i] = wd[i] = wp[i];
    const double lm[3] = {A.lmp[3 * (size_t)l], A.lmp[3 * (size_t)l + 1], A.lmp[3 * (size_t)l + 2]};
    const double *dxw = A.dx + W.st0, *yw = y + W.st0;
    const int oh = A.slot_off[W.slot0 + h];
    // pass 1: Hlp dx_p and Hlp y_p
    double sdx[3] = {0, 0, 0}, sy[3] = {0, 0, 0};
    const bool ok = A.lmok[l] != 0;
    if (ok)
        for (int k = o0; k < o1; k++) {
            AdjRow R;
            if (!adj_row(A, k, h, W.slot0, wp, sh64, sh, wc, wd, lm, R) || R.t == h) continue;
            const int ot = A.slot_off[W.slot0 + R.t];
            const double ww = R.wh * R.om;
            for (int r = 0; r < R.nd; r++) {
                double jd = 0.0, jy = 0.0;
                for (int j = 0; j < 6; j++) {
                    jd += R.Jt[r][j] * dxw[ot + j] + R.Jh[r][j] * dxw[oh + j];
                    jy += R.Jt[r][j] * yw[ot + j] + R.Jh[r][j] * yw[oh + j];
                }
                for (int i = 0; i < 3; i++) {
                    sdx[i] += ww * R.Jl[r][i] * jd;
                    sy[i] += ww * R.Jl[r][i] * jy;
                }
            }
        }
    double dl[3] = {0, 0, 0}, yl[3] = {0, 0, 0};
    if (ok) {
        const double *Hs = A.lmHi + 6 * (size_t)l, *bl = A.lmbl + 3 * (size_t)l;
        const double Hi[9] = {Hs[0], Hs[1], Hs[2], Hs[1], Hs[3], Hs[4], Hs[2], Hs[4], Hs[5]};
        for (int i = 0; i < 3; i++) {
            dl[i] = -(Hi[3 * i] * (bl[0] + sdx[0]) + Hi[3 * i + 1] * (bl[1] + sdx[1]) + Hi[3 * i + 2] * (bl[2] + sdx[2]));
            yl[i] = -(Hi[3 * i] * sy[0] + Hi[3 * i + 1] * sy[1] + Hi[3 * i + 2] * sy[2]);
        }
    }
    // pass 2: per row u = J dx, v = J y and the two gradients
    const double K = wd[5];
    for (int k = o0; k < o1; k++) {
        AdjRow R;
        double g0 = 0.0, g1 = 0.0, gwk = 0.0;
        if (adj_row(A, k, h, W.slot0, wp, sh64, sh, wc, wd, lm, R)) {
            const int ot = A.slot_off[W.slot0 + R.t];
            double u[2] = {0, 0}, v[2] = {0, 0};
            for (int r = 0; r < R.nd; r++) {
                for (int i = 0; i < 3; i++) {
                    u[r] += R.Jl[r][i] * dl[i];
                    v[r] += R.Jl[r][i] * yl[i];
                }
                if (R.t != h)
                    for (int j = 0; j < 6; j++) {
                        u[r] += R.Jt[r][j] * dxw[ot + j] + R.Jh[r][j] * dxw[oh + j];
                        v[r] += R.Jt[r][j] * yw[ot + j] + R.Jh[r][j] * yw[oh + j];
                    }
            }
            double s = 0.0;
            for (int r = 0; r < R.nd; r++) s += v[r] * (u[r] + R.e[r]);
            double dw[2] = {0, 0};
            if (R.en > K) {
                const double c3 = -K / (R.en * R.en * R.en);
                for (int r = 0; r < R.nd; r++) dw[r] = c3 * R.e[r];
            }
            g0 = R.om * R.isg * (dw[0] * s + R.wh * v[0]);
            g1 = R.nd == 2 ? R.om * R.isg * (dw[1] * s + R.wh * v[1]) : 0.0;
            gwk = -R.wh * s;
        }
        gz[2 * (size_t)k] = g0;
        gz[2 * (size_t)k + 1] = g1;
        gw[k] = gwk;
    }
}

}  // namespace xba
