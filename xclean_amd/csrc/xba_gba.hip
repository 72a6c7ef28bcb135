// Run-end global BA in inverse depth on the device (backend/global_ba.py:694-1785 global_ba_id, default path) with
// its reduced camera system solved by two-level deflated PCG (backend/gba_linsolve.py:129-478).
//
// Layout in HBM (FP64 unless noted):
//   state        Tcw (nk, 16) row-major 4x4, rho (nl), mb (nl, 3); trial and best copies of each
//   rows         grouped by landmark: lstart (nl + 1) int32, per row rkf / rlm int32, ruv (2), rd (disparity)
//   visits       per keyframe the (landmark, role) pairs that touch it: vstart (nk + 1), visit int32 (row k >= 0 for
//                an observer visit, -(l + 1) for the host visit of landmark l)
//   reduced H    block-sparse rows of 6x6 blocks, both triangles: rowptr (nk + 1), col (nnzb) int32, Hb (nnzb, 36)
//                row-major; bp (6 nk)
//   landmarks    Hlli (nl, 6) packed symmetric damped inverse, bl (nl, 3), kept for the back-substitution
//
// Kernels of one LM iteration:
//   k_gba_assemble  one workgroup per chunk of a block row p: every visit of keyframe p re-linearises its landmark
//                   (two passes over the landmark's rows, no per-row intermediates in HBM), inverts the damped 3x3
//                   landmark block in registers and adds the landmark's Hpp and Schur terms of row p into the row's
//                   blocks staged in LDS (ds_add_f64); the row is written to HBM once
//   k_gba_gauge     gauge rows / columns of fixed keyframes to identity, then the Marquardt diagonal
//   (DPCG)          the deflated PCG of the damped system (below)
//   k_gba_backsub   per landmark: dl = Hlli (bl - sum_rows Jl^T (Jo dx_o + Jh dx_h)), the trial landmark state
//   k_gba_trial     per keyframe: T <- exp(dx) T for the free ones
//   k_gba_cost      per landmark: sum of squared pixel residuals plus the Huber-weighted disparity chi2, reduced
//                   in a fixed order (deterministic)
#pragma once
#include <chrono>
#include <cmath>
#include <vector>

#include "xba_internal.h"

namespace gba {

#define GBA_WG 256
#define GBA_SEG_MAX 23            // keyframes per segment: a 138 x 138 FP64 segment block fits one workgroup's LDS
#define GBA_LDS_CAP (150 * 1024)  // LDS of one assembly chunk (accumulators + column list)

struct Dev {
    int nk, nl, N;
    const double *Tcw, *rho, *mb;  // state the kernels read
    const int32_t *host, *lstart, *rkf, *rlm;
    const double *ruv, *rd, *vtw;  // rd, vtw may be null
    const uint8_t *free_mask;
    const double *pm0, *prho0, *pwb, *pisg, *pval;  // null: no prior
    double fx, cx, cy, fB, isg_s, huber_k, lam;
    const int32_t *vstart, *visit, *rowptr, *col;
    double *Hb, *bp, *Hlli, *bl;
};

__device__ inline double huber_w(double a, double k) { return a <= k ? 1.0 : k / fmax(a, 1e-9); }

// host camera of a landmark: R_wc = R_cw^T, t_wc = -R_cw^T t_cw, and the host-frame point Ph = m / rho
struct HostFrame {
    double R[9], t[3], Ph[3], Pw[3], irho;
};

__device__ inline void host_frame(const Dev &G, int l, HostFrame &h) {
    const double *T = G.Tcw + 16 * (size_t)G.host[l];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) h.R[3 * i + j] = T[4 * j + i];
    for (int i = 0; i < 3; i++) h.t[i] = -(h.R[3 * i] * T[3] + h.R[3 * i + 1] * T[7] + h.R[3 * i + 2] * T[11]);
    const double rr = G.rho[l];
    h.irho = 1.0 / fmax(rr, 1e-12);
    for (int i = 0; i < 3; i++) h.Ph[i] = G.mb[3 * l + i] * h.irho;
    for (int i = 0; i < 3; i++) h.Pw[i] = h.R[3 * i] * h.Ph[0] + h.R[3 * i + 1] * h.Ph[1] + h.R[3 * i + 2] * h.Ph[2] + h.t[i];
}

// One observation row linearised (global_ba.py:1003-1046 and, for a disparity, 1063-1083): whitened Jacobians
// with respect to the observer twist (Jo), the host twist (Jh) and the landmark (a, b, rho) (Jl); rows 0-1 the
// pixel, row 2 the disparity (zero when the row has none).
struct Row {
    double Jo[3][6], Jh[3][6], Jl[3][3], rw[3];
};

__device__ inline void lin_row(const Dev &G, const HostFrame &h, int k, Row &o) {
    const int kf = G.rkf[k];
    const double *T = G.Tcw + 16 * (size_t)kf;
    double Pc[3];
    for (int i = 0; i < 3; i++) Pc[i] = T[4 * i] * h.Pw[0] + T[4 * i + 1] * h.Pw[1] + T[4 * i + 2] * h.Pw[2] + T[4 * i + 3];
    const double x = Pc[0], y = Pc[1], z = fmax(Pc[2], 1e-6), iz = 1.0 / z;
    const double r0 = G.fx * x * iz + G.cx - G.ruv[2 * k], r1 = G.fx * y * iz + G.cy - G.ruv[2 * k + 1];
    double w = huber_w(sqrt(r0 * r0 + r1 * r1), G.huber_k);
    if (G.vtw) w *= G.vtw[kf];
    const double sw = sqrt(w);
    // A = d proj / d Pc
    const double a00 = G.fx * iz, a02 = -G.fx * x * iz * iz, a12 = -G.fx * y * iz * iz;
    // Rch = R_cw(obs) R_wc(host)
    double Rch[9];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
            Rch[3 * i + j] = T[4 * i] * h.R[j] + T[4 * i + 1] * h.R[3 + j] + T[4 * i + 2] * h.R[6 + j];
    // dPc/d(obs twist) = [I | -[P]x], P = (x, y, z clamped); dPc/d(host twist) = Rch [-I | [Ph]x];
    // dPc/d(a, b, rho) = [Rch e0 / rho, Rch e1 / rho, -Rch Ph / rho]
    double dPo[3][6] = {{1, 0, 0, 0, z, -y}, {0, 1, 0, -z, 0, x}, {0, 0, 1, y, -x, 0}};
    const double *P = h.Ph;
    const double hx[9] = {0, -P[2], P[1], P[2], 0, -P[0], -P[1], P[0], 0};
    double dPh[3][6], dPl[3][3];
    for (int i = 0; i < 3; i++) {
        for (int j = 0; j < 3; j++) {
            dPh[i][j] = -Rch[3 * i + j];
            dPh[i][3 + j] = Rch[3 * i] * hx[j] + Rch[3 * i + 1] * hx[3 + j] + Rch[3 * i + 2] * hx[6 + j];
        }
        dPl[i][0] = Rch[3 * i] * h.irho;
        dPl[i][1] = Rch[3 * i + 1] * h.irho;
        dPl[i][2] = -(Rch[3 * i] * P[0] + Rch[3 * i + 1] * P[1] + Rch[3 * i + 2] * P[2]) * h.irho;
    }
    for (int j = 0; j < 6; j++) {
        o.Jo[0][j] = (a00 * dPo[0][j] + a02 * dPo[2][j]) * sw;
        o.Jo[1][j] = (a00 * dPo[1][j] + a12 * dPo[2][j]) * sw;
        o.Jh[0][j] = (a00 * dPh[0][j] + a02 * dPh[2][j]) * sw;
        o.Jh[1][j] = (a00 * dPh[1][j] + a12 * dPh[2][j]) * sw;
    }
    for (int j = 0; j < 3; j++) {
        o.Jl[0][j] = (a00 * dPl[0][j] + a02 * dPl[2][j]) * sw;
        o.Jl[1][j] = (a00 * dPl[1][j] + a12 * dPl[2][j]) * sw;
    }
    o.rw[0] = r0 * sw;
    o.rw[1] = r1 * sw;
    const double d = G.rd ? G.rd[k] : 0.0;
    if (d > 1e-6) {
        const double rs = (iz - d / G.fB) * G.isg_s;
        const double sws = sqrt(huber_w(fabs(rs), G.huber_k));
        const double cz = -iz * iz * G.isg_s * sws;
        for (int j = 0; j < 6; j++) {
            o.Jo[2][j] = dPo[2][j] * cz;
            o.Jh[2][j] = dPh[2][j] * cz;
        }
        for (int j = 0; j < 3; j++) o.Jl[2][j] = dPl[2][j] * cz;
        o.rw[2] = rs * sws;
    } else {
        for (int j = 0; j < 6; j++) o.Jo[2][j] = o.Jh[2][j] = 0.0;
        for (int j = 0; j < 3; j++) o.Jl[2][j] = 0.0;
        o.rw[2] = 0.0;
    }
}

// landmark prior rows (global_ba.py:540-566 and 1158-1166): Hll diagonal and bl
__device__ inline void prior_terms(const Dev &G, int l, double *Hll6, double *bl) {
    if (!G.pm0 || !(G.pval[l] > 0.5)) return;
    const double wb = G.pwb[l], ig = G.pisg[l];
    const double rp[3] = {(G.mb[3 * l] - G.pm0[2 * l]) * wb, (G.mb[3 * l + 1] - G.pm0[2 * l + 1]) * wb,
                          (G.rho[l] - G.prho0[l]) * ig};
    const double J0[3] = {wb, wb, ig};
    const double sw[3] = {1.0, 1.0, sqrt(huber_w(fabs(rp[2]), G.huber_k))};
    const int dg[3] = {0, 3, 5};
    for (int c = 0; c < 3; c++) {
        const double j = J0[c] * sw[c];
        Hll6[dg[c]] += j * j;
        bl[c] -= j * rp[c] * sw[c];
    }
}

// packed symmetric 3x3: 0 (0,0) 1 (0,1) 2 (0,2) 3 (1,1) 4 (1,2) 5 (2,2)
__device__ inline void damped_inverse(const double *H, double lam, double *Hi) {
    const double s = lam * fmax((H[0] + H[3] + H[5]) / 3.0, 1e-9);
    const double a = H[0] + s, b = H[1], c = H[2], d = H[3] + s, e = H[4], f = H[5] + s;
    const double c00 = d * f - e * e, c01 = c * e - b * f, c02 = b * e - c * d;
    const double det = a * c00 + b * c01 + c * c02, id = 1.0 / det;
    Hi[0] = c00 * id;
    Hi[1] = c01 * id;
    Hi[2] = c02 * id;
    Hi[3] = (a * f - c * c) * id;
    Hi[4] = (b * c - a * e) * id;
    Hi[5] = (a * d - b * b) * id;
}

__device__ inline double sym3(const double *S, int i, int j) {
    const int idx[9] = {0, 1, 2, 1, 3, 4, 2, 4, 5};
    return S[idx[3 * i + j]];
}

// ------------------------------------------------------------------------------------------- set-up kernels --
// rows kept: not at the landmark's own host (global_ba.py:733-736), indices in range (err |= 1 otherwise)
__global__ void k_gba_count(int N, int nk, int nl, const int32_t *okf, const int32_t *olm, const int32_t *host,
                            const uint8_t *keep_in, int32_t *lcount, int32_t *err) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= N) return;
    const int o = okf[k], l = olm[k];
    if (o < 0 || o >= nk || l < 0 || l >= nl) {
        atomicOr(err, 1);
        return;
    }
    if (keep_in && !keep_in[k]) return;
    if (host[l] == o) return;
    atomicAdd(lcount + l, 1);
}

// exclusive scan in place of n int32 counts into n + 1 offsets (one workgroup of 1024)
__global__ __launch_bounds__(1024) void k_gba_scan(int n, int32_t *a) {
    __shared__ int32_t part[1024];
    __shared__ int32_t carry;
    const int tid = threadIdx.x;
    if (tid == 0) carry = 0;
    __syncthreads();
    for (int base = 0; base <= n; base += 1024) {
        const int i = base + tid;
        const int32_t v = i < n ? a[i] : 0;
        part[tid] = v;
        __syncthreads();
        for (int d = 1; d < 1024; d <<= 1) {
            const int32_t add = tid >= d ? part[tid - d] : 0;
            __syncthreads();
            part[tid] += add;
            __syncthreads();
        }
        const int32_t c0 = carry;
        if (i <= n) a[i] = c0 + part[tid] - v;
        __syncthreads();
        if (tid == 1023) carry = c0 + part[1023];
        __syncthreads();
    }
}

// every landmark's host keyframe in range (err |= 1 otherwise)
__global__ void k_gba_check_host(int nl, int nk, const int32_t *host, int32_t *err) {
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l < nl && (host[l] < 0 || host[l] >= nk)) atomicOr(err, 1);
}

// every pose factor's keyframes in range (err |= 1 otherwise)
__global__ void k_gba_check_pairs(int nf, int nk, const int32_t *ab, int32_t *err) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nf) return;
    const int a = ab[2 * f], b = ab[2 * f + 1];
    if (a < 0 || a >= nk || b < 0 || b >= nk) atomicOr(err, 1);
}

// every kept row to its landmark's segment (arbitrary order; k_gba_sort_rows orders each segment by input index)
__global__ void k_gba_scatter(int N, const int32_t *okf, const int32_t *olm, const int32_t *host,
                              const uint8_t *keep_in, int32_t *fill, int32_t *src) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= N) return;
    const int o = okf[k], l = olm[k];
    if (keep_in && !keep_in[k]) return;
    if (host[l] == o) return;
    src[atomicAdd(fill + l, 1)] = k;
}

__global__ void k_gba_sort_rows(int nl, const int32_t *lstart, int32_t *src) {
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= nl) return;
    const int a = lstart[l], b = lstart[l + 1];
    for (int i = a + 1; i < b; i++) {
        const int v = src[i];
        int j = i - 1;
        while (j >= a && src[j] > v) {
            src[j + 1] = src[j];
            j--;
        }
        src[j + 1] = v;
    }
}

// the grouped row arrays from the input rows
__global__ void k_gba_gather(int M, const int32_t *src, const int32_t *okf, const int32_t *olm, const double *ouv,
                             const double *odisp, int32_t *rkf, int32_t *rlm, double *ruv, double *rd, int32_t *rsrc) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= M) return;
    const int s = src[k];
    rkf[k] = okf[s];
    rlm[k] = olm[s];
    ruv[2 * k] = ouv[2 * s];
    ruv[2 * k + 1] = ouv[2 * s + 1];
    if (rd) rd[k] = odisp ? odisp[s] : 0.0;
    if (rsrc) rsrc[k] = s;
}

// visits: one per row (its observer) and one per landmark (its host)
__global__ void k_gba_visit_count(int M, int nl, const int32_t *rkf, const int32_t *host, int32_t *vcount) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < M) atomicAdd(vcount + rkf[i], 1);
    else if (i < M + nl) atomicAdd(vcount + host[i - M], 1);
}

__global__ void k_gba_visit_fill(int M, int nl, const int32_t *rkf, const int32_t *host, int32_t *vfill,
                                 int32_t *visit) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < M) visit[atomicAdd(vfill + rkf[i], 1)] = i;
    else if (i < M + nl) visit[atomicAdd(vfill + host[i - M], 1)] = -(i - M) - 1;
}

// block pattern: bit (p, q) of an nk x nk bitmap (words of 32 columns) for every pair of keyframes one landmark
// touches (its host and its observers), and every diagonal
__global__ void k_gba_pattern(int nk, int nl, const int32_t *lstart, const int32_t *rkf, const int32_t *host,
                              uint32_t *bits, int64_t wpr) {
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l < nk) atomicOr(bits + (size_t)l * wpr + (l >> 5), 1u << (l & 31));
    if (l >= nl) return;
    const int a = lstart[l], b = lstart[l + 1], h = host[l];
    for (int i = a - 1; i < b; i++) {
        const int p = i < a ? h : rkf[i];
        for (int j = a - 1; j < b; j++) {
            const int q = j < a ? h : rkf[j];
            atomicOr(bits + (size_t)p * wpr + (q >> 5), 1u << (q & 31));
        }
    }
}

// blocks per row (one wavefront per row)
__global__ void k_gba_rowcount(int nk, const uint32_t *bits, int64_t wpr, int32_t *rcount) {
    const int p = blockIdx.x * (blockDim.x / XBA_WAVE) + threadIdx.x / XBA_WAVE;
    const int lane = threadIdx.x & (XBA_WAVE - 1);
    if (p >= nk) return;
    int c = 0;
    for (int64_t w = lane; w < wpr; w += XBA_WAVE) c += __popc(bits[(size_t)p * wpr + w]);
    for (int d = XBA_WAVE / 2; d > 0; d >>= 1) c += __shfl_xor(c, d, XBA_WAVE);
    if (lane == 0) rcount[p] = c;
}

// column indices of every row, ascending (one thread per row; rows are short relative to nk / 32 words)
__global__ void k_gba_rowfill(int nk, const uint32_t *bits, int64_t wpr, const int32_t *rowptr, int32_t *col) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= nk) return;
    int o = rowptr[p];
    for (int64_t w = 0; w < wpr; w++) {
        uint32_t v = bits[(size_t)p * wpr + w];
        while (v) {
            const int b = __ffs(v) - 1;
            col[o++] = (int)(32 * w + b);
            v &= v - 1;
        }
    }
}


// --------------------------------------------------------------------------------------- the reduced system --
__device__ inline int sym6(int i, int j) {  // packed upper triangle of a symmetric 6x6 (i <= j)
    return i * 6 - i * (i - 1) / 2 + (j - i);
}

__device__ inline int find_col(const int *scol, int n, int q) {
    int lo = 0, hi = n;
    while (lo < hi) {
        const int m = (lo + hi) >> 1;
        if (scol[m] < q) lo = m + 1;
        else hi = m;
    }
    return (lo < n && scol[lo] == q) ? lo : -1;
}

// One workgroup per chunk (p, c0, c1, owns_b): the blocks [c0, c1) of block row p (global_ba.py:1176-1180 and
// 1326-1334 for this row): every landmark that touches keyframe p adds, through each of its visits,
//   host visit (p = h):      (h,h) += sum_t Jh_t^T Jh_t - W_h Hlli W_h^T,   (h,o_t) += Jh_t^T Jo_t - W_h Hlli W_t^T,
//                            b_h += -sum_t Jh_t^T r_t - W_h Hlli bl
//   observer visit (row k):  (p,p) += Jo_k^T Jo_k - W_k Hlli W_k^T,  (p,h) += Jo_k^T Jh_k - W_k Hlli W_h^T,
//                            (p,o_t) += -W_k Hlli W_t^T (t != k),   b_p += -Jo_k^T r_k - W_k Hlli bl
// with W_t = Jo_t^T Jl_t, W_h = sum_t Jh_t^T Jl_t and Hlli the damped inverse of the landmark block.
// Per-row products of one LM iteration, written once by k_gba_rows (one thread per landmark, every row linearised
// once) and gathered by the block rows in k_gba_assemble:
//   RR_W   Jo^T Jl (6x3)        RR_OO  Jo^T Jo (packed upper 21)   RR_OH  Jo^T Jh (6x6)
//   RR_HH  Jh^T Jh (packed 21)  RR_GO  -Jo^T r                     RR_GH  -Jh^T r
// and per landmark Hlli (packed 6, damped inverse), bl (3) and W_h = sum_t Jh_t^T Jl_t (6x3).
enum { RR_W = 0, RR_OO = 18, RR_OH = 39, RR_HH = 75, RR_GO = 96, RR_GH = 102, RR_N = 108 };

__global__ __launch_bounds__(128) void k_gba_rows(Dev G, double *rr, double *wh) {
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= G.nl) return;
    HostFrame hf;
    host_frame(G, l, hf);
    double Hll[6] = {0, 0, 0, 0, 0, 0}, bl[3] = {0, 0, 0}, Wh[18];
    for (int i = 0; i < 18; i++) Wh[i] = 0.0;
    Row R;
    for (int k = G.lstart[l]; k < G.lstart[l + 1]; k++) {
        lin_row(G, hf, k, R);
        double *o = rr + RR_N * (size_t)k;
        for (int m = 0; m < 3; m++) {
            const double *jl = R.Jl[m];
            Hll[0] += jl[0] * jl[0];
            Hll[1] += jl[0] * jl[1];
            Hll[2] += jl[0] * jl[2];
            Hll[3] += jl[1] * jl[1];
            Hll[4] += jl[1] * jl[2];
            Hll[5] += jl[2] * jl[2];
            for (int c = 0; c < 3; c++) bl[c] -= jl[c] * R.rw[m];
        }
        for (int i = 0; i < 6; i++) {
            for (int c = 0; c < 3; c++) {
                o[RR_W + 3 * i + c] = R.Jo[0][i] * R.Jl[0][c] + R.Jo[1][i] * R.Jl[1][c] + R.Jo[2][i] * R.Jl[2][c];
                Wh[3 * i + c] += R.Jh[0][i] * R.Jl[0][c] + R.Jh[1][i] * R.Jl[1][c] + R.Jh[2][i] * R.Jl[2][c];
            }
            for (int j = i; j < 6; j++) {
                o[RR_OO + sym6(i, j)] = R.Jo[0][i] * R.Jo[0][j] + R.Jo[1][i] * R.Jo[1][j] + R.Jo[2][i] * R.Jo[2][j];
                o[RR_HH + sym6(i, j)] = R.Jh[0][i] * R.Jh[0][j] + R.Jh[1][i] * R.Jh[1][j] + R.Jh[2][i] * R.Jh[2][j];
            }
            for (int j = 0; j < 6; j++)
                o[RR_OH + 6 * i + j] = R.Jo[0][i] * R.Jh[0][j] + R.Jo[1][i] * R.Jh[1][j] + R.Jo[2][i] * R.Jh[2][j];
            o[RR_GO + i] = -(R.Jo[0][i] * R.rw[0] + R.Jo[1][i] * R.rw[1] + R.Jo[2][i] * R.rw[2]);
            o[RR_GH + i] = -(R.Jh[0][i] * R.rw[0] + R.Jh[1][i] * R.rw[1] + R.Jh[2][i] * R.rw[2]);
        }
    }
    prior_terms(G, l, Hll, bl);
    double Hi[6];
    damped_inverse(Hll, G.lam, Hi);
    for (int i = 0; i < 6; i++) G.Hlli[6 * (size_t)l + i] = Hi[i];
    for (int i = 0; i < 3; i++) G.bl[3 * (size_t)l + i] = bl[i];
    for (int i = 0; i < 18; i++) wh[18 * (size_t)l + i] = Wh[i];
}

// One workgroup per chunk (p, c0, c1, owns_b): the blocks [c0, c1) of block row p (global_ba.py:1176-1180 and
// 1326-1334 for this row), from every visit of keyframe p:
//   host visit (p = h):      (h,h) += sum_t Jh_t^T Jh_t - W_h Hlli W_h^T,   (h,o_t) += Jh_t^T Jo_t - W_h Hlli W_t^T,
//                            b_h += -sum_t Jh_t^T r_t - W_h Hlli bl
//   observer visit (row k):  (p,p) += Jo_k^T Jo_k - W_k Hlli W_k^T,  (p,h) += Jo_k^T Jh_k - W_k Hlli W_h^T,
//                            (p,o_t) += -W_k Hlli W_t^T (t != k),   b_p += -Jo_k^T r_k - W_k Hlli bl
// with W_t = Jo_t^T Jl_t.  The row's blocks are staged in LDS (ds_add_f64; the diagonal block as its upper triangle,
// mirrored when the row is written) and written to HBM once.
__global__ __launch_bounds__(GBA_WG) void k_gba_assemble(Dev G, const int4 *chunks, const double *rr,
                                                         const double *whs) {
    extern __shared__ double smem[];
    const int4 ch = chunks[blockIdx.x];
    const int p = ch.x, c0 = ch.y, nb = ch.z - ch.y;
    const bool owns_b = ch.w != 0;
    const int tid = threadIdx.x, nt = blockDim.x;
    double *sacc = smem;
    double *sbv = sacc + 36 * nb;
    int *scol = (int *)(sbv + 6);
    const int r0 = G.rowptr[p];
    for (int i = tid; i < nb; i += nt) scol[i] = G.col[r0 + c0 + i];
    for (int i = tid; i < 36 * nb + 6; i += nt) sacc[i] = 0.0;
    __syncthreads();
    const int jpp = find_col(scol, nb, p);

    for (int v = G.vstart[p] + tid; v < G.vstart[p + 1]; v += nt) {
        const int vv = G.visit[v];
        const bool is_host = vv < 0;
        const int l = is_host ? -vv - 1 : G.rlm[vv];
        const int a = G.lstart[l], b = G.lstart[l + 1];
        double Hi[6], bl[3], Wh[18], W[18], Y[18], Dg[21], gv[6];
        for (int i = 0; i < 6; i++) Hi[i] = G.Hlli[6 * (size_t)l + i];
        for (int i = 0; i < 3; i++) bl[i] = G.bl[3 * (size_t)l + i];
        for (int i = 0; i < 18; i++) Wh[i] = whs[18 * (size_t)l + i];
        const double *own = is_host ? nullptr : rr + RR_N * (size_t)vv;
        for (int i = 0; i < 18; i++) W[i] = is_host ? Wh[i] : own[RR_W + i];
        for (int i = 0; i < 6; i++)
            for (int c = 0; c < 3; c++)
                Y[3 * i + c] = W[3 * i] * sym3(Hi, 0, c) + W[3 * i + 1] * sym3(Hi, 1, c) + W[3 * i + 2] * sym3(Hi, 2, c);
        for (int i = 0; i < 6; i++) {
            for (int j = i; j < 6; j++)
                Dg[sym6(i, j)] = (is_host ? 0.0 : own[RR_OO + sym6(i, j)]) -
                                 (Y[3 * i] * W[3 * j] + Y[3 * i + 1] * W[3 * j + 1] + Y[3 * i + 2] * W[3 * j + 2]);
            gv[i] = (is_host ? 0.0 : own[RR_GO + i]) - (Y[3 * i] * bl[0] + Y[3 * i + 1] * bl[1] + Y[3 * i + 2] * bl[2]);
        }
        if (!is_host && G.host[l] > p) {
            const int jph = find_col(scol, nb, G.host[l]);
            if (jph >= 0)
                for (int i = 0; i < 6; i++)
                    for (int j = 0; j < 6; j++)
                        atomicAdd(sacc + 36 * jph + 6 * i + j,
                                  own[RR_OH + 6 * i + j] -
                                      (Y[3 * i] * Wh[3 * j] + Y[3 * i + 1] * Wh[3 * j + 1] + Y[3 * i + 2] * Wh[3 * j + 2]));
        }
        for (int k = a; k < b; k++) {
            if (k == vv) continue;
            const double *t = rr + RR_N * (size_t)k;
            if (is_host) {
                for (int i = 0; i < 21; i++) Dg[i] += t[RR_HH + i];
                for (int i = 0; i < 6; i++) gv[i] += t[RR_GH + i];
            }
            const int q = G.rkf[k];
            if (q < p) continue;  // the lower blocks are the transposes of upper ones (k_gba_mirror)
            const int jq = find_col(scol, nb, q);
            if (jq < 0) continue;
            // a second row of this keyframe lands on the diagonal block, upper triangle only: the visit of that row
            // adds the transposed term of the pair
            const bool diag = q == p;
            for (int i = 0; i < 6; i++)
                for (int j = diag ? i : 0; j < 6; j++) {
                    double v2 = -(Y[3 * i] * t[RR_W + 3 * j] + Y[3 * i + 1] * t[RR_W + 3 * j + 1] +
                                  Y[3 * i + 2] * t[RR_W + 3 * j + 2]);
                    if (is_host) v2 += t[RR_OH + 6 * j + i];
                    atomicAdd(sacc + 36 * jq + 6 * i + j, v2);
                }
        }
        if (jpp >= 0)
            for (int i = 0; i < 6; i++)
                for (int j = i; j < 6; j++) atomicAdd(sacc + 36 * jpp + 6 * i + j, Dg[sym6(i, j)]);
        if (owns_b)
            for (int i = 0; i < 6; i++) atomicAdd(sbv + i, gv[i]);
    }
    __syncthreads();
    double *out = G.Hb + 36 * (size_t)(r0 + c0);
    for (int i = tid; i < 36 * nb; i += nt) {
        const int bk = i / 36, e = i - 36 * bk, r = e / 6, c = e - 6 * r;
        out[i] = (bk == jpp && r > c) ? sacc[36 * bk + 6 * c + r] : sacc[i];
    }
    if (owns_b && tid < 6) G.bp[6 * (size_t)p + tid] = sbv[tid];
}

// the blocks below the diagonal: H_qp = H_pq^T from the upper block of row p (one thread per block)
__global__ void k_gba_mirror(int64_t nnzb, const int32_t *brow, const int32_t *rowptr, const int32_t *col, double *Hb) {
    const int64_t bi = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (bi >= nnzb) return;
    const int q = brow[bi], p = col[bi];
    if (p >= q) return;
    int lo = rowptr[p], hi = rowptr[p + 1];
    while (lo < hi) {
        const int m = (lo + hi) >> 1;
        if (col[m] < q) lo = m + 1;
        else hi = m;
    }
    const double *U = Hb + 36 * (size_t)lo;
    double *Lb = Hb + 36 * bi;
    for (int i = 0; i < 6; i++)
        for (int j = 0; j < 6; j++) Lb[6 * i + j] = U[6 * j + i];
}

// keyframe pairs of the pose factors into the block pattern: (a,a), (a,b), (b,a), (b,b)
__global__ void k_gba_pattern_pairs(int nf, const int32_t *ab, uint32_t *bits, int64_t wpr) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nf) return;
    const int a = ab[2 * f], b = ab[2 * f + 1];
    const int p[2] = {a, b};
    for (int i = 0; i < 2; i++)
        for (int j = 0; j < 2; j++) atomicOr(bits + (size_t)p[i] * wpr + (p[j] >> 5), 1u << (p[j] & 31));
}

__device__ inline double *block_of(const Dev &G, int p, int q) {
    int lo = G.rowptr[p], hi = G.rowptr[p + 1];
    while (lo < hi) {
        const int m = (lo + hi) >> 1;
        if (G.col[m] < q) lo = m + 1;
        else hi = m;
    }
    return G.Hb + 36 * (size_t)lo;
}

// the pose factors' Hessian blocks and gradient, added to the assembled system (global atomics; one thread per
// factor).  Rotation factor (a, b, M, sig) (global_ba.py:1921-1960): e = Log(M^T R_a R_b^T), w = 1 / max(sig, 1e-9)
// (Huber-reduced by sqrt(k / (w |e|)) above k), Ja = w Jr^-1(e) A^T, Jb = -w Jr^-1(e) on the rotation part of each
// twist.  Weld (a, b, Z, W) (global_ba.py:2034-2067): e = Log_SE3(Z^-1 T_cw_a T_wc_b), Ja = Ad(T_cw_b T_wc_a),
// Jb = -I on the whole twist.
__global__ void k_gba_pose_factors(Dev G, int n_rot, const int32_t *rot_ab, const double *rot_M, const double *rot_sig,
                                   double rot_huber, int n_weld, const int32_t *weld_ab, const double *weld_Z,
                                   const double *weld_w) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= n_rot + n_weld) return;
    if (f < n_rot) {
        const int a = rot_ab[2 * f], b = rot_ab[2 * f + 1];
        const double *Ta = G.Tcw + 16 * (size_t)a, *Tb = G.Tcw + 16 * (size_t)b;
        double Ra[9], Rb[9], Am[9], E[9], e[3], Jr[9], Ja[9], Jb[9];
        for (int i = 0; i < 3; i++)
            for (int c = 0; c < 3; c++) {
                Ra[3 * i + c] = Ta[4 * i + c];
                Rb[3 * i + c] = Tb[4 * i + c];
            }
        xba::mmt3(Ra, Rb, Am);
        xba::mtm3(rot_M + 9 * (size_t)f, Am, E);
        xba::log_quat(E, e);
        double w = 1.0 / fmax(rot_sig[f], 1e-9);
        if (rot_huber > 0.0) {
            const double en = w * sqrt(e[0] * e[0] + e[1] * e[1] + e[2] * e[2]);
            if (en > rot_huber) w *= sqrt(rot_huber / en);
        }
        xba::jr_inv(e, Jr);
        xba::mmt3(Jr, Am, Ja);
        for (int i = 0; i < 9; i++) {
            Ja[i] *= w;
            Jb[i] = -w * Jr[i];
        }
        const double r[3] = {w * e[0], w * e[1], w * e[2]};
        const int p[2] = {a, b};
        const double *J[2] = {Ja, Jb};
        for (int u = 0; u < 2; u++) {
            for (int i = 0; i < 3; i++) {
                double gi = 0.0;
                for (int m = 0; m < 3; m++) gi += J[u][3 * m + i] * r[m];
                atomicAdd(G.bp + 6 * (size_t)p[u] + 3 + i, -gi);
            }
            for (int v = 0; v < 2; v++) {
                double *B = block_of(G, p[u], p[v]);
                for (int i = 0; i < 3; i++)
                    for (int c = 0; c < 3; c++) {
                        double hv = 0.0;
                        for (int m = 0; m < 3; m++) hv += J[u][3 * m + i] * J[v][3 * m + c];
                        atomicAdd(B + 6 * (3 + i) + 3 + c, hv);
                    }
            }
        }
        return;
    }
    const int fw = f - n_rot;
    const int a = weld_ab[2 * fw], b = weld_ab[2 * fw + 1];
    const double *Ta = G.Tcw + 16 * (size_t)a, *Tb = G.Tcw + 16 * (size_t)b, *Z = weld_Z + 16 * (size_t)fw;
    // rel = T_cw_a T_wc_b; X = Z^-1 rel
    double Rb_t[9], tb_w[3], Rrel[9], trel[3];
    for (int i = 0; i < 3; i++)
        for (int c = 0; c < 3; c++) Rb_t[3 * i + c] = Tb[4 * c + i];
    for (int i = 0; i < 3; i++) tb_w[i] = -(Rb_t[3 * i] * Tb[3] + Rb_t[3 * i + 1] * Tb[7] + Rb_t[3 * i + 2] * Tb[11]);
    for (int i = 0; i < 3; i++) {
        for (int c = 0; c < 3; c++)
            Rrel[3 * i + c] = Ta[4 * i] * Rb_t[c] + Ta[4 * i + 1] * Rb_t[3 + c] + Ta[4 * i + 2] * Rb_t[6 + c];
        trel[i] = Ta[4 * i] * tb_w[0] + Ta[4 * i + 1] * tb_w[1] + Ta[4 * i + 2] * tb_w[2] + Ta[4 * i + 3];
    }
    double RX[9], tX[3];
    for (int i = 0; i < 3; i++) {
        for (int c = 0; c < 3; c++)
            RX[3 * i + c] = Z[i] * Rrel[c] + Z[4 + i] * Rrel[3 + c] + Z[8 + i] * Rrel[6 + c];  // Z_R^T Rrel
        tX[i] = Z[i] * (trel[0] - Z[3]) + Z[4 + i] * (trel[1] - Z[7]) + Z[8 + i] * (trel[2] - Z[11]);
    }
    double e[6], phi[3], mphi[3], Vi[9];
    xba::log_quat(RX, phi);
    for (int i = 0; i < 3; i++) mphi[i] = -phi[i];
    xba::jr_inv(mphi, Vi);  // V(phi)^-1 = Jr^-1(-phi)
    for (int i = 0; i < 3; i++) e[i] = Vi[3 * i] * tX[0] + Vi[3 * i + 1] * tX[1] + Vi[3 * i + 2] * tX[2];
    for (int i = 0; i < 3; i++) e[3 + i] = phi[i];
    // Ja = Ad(T_cw_b T_wc_a): R = R_b R_a^T, t = t_b - R t_a
    double R[9], tt[3], tx[9], Ja[36];
    for (int i = 0; i < 3; i++)
        for (int c = 0; c < 3; c++) R[3 * i + c] = Tb[4 * i] * Ta[4 * c] + Tb[4 * i + 1] * Ta[4 * c + 1] + Tb[4 * i + 2] * Ta[4 * c + 2];
    for (int i = 0; i < 3; i++) tt[i] = Tb[4 * i + 3] - (R[3 * i] * Ta[3] + R[3 * i + 1] * Ta[7] + R[3 * i + 2] * Ta[11]);
    const double hx[9] = {0, -tt[2], tt[1], tt[2], 0, -tt[0], -tt[1], tt[0], 0};
    xba::mm3(hx, R, tx);
    for (int i = 0; i < 36; i++) Ja[i] = 0.0;
    for (int i = 0; i < 3; i++)
        for (int c = 0; c < 3; c++) {
            Ja[6 * i + c] = R[3 * i + c];
            Ja[6 * i + 3 + c] = tx[3 * i + c];
            Ja[6 * (3 + i) + 3 + c] = R[3 * i + c];
        }
    const double *Wd = weld_w + 6 * (size_t)fw;
    double We[6];
    for (int i = 0; i < 6; i++) We[i] = Wd[i] * e[i];
    // gradient: -Ja^T W e into a, +W e into b (Jb = -I)
    for (int i = 0; i < 6; i++) {
        double gi = 0.0;
        for (int m = 0; m < 6; m++) gi += Ja[6 * m + i] * We[m];
        atomicAdd(G.bp + 6 * (size_t)a + i, -gi);
        atomicAdd(G.bp + 6 * (size_t)b + i, We[i]);
    }
    double *Baa = block_of(G, a, a), *Bab = block_of(G, a, b), *Bba = block_of(G, b, a), *Bbb = block_of(G, b, b);
    for (int i = 0; i < 6; i++)
        for (int c = 0; c < 6; c++) {
            double haa = 0.0;
            for (int m = 0; m < 6; m++) haa += Ja[6 * m + i] * Wd[m] * Ja[6 * m + c];
            atomicAdd(Baa + 6 * i + c, haa);
            atomicAdd(Bab + 6 * i + c, -Ja[6 * c + i] * Wd[c]);  // Ja^T W (-I)
            atomicAdd(Bba + 6 * i + c, -Wd[i] * Ja[6 * i + c]);  // (-I) W Ja
            if (i == c) atomicAdd(Bbb + 6 * i + c, Wd[i]);
        }
}

// gauge rows and columns of the fixed keyframes to identity, their b to zero, then H += diag(lam max(diag, 1e-9))
// (global_ba.py:1326-1334)
__global__ void k_gba_gauge(Dev G, const int32_t *brow, int64_t nnzb) {
    const int64_t bi = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (bi < G.nk && !G.free_mask[bi])
        for (int i = 0; i < 6; i++) G.bp[6 * bi + i] = 0.0;
    if (bi >= nnzb) return;
    const int p = brow[bi], q = G.col[bi];
    double *B = G.Hb + 36 * bi;
    if (!G.free_mask[p] || !G.free_mask[q]) {
        for (int i = 0; i < 36; i++) B[i] = 0.0;
        if (p == q)
            for (int i = 0; i < 6; i++) B[7 * i] = 1.0;
    }
    if (p == q)
        for (int i = 0; i < 6; i++) B[7 * i] += G.lam * fmax(B[7 * i], 1e-9);
}

// landmark back-substitution and trial landmark state (global_ba.py:1381-1398):
// dl = Hlli (bl - sum_t Jl_t^T (Jo_t dx_o + Jh_t dx_h)); a, b += dl[0:2]; rho <- max(rho + dl[2], rho_floor)
__global__ __launch_bounds__(128) void k_gba_backsub(Dev G, const double *dx, double rho_floor, double *rho_n, double *mb_n, int32_t *bad) {
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= G.nl) return;
    HostFrame hf;
    host_frame(G, l, hf);
    const int h = G.host[l];
    double acc[3] = {0, 0, 0};
    Row R;
    for (int k = G.lstart[l]; k < G.lstart[l + 1]; k++) {
        lin_row(G, hf, k, R);
        const double *xo = dx + 6 * (size_t)G.rkf[k], *xh = dx + 6 * (size_t)h;
        for (int m = 0; m < 3; m++) {
            double t = 0.0;
            for (int j = 0; j < 6; j++) t += R.Jo[m][j] * xo[j] + R.Jh[m][j] * xh[j];
            for (int c = 0; c < 3; c++) acc[c] += R.Jl[m][c] * t;
        }
    }
    const double *Hi = G.Hlli + 6 * (size_t)l;
    double r[3];
    for (int c = 0; c < 3; c++) r[c] = G.bl[3 * (size_t)l + c] - acc[c];
    double dl[3];
    for (int c = 0; c < 3; c++) dl[c] = sym3(Hi, c, 0) * r[0] + sym3(Hi, c, 1) * r[1] + sym3(Hi, c, 2) * r[2];
    mb_n[3 * l] = G.mb[3 * l] + dl[0];
    mb_n[3 * l + 1] = G.mb[3 * l + 1] + dl[1];
    mb_n[3 * l + 2] = G.mb[3 * l + 2];
    rho_n[l] = fmax(G.rho[l] + dl[2], rho_floor);
    if (!(isfinite(dl[0]) && isfinite(dl[1]) && isfinite(dl[2]))) atomicOr(bad, 1);
}

// trial poses: T <- exp(dx) T for the free keyframes (global_ba.py:1393-1395), the others copied
__global__ __launch_bounds__(128) void k_gba_trial(int nk, const double *Tcw, const uint8_t *free_mask, const double *dx, double *Tn,
                            int32_t *bad) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nk) return;
    const double *T = Tcw + 16 * (size_t)k;
    double *O = Tn + 16 * (size_t)k;
    if (!free_mask[k]) {
        for (int i = 0; i < 16; i++) O[i] = T[i];
        return;
    }
    const double *xi = dx + 6 * (size_t)k;
    for (int i = 0; i < 6; i++)
        if (!isfinite(xi[i])) atomicOr(bad, 1);
    double a, b, c, K[9] = {0, -xi[5], xi[4], xi[5], 0, -xi[3], -xi[4], xi[3], 0}, KK[9], Rm[9], V[9];
    xba::abc(xi[3] * xi[3] + xi[4] * xi[4] + xi[5] * xi[5], &a, &b, &c);
    xba::mm3(K, K, KK);
    for (int i = 0; i < 9; i++) {
        const double e = (i % 4 == 0) ? 1.0 : 0.0;
        Rm[i] = e + a * K[i] + b * KK[i];
        V[i] = e + b * K[i] + c * KK[i];
    }
    double E[16] = {0};
    for (int i = 0; i < 3; i++) {
        for (int j = 0; j < 3; j++) E[4 * i + j] = Rm[3 * i + j];
        E[4 * i + 3] = V[3 * i] * xi[0] + V[3 * i + 1] * xi[1] + V[3 * i + 2] * xi[2];
    }
    E[15] = 1.0;
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) {
            double acc = 0.0;
            for (int m = 0; m < 4; m++) acc += E[4 * i + m] * T[4 * m + j];
            O[4 * i + j] = acc;
        }
}

// cost at a state: per block the sum of squared pixel residuals and the Huber-weighted disparity chi2
// (global_ba.py:1004-1005, 338-354), summed in a fixed tree; e (optional) the pixel-residual norm of every row
__global__ __launch_bounds__(256) void k_gba_cost(Dev G, double *part, double *e) {
    __shared__ double sv[256], ss[256];
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    double cv = 0.0, cs = 0.0;
    if (l < G.nl) {
        HostFrame hf;
        host_frame(G, l, hf);
        for (int k = G.lstart[l]; k < G.lstart[l + 1]; k++) {
            const double *T = G.Tcw + 16 * (size_t)G.rkf[k];
            double Pc[3];
            for (int i = 0; i < 3; i++)
                Pc[i] = T[4 * i] * hf.Pw[0] + T[4 * i + 1] * hf.Pw[1] + T[4 * i + 2] * hf.Pw[2] + T[4 * i + 3];
            const double z = fmax(Pc[2], 1e-6);
            const double r0 = G.fx * Pc[0] / z + G.cx - G.ruv[2 * k], r1 = G.fx * Pc[1] / z + G.cy - G.ruv[2 * k + 1];
            cv += r0 * r0 + r1 * r1;
            if (e) e[k] = sqrt(r0 * r0 + r1 * r1);
            const double d = G.rd ? G.rd[k] : 0.0;
            if (d > 1e-6) {
                const double rs = (1.0 / z - d / G.fB) * G.isg_s;
                cs += huber_w(fabs(rs), G.huber_k) * rs * rs;
            }
        }
    }
    sv[threadIdx.x] = cv;
    ss[threadIdx.x] = cs;
    __syncthreads();
    for (int d = 128; d > 0; d >>= 1) {
        if (threadIdx.x < d) {
            sv[threadIdx.x] += sv[threadIdx.x + d];
            ss[threadIdx.x] += ss[threadIdx.x + d];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = sv[0];
        part[2 * blockIdx.x + 1] = ss[0];
    }
}

// out[c] = sum over np partials of part[ncomp * i + c], one workgroup, fixed order
__global__ __launch_bounds__(1024) void k_gba_sum(const double *part, int np, int ncomp, double *out) {
    __shared__ double s[1024];
    for (int c = 0; c < ncomp; c++) {
        double a = 0.0;
        for (int i = threadIdx.x; i < np; i += 1024) a += part[(size_t)ncomp * i + c];
        s[threadIdx.x] = a;
        __syncthreads();
        for (int d = 512; d > 0; d >>= 1) {
            if (threadIdx.x < d) s[threadIdx.x] += s[threadIdx.x + d];
            __syncthreads();
        }
        if (threadIdx.x == 0) out[c] = s[0];
        __syncthreads();
    }
}


// ------------------------------------------------------------------------------------- deflated PCG (DPCG) --
// gba_linsolve.py:160-478 on the block rows: segment smoother M (exact inverses of the segments' diagonal blocks),
// deflation space Z (per segment an orthonormal basis of its keyframes' rigid-motion and scale modes), coarse
// operator E = Z^T H Z (dense, Jacobi-scaled, inverted once per solve), and the CG recursion with its true-residual
// checks, replacement restarts and stopping rules.  Every iteration kernel reads the recursion's scalars from one
// device record and returns at once when the solve has ended or waits for the host (a restart), so a fixed sequence
// of launches can be replayed without reading anything back.
enum { PR_RUN = 0, PR_TOL, PR_MAXITER, PR_BREAKDOWN, PR_DIVERGENCE, PR_STAGNATION, PR_ZERO_RHS, PR_NONFINITE_RHS,
       PR_SETUP };

struct Pcg {
    double rz, alpha, beta, rn, best, tol, nb, truev;
    int32_t it, stall, check, done, paused, reason, matvecs, restarts, copybest, use_best, maxiter, pad;
};

__device__ inline bool pcg_idle(const Pcg *st) { return st->done || st->paused; }

#define GBA_RED 1024  // partial sums per reduction (fixed: deterministic order)

// y = H x (mode 0, partial x.y) or y = b - H x (mode 1, partial |y|^2; only when the recursion asked for a check,
// or always with force), one wavefront per block row, lanes (block group, row)
__global__ __launch_bounds__(256) void k_spmv(int nk, const int32_t *rowptr, const int32_t *col, const double *Hb,
                                              const double *x, double *y, const double *b, int mode, int force,
                                              const Pcg *st, double *part) {
    __shared__ double sp[4][64];
    __shared__ double sdot[4];
    if (!force && (pcg_idle(st) || (mode == 1 && !st->check))) return;
    const int wv = threadIdx.x / XBA_WAVE, lane = threadIdx.x & (XBA_WAVE - 1);
    const int g = lane / 6, i = lane - 6 * g;  // lanes 0..59: 10 groups of 6 rows
    double dacc = 0.0;
    for (int p = blockIdx.x * 4 + wv; p < nk; p += gridDim.x * 4) {
        double acc = 0.0;
        if (lane < 60)
            for (int bi = rowptr[p] + g; bi < rowptr[p + 1]; bi += 10) {
                const double *B = Hb + 36 * (size_t)bi + 6 * i;
                const double *xq = x + 6 * (size_t)col[bi];
                acc += B[0] * xq[0] + B[1] * xq[1] + B[2] * xq[2] + B[3] * xq[3] + B[4] * xq[4] + B[5] * xq[5];
            }
        sp[wv][lane] = acc;
        __builtin_amdgcn_wave_barrier();
        if (lane < 6) {
            double s = 0.0;
            for (int gg = 0; gg < 10; gg++) s += sp[wv][6 * gg + lane];
            const size_t o = 6 * (size_t)p + lane;
            if (mode == 0) {
                y[o] = s;
                dacc += x[o] * s;
            } else {
                const double r = b[o] - s;
                y[o] = r;
                dacc += r * r;
            }
        }
        __builtin_amdgcn_wave_barrier();
    }
    for (int d = 32; d > 0; d >>= 1) dacc += __shfl_xor(dacc, d, XBA_WAVE);
    if (lane == 0) sdot[wv] = dacc;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = ((sdot[0] + sdot[1]) + sdot[2]) + sdot[3];
}

// control stages (one workgroup: the partial sums in a fixed tree, then thread 0)
enum { CT_INIT = 0, CT_START, CT_ALPHA, CT_CHECK, CT_TRUE, CT_BETA, CT_RZ };

__global__ __launch_bounds__(1024) void k_pcg_ctl(Pcg *st, const double *part, int np, int stage) {
    __shared__ double s[1024];
    if (stage != CT_INIT && stage != CT_RZ && pcg_idle(st)) return;
    if (stage == CT_TRUE && !st->check) return;
    double a = 0.0;
    for (int i = threadIdx.x; i < np; i += 1024) a += part[i];
    s[threadIdx.x] = a;
    __syncthreads();
    for (int d = 512; d > 0; d >>= 1) {
        if (threadIdx.x < d) s[threadIdx.x] += s[threadIdx.x + d];
        __syncthreads();
    }
    if (threadIdx.x) return;
    const double v = s[0];
    switch (stage) {
        case CT_INIT: {  // v = |r0|^2 after x0 = coarse part of b
            st->matvecs += 1;
            st->best = sqrt(v);
            if (st->best <= st->tol) {
                st->done = 1;
                st->reason = PR_TOL;
                st->truev = st->best;
            }
            break;
        }
        case CT_START:
            st->copybest = 0;
            if (st->it >= st->maxiter) {
                st->done = 1;
                st->reason = PR_MAXITER;
                st->use_best = 1;
            }
            break;
        case CT_ALPHA:  // v = p.w
            st->matvecs += 1;
            st->it += 1;
            if (!(v > 0.0) || !isfinite(v)) {
                st->done = 1;
                st->reason = PR_BREAKDOWN;
                st->use_best = 1;
            } else {
                st->alpha = st->rz / v;
            }
            break;
        case CT_CHECK:  // v = |r|^2 of the recursion
            st->rn = sqrt(v);
            st->check = (st->rn <= st->tol || st->rn <= 0.1 * st->best || st->it % 50 == 0) ? 1 : 0;
            break;
        case CT_TRUE: {  // v = |b - H x|^2
            st->matvecs += 1;
            const double t = sqrt(v);
            st->truev = t;
            if (t <= st->tol) {
                st->done = 1;
                st->reason = PR_TOL;
                break;
            }
            if (t < st->best) {
                st->best = t;
                st->stall = 0;
                st->copybest = 1;
            } else {
                st->stall += 1;
            }
            if (t > 10.0 * st->best) {
                st->done = 1;
                st->reason = PR_DIVERGENCE;
                st->use_best = 1;
            } else if (st->stall >= 4) {
                st->done = 1;
                st->reason = PR_STAGNATION;
                st->use_best = 1;
            } else if (st->rn <= st->tol || t > 10.0 * st->rn) {
                st->restarts += 1;
                st->paused = 1;  // residual replacement: the host runs the restart sequence
            }
            break;
        }
        case CT_BETA:  // v = r.z
            st->beta = v / st->rz;
            st->rz = v;
            break;
        case CT_RZ:  // v = r.z (start or restart of the recursion)
            st->rz = v;
            break;
    }
}

// x += alpha p, r -= alpha w; partial |r|^2
__global__ __launch_bounds__(256) void k_axpy(int n, const Pcg *st, double *x, double *r, const double *p,
                                              const double *w, double *part) {
    __shared__ double s[256];
    if (pcg_idle(st)) return;
    const double al = st->alpha;
    double a = 0.0;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
        x[i] += al * p[i];
        const double v = r[i] - al * w[i];
        r[i] = v;
        a += v * v;
    }
    s[threadIdx.x] = a;
    __syncthreads();
    for (int d = 128; d > 0; d >>= 1) {
        if (threadIdx.x < d) s[threadIdx.x] += s[threadIdx.x + d];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}

// x_best = x after a true-residual check that found a new smallest residual
__global__ void k_copybest(int n, const Pcg *st, const double *x, double *xb) {
    if (!st->copybest) return;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) xb[i] = x[i];
}

// z = M^-1 r, segment by segment (one workgroup per segment, Binv row-major (6 m_s)^2); partial r.z per segment
__global__ __launch_bounds__(256) void k_minv(const int32_t *segk, const int64_t *boff, const double *Binv,
                                              const double *r, double *z, const Pcg *st, int force, double *part) {
    __shared__ double sr[6 * GBA_SEG_MAX];
    __shared__ double s[256];
    if (!force && pcg_idle(st)) return;
    const int sg = blockIdx.x;
    const int k0 = segk[sg], n = 6 * (segk[sg + 1] - k0);
    const double *rr = r + 6 * (size_t)k0;
    for (int i = threadIdx.x; i < n; i += 256) sr[i] = rr[i];
    __syncthreads();
    const double *B = Binv + boff[sg];
    double a = 0.0;
    for (int i = threadIdx.x; i < n; i += 256) {
        double acc = 0.0;
        for (int j = 0; j < n; j++) acc += B[(size_t)i * n + j] * sr[j];
        z[6 * (size_t)k0 + i] = acc;
        a += sr[i] * acc;
    }
    s[threadIdx.x] = a;
    __syncthreads();
    for (int d = 128; d > 0; d >>= 1) {
        if (threadIdx.x < d) s[threadIdx.x] += s[threadIdx.x + d];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[sg] = s[0];
}

// y = Z^T v (coarse vector; one wavefront per segment)
__global__ void k_zt(const int32_t *segk, const int32_t *coff, const double *Zg, const double *v, double *y,
                     const Pcg *st, int force) {
    if (!force && pcg_idle(st)) return;
    const int sg = blockIdx.x, lane = threadIdx.x;
    const int k0 = segk[sg], n = 6 * (segk[sg + 1] - k0), c0 = coff[sg], rs = coff[sg + 1] - c0;
    double acc[7] = {0, 0, 0, 0, 0, 0, 0};
    for (int i = lane; i < n; i += XBA_WAVE) {
        const size_t gi = 6 * (size_t)k0 + i;
        const double vi = v[gi];
        for (int j = 0; j < 7; j++) acc[j] += Zg[7 * gi + j] * vi;
    }
    for (int j = 0; j < 7; j++)
        for (int d = 32; d > 0; d >>= 1) acc[j] += __shfl_xor(acc[j], d, XBA_WAVE);
    if (lane < rs) {
        double a = 0.0;
        for (int j = 0; j < 7; j++) a = (j == lane) ? acc[j] : a;
        y[c0 + lane] = a;
    }
}

// c = Einv y (dense nc x nc, one wavefront per row)
__global__ __launch_bounds__(256) void k_einv(int nc, int ld, const double *Einv, const double *y, double *c,
                                              const Pcg *st, int force) {
    if (!force && pcg_idle(st)) return;
    const int row = blockIdx.x * 4 + threadIdx.x / XBA_WAVE, lane = threadIdx.x & (XBA_WAVE - 1);
    if (row >= nc) return;
    double a = 0.0;
    for (int j = lane; j < nc; j += XBA_WAVE) a += Einv[(size_t)row * ld + j] * y[j];
    for (int d = 32; d > 0; d >>= 1) a += __shfl_xor(a, d, XBA_WAVE);
    if (lane == 0) c[row] = a;
}

// per segment: mode 0  p = (z - Z c) + beta p   (beta from the record; 0 with nobeta)
//              mode 1  x = Z c                  (the coarse part)
//              mode 2  x += Z c
__global__ void k_zapply(const int32_t *segk, const int32_t *coff, const double *Zg, const double *c, const double *z,
                         double *p, const Pcg *st, int mode, int nobeta, int force) {
    if (!force && pcg_idle(st)) return;
    const int sg = blockIdx.x;
    const int k0 = segk[sg], n = 6 * (segk[sg + 1] - k0), c0 = coff[sg], rs = coff[sg + 1] - c0;
    const double be = st->beta;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        const size_t gi = 6 * (size_t)k0 + i;
        double zc = 0.0;
        for (int j = 0; j < rs; j++) zc += Zg[7 * gi + j] * c[c0 + j];
        if (mode == 0) p[gi] = nobeta ? z[gi] - zc : (z[gi] - zc) + be * p[gi];  // no read of p on a (re)start
        else if (mode == 1) p[gi] = zc;
        else p[gi] += zc;
    }
}

// |v|^2 partials (any state)
__global__ __launch_bounds__(256) void k_norm2(int n, const double *v, double *part) {
    __shared__ double s[256];
    double a = 0.0;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) a += v[i] * v[i];
    s[threadIdx.x] = a;
    __syncthreads();
    for (int d = 128; d > 0; d >>= 1) {
        if (threadIdx.x < d) s[threadIdx.x] += s[threadIdx.x + d];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}

// ------------------------------------------------------------------------------------------- DPCG set-up --
// Scalar sweep of a symmetric positive definite n x n matrix (A <- -A^-1) with the matrix in registers: thread t
// owns column j = t % n of the rows i0 + R u (R = nt / n row groups, i0 = t / n); step k publishes column k through
// LDS (ck) and updates every entry in place.  Loads from and stores back to the LDS copy A (row-major).  Returns
// false when a pivot is not positive (the matrix is not numerically positive definite).
template <int MAXR, int RMAX>
__device__ inline bool reg_sweep(double *A, int n, double *ck) {
    // R = min(nt / n, RMAX) row groups; ck holds RMAX * MAXR entries, the rows past n padding (never stored)
    const int tid = threadIdx.x, R = min((int)blockDim.x / n, RMAX);
    const int j = tid % n, i0 = tid / n;
    const bool act = i0 < R;
    double v[MAXR];
#pragma unroll
    for (int u = 0; u < MAXR; u++) {
        const int i = i0 + R * u;
        v[u] = (act && i < n) ? A[i * n + j] : 0.0;
    }
    bool ok = true;
    for (int k = 0; k < n; k++) {
        if (act && j == k)
#pragma unroll
            for (int u = 0; u < MAXR; u++) ck[i0 + R * u] = v[u];
        __syncthreads();
        const double piv = ck[k];
        if (!(piv > 0.0) || !isfinite(piv)) {
            ok = false;
            break;
        }
        const double ip = 1.0 / piv;
        const bool colk = j == k;
        const double f = colk ? ip : ck[j] * ip;
#pragma unroll
        for (int u = 0; u < MAXR; u++) {
            const int i = i0 + R * u;
            const double c = ck[i];
            const double upd = colk ? c * ip : v[u] - c * f;
            v[u] = (i == k) ? (colk ? -ip : f) : upd;
        }
        __syncthreads();
    }
#pragma unroll
    for (int u = 0; u < MAXR; u++) {
        const int i = i0 + R * u;
        if (act && i < n) A[i * n + j] = v[u];
    }
    __syncthreads();
    return ok;
}

// one workgroup per segment: its diagonal block of H (symmetrised), inverted in LDS; fail |= 1 on a block that is
// not positive definite
__global__ __launch_bounds__(1024) void k_seg_inv(const int32_t *segk, const int64_t *boff, const int32_t *rowptr,
                                                  const int32_t *col, const int32_t *brow, const double *Hb,
                                                  double *Binv, int32_t *fail) {
    extern __shared__ double sB[];
    __shared__ double ck[7 * 20];  // 7 row groups x 20 rows >= 6 GBA_SEG_MAX
    const int sg = blockIdx.x, tid = threadIdx.x, nt = blockDim.x;
    const int k0 = segk[sg], k1 = segk[sg + 1], n = 6 * (k1 - k0);
    for (int e = tid; e < n * n; e += nt) sB[e] = 0.0;
    __syncthreads();
    // the blocks of the segment's rows with a column inside the segment, 36 threads per block
    const int bi0 = rowptr[k0], nbk = rowptr[k1] - bi0;
    for (int e = tid; e < 36 * nbk; e += nt) {
        const int bi = bi0 + e / 36, w = e - 36 * (e / 36);
        const int q = col[bi];
        if (q < k0 || q >= k1) continue;
        const int p = brow[bi];
        sB[(size_t)(6 * (p - k0) + w / 6) * n + 6 * (q - k0) + w % 6] = Hb[36 * (size_t)bi + w];
    }
    __syncthreads();
    for (int e = tid; e < n * n; e += nt) {
        const int i = e / n, j = e - i * n;
        if (i < j) {
            const double v = 0.5 * (sB[e] + sB[(size_t)j * n + i]);
            sB[e] = v;
            sB[(size_t)j * n + i] = v;
        }
    }
    __syncthreads();
    // n <= 6 GBA_SEG_MAX = 138 columns: at least 7 row groups of 1024 threads, 20 rows each
    const bool ok = reg_sweep<20, 7>(sB, n, ck);
    if (!ok) {
        if (tid == 0) atomicOr(fail, 1);
        return;
    }
    double *out = Binv + boff[sg];
    for (int e = tid; e < n * n; e += nt) out[e] = -sB[e];
}

// one wavefront per segment: the mode vectors of its keyframes (gba_linsolve.py:249-270, no velocity; rows of fixed
// keyframes zero) reduced to an orthonormal basis by modified Gram-Schmidt with column pivoting and a second
// orthogonalisation pass; columns whose remaining norm is at or below |first| * max(rows, 7) * eps are dropped
// (the rank rule of gba_linsolve.py:294).  Zg rows of the segment (7 columns, unused ones zero), rcount[s] = rank.
__global__ void k_seg_z(const int32_t *segk, const double *Tcw, const uint8_t *free_mask, double *Zg,
                        int32_t *rcount) {
    const int sg = blockIdx.x, lane = threadIdx.x;
    const int k0 = segk[sg], k1 = segk[sg + 1], n = 6 * (k1 - k0);
    // per lane up to 3 rows (n <= 138 < 192)
    double M[3][7];
    int nzr = 0;
    for (int q = 0; q < 3; q++) {
        const int i = lane + XBA_WAVE * q;
        for (int j = 0; j < 7; j++) M[q][j] = 0.0;
        if (i >= n) continue;
        const int k = k0 + i / 6, rr = i % 6;
        if (!free_mask[k]) continue;
        const double *T = Tcw + 16 * (size_t)k;
        const double R[9] = {T[0], T[1], T[2], T[4], T[5], T[6], T[8], T[9], T[10]};
        const double t[3] = {T[3], T[7], T[11]};
        if (rr < 3) {
            // [-R | -[t]x R | t]
            for (int j = 0; j < 3; j++) M[q][j] = -R[3 * rr + j];
            const double tx[9] = {0, -t[2], t[1], t[2], 0, -t[0], -t[1], t[0], 0};
            for (int j = 0; j < 3; j++)
                M[q][3 + j] = -(tx[3 * rr] * R[j] + tx[3 * rr + 1] * R[3 + j] + tx[3 * rr + 2] * R[6 + j]);
            M[q][6] = t[rr];
        } else {
            for (int j = 0; j < 3; j++) M[q][3 + j] = -R[3 * (rr - 3) + j];
        }
        nzr++;
    }
    for (int d = 32; d > 0; d >>= 1) nzr += __shfl_xor(nzr, d, XBA_WAVE);
    auto dot = [&](int a, int b) {
        double s = 0.0;
        for (int q = 0; q < 3; q++) s += M[q][a] * M[q][b];
        for (int d = 32; d > 0; d >>= 1) s += __shfl_xor(s, d, XBA_WAVE);
        return s;
    };
    bool used[7] = {false, false, false, false, false, false, false};
    int order[7];
    int r = 0;
    double first = 0.0;
    const double thr_f = (double)(nzr > 7 ? nzr : 7) * 2.220446049250313e-16;
    for (int step = 0; step < 7; step++) {
        int best = -1;
        double bn = -1.0;
        for (int j = 0; j < 7; j++) {
            if (used[j]) continue;
            const double v = dot(j, j);
            if (v > bn) {
                bn = v;
                best = j;
            }
        }
        const double nrm = sqrt(fmax(bn, 0.0));
        if (step == 0) first = nrm;
        if (best < 0 || !(nrm > first * thr_f) || nzr == 0) break;
        used[best] = true;
        for (int q = 0; q < 3; q++) M[q][best] /= nrm;
        for (int pass = 0; pass < 2; pass++)
            for (int j = 0; j < 7; j++) {
                if (used[j]) continue;
                const double c = dot(best, j);
                for (int q = 0; q < 3; q++) M[q][j] -= c * M[q][best];
            }
        order[r++] = best;
    }
    for (int q = 0; q < 3; q++) {
        const int i = lane + XBA_WAVE * q;
        if (i >= n) continue;
        double *z = Zg + 7 * (6 * (size_t)k0 + i);
        for (int j = 0; j < 7; j++) {
            double v = 0.0;
            if (j < r)
                for (int u = 0; u < 7; u++) v = (u == order[j]) ? M[q][u] : v;
            z[j] = v;
        }
    }
    if (lane == 0) rcount[sg] = r;
}

// E[coff_s + a][coff_t + b] += Z_p^T H_pq Z_q: one wavefront per block row p, a lane per block; the lanes of one
// column segment t (contiguous: segments are keyframe ranges and columns ascend) are summed by wavefront
// reductions and flushed with one 7 x 7 set of global atomics per (p, t)
__global__ __launch_bounds__(256) void k_coarse_E(int nk, const int32_t *rowptr, const int32_t *col, const double *Hb,
                                                  const int32_t *seg_of, const int32_t *coff, const double *Zg,
                                                  double *E, int ldE) {
    const int p = blockIdx.x * 4 + threadIdx.x / XBA_WAVE, lane = threadIdx.x & (XBA_WAVE - 1);
    if (p >= nk) return;
    const int s = seg_of[p], rs = coff[s + 1] - coff[s];
    if (!rs) return;
    const double *Zp = Zg + 42 * (size_t)p;
    for (int base = rowptr[p]; base < rowptr[p + 1]; base += XBA_WAVE) {
        const int bi = base + lane;
        const bool on = bi < rowptr[p + 1];
        const int q = on ? col[bi] : 0, t = on ? seg_of[q] : -1;
        double C[7][7];
#pragma unroll
        for (int a = 0; a < 7; a++)
#pragma unroll
            for (int b = 0; b < 7; b++) C[a][b] = 0.0;
        if (on) {
            const double *B = Hb + 36 * (size_t)bi;
            const double *Zq = Zg + 42 * (size_t)q;
            double HZ[6][7];
            for (int i = 0; i < 6; i++)
                for (int b = 0; b < 7; b++) {
                    double acc = 0.0;
                    for (int j = 0; j < 6; j++) acc += B[6 * i + j] * Zq[7 * j + b];
                    HZ[i][b] = acc;
                }
            for (int a = 0; a < 7; a++)
                for (int b = 0; b < 7; b++) {
                    double v = 0.0;
                    for (int i = 0; i < 6; i++) v += Zp[7 * i + a] * HZ[i][b];
                    C[a][b] = v;
                }
        }
        // the distinct segments among the lanes, ascending (wavefront minima)
        auto wmin = [&](int v) {
            for (int d = 32; d > 0; d >>= 1) v = min(v, __shfl_xor(v, d, XBA_WAVE));
            return v;
        };
        for (int tcur = wmin(on ? t : INT_MAX); tcur != INT_MAX; tcur = wmin(on && t > tcur ? t : INT_MAX)) {
            const int rt = coff[tcur + 1] - coff[tcur];
#pragma unroll
            for (int a = 0; a < 7; a++)
#pragma unroll
                for (int b = 0; b < 7; b++) {
                    if (a >= rs || b >= rt) continue;
                    double v = t == tcur ? C[a][b] : 0.0;
                    for (int d = 32; d > 0; d >>= 1) v += __shfl_xor(v, d, XBA_WAVE);
                    if (lane == 0 && v != 0.0) atomicAdd(E + (size_t)(coff[s] + a) * ldE + coff[tcur] + b, v);
                }
        }
    }
}

// d = sqrt(max(diag E, 1e-300)) (gba_linsolve.py:318-321)
__global__ void k_coarse_diag(int nc, int ldE, const double *E, double *dsc) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nc) dsc[i] = sqrt(fmax(E[(size_t)i * ldE + i], 1e-300));
}

// E <- (E + E^T) / 2 / (d d^T); identity on the padding [nc, ldE)
__global__ void k_coarse_scale(int nc, int ldE, double *E, const double *dsc) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (int64_t)ldE * ldE) return;
    const int i = (int)(e / ldE), j = (int)(e - (int64_t)i * ldE);
    if (i >= nc || j >= nc) {
        E[e] = (i == j) ? 1.0 : 0.0;
        return;
    }
    if (i > j) return;  // the pair (i, j), (j, i) belongs to the thread of i <= j
    const double v = 0.5 * (E[e] + E[(size_t)j * ldE + i]) / (dsc[i] * dsc[j]);
    E[e] = v;
    E[(size_t)j * ldE + i] = v;
}

// Blocked sweep of the scaled coarse matrix over 64 x 64 tiles (A <- -A^-1 after every tile step k):
//   P = A_kk^-1;  S_i = A_ik P (i != k);  A_ij -= S_i A_kj (i, j != k);  A_ik = S_i, A_ki = S_i^T, A_kk = -P
// on the lower triangle of tiles (below)
#define SW_T 64
__device__ inline void tile_load(double *dst, const double *A, int ld, int ti, int tj) {
    for (int e = threadIdx.x; e < SW_T * SW_T; e += blockDim.x) {
        const int i = e / SW_T, j = e - SW_T * (e / SW_T);
        dst[e] = A[(size_t)(SW_T * ti + i) * ld + SW_T * tj + j];
    }
}

__global__ __launch_bounds__(1024) void k_sw_pivot(double *A, int ld, int k, double *P, int32_t *fail) {
    __shared__ double sT[SW_T * SW_T];
    __shared__ double ck[SW_T];  // R * MAXR = 16 * 4 = 64 rows
    tile_load(sT, A, ld, k, k);
    __syncthreads();
    for (int e = threadIdx.x; e < SW_T * SW_T; e += blockDim.x) {  // symmetrise the pivot tile
        const int i = e / SW_T, j = e - SW_T * (e / SW_T);
        if (i < j) {
            const double v = 0.5 * (sT[e] + sT[j * SW_T + i]);
            sT[e] = v;
            sT[j * SW_T + i] = v;
        }
    }
    __syncthreads();
    if (!reg_sweep<SW_T * SW_T / 1024, 16>(sT, SW_T, ck)) {  // 16 row groups of 64 columns, 4 rows each
        if (threadIdx.x == 0) atomicOr(fail, 1);
        return;
    }
    for (int e = threadIdx.x; e < SW_T * SW_T; e += blockDim.x) P[e] = -sT[e];
}

// C = X Y (64 x 64 tiles in LDS, 256 threads x 4 x 4 outputs)
__device__ inline void tile_mm(const double *X, const double *Y, double acc[4][4]) {
    const int ti = threadIdx.x / 16, tj = threadIdx.x - 16 * (threadIdx.x / 16);
    for (int a = 0; a < 4; a++)
        for (int b = 0; b < 4; b++) acc[a][b] = 0.0;
    for (int m = 0; m < SW_T; m++) {
        double xv[4], yv[4];
        for (int a = 0; a < 4; a++) xv[a] = X[(4 * ti + a) * SW_T + m];
        for (int b = 0; b < 4; b++) yv[b] = Y[m * SW_T + 4 * tj + b];
        for (int a = 0; a < 4; a++)
            for (int b = 0; b < 4; b++) acc[a][b] += xv[a] * yv[b];
    }
}

// The sweep keeps the lower triangle of tiles only (the matrix stays symmetric through every step): an upper tile
// (i, j), i < j, is read as the transpose of the lower tile (j, i).
__device__ inline void tile_load_sym(double *dst, const double *A, int ld, int ti, int tj) {
    if (ti >= tj) {
        tile_load(dst, A, ld, ti, tj);
        return;
    }
    for (int e = threadIdx.x; e < SW_T * SW_T; e += blockDim.x) {
        const int i = e / SW_T, j = e - SW_T * (e / SW_T);
        dst[e] = A[(size_t)(SW_T * tj + j) * ld + SW_T * ti + i];
    }
}

__global__ __launch_bounds__(256) void k_sw_strip(const double *A, int ld, int k, const double *P, double *S,
                                                  const int32_t *fail) {
    __shared__ double sX[SW_T * SW_T], sY[SW_T * SW_T];
    const int i = blockIdx.x;
    if (*fail || i == k) return;
    tile_load_sym(sX, A, ld, i, k);
    for (int e = threadIdx.x; e < SW_T * SW_T; e += blockDim.x) sY[e] = P[e];
    __syncthreads();
    double acc[4][4];
    tile_mm(sX, sY, acc);
    const int ti = threadIdx.x / 16, tj = threadIdx.x - 16 * (threadIdx.x / 16);
    for (int a = 0; a < 4; a++)
        for (int b = 0; b < 4; b++) S[(size_t)(SW_T * i + 4 * ti + a) * SW_T + 4 * tj + b] = acc[a][b];
}

// the lower tiles (i >= j, neither k): A_ij -= S_i A_kj; one workgroup per tile, x = i (i + 1) / 2 + j
__global__ __launch_bounds__(256) void k_sw_update(double *A, int ld, int k, const double *S, const int32_t *fail) {
    __shared__ double sX[SW_T * SW_T], sY[SW_T * SW_T];
    const int x = blockIdx.x;
    int i = (int)((sqrtf(8.0f * (float)x + 1.0f) - 1.0f) * 0.5f);
    while ((i + 1) * (i + 2) / 2 <= x) i++;
    while (i * (i + 1) / 2 > x) i--;
    const int j = x - i * (i + 1) / 2;
    if (*fail || i == k || j == k) return;
    for (int e = threadIdx.x; e < SW_T * SW_T; e += blockDim.x) sX[e] = S[(size_t)SW_T * SW_T * i + e];
    tile_load_sym(sY, A, ld, k, j);
    __syncthreads();
    double acc[4][4];
    tile_mm(sX, sY, acc);
    const int ti = threadIdx.x / 16, tj = threadIdx.x - 16 * (threadIdx.x / 16);
    for (int a = 0; a < 4; a++)
        for (int b = 0; b < 4; b++) A[(size_t)(SW_T * i + 4 * ti + a) * ld + SW_T * j + 4 * tj + b] -= acc[a][b];
}

// column k of the lower triangle: A_ik = S_i below the diagonal tile, A_ki = S_i^T left of it; A_kk = -P
__global__ void k_sw_finish(double *A, int ld, int k, const double *P, const double *S, const int32_t *fail) {
    const int i = blockIdx.x;
    if (*fail) return;
    for (int e = threadIdx.x; e < SW_T * SW_T; e += blockDim.x) {
        const int a = e / SW_T, b = e - SW_T * (e / SW_T);
        if (i == k) {
            A[(size_t)(SW_T * k + a) * ld + SW_T * k + b] = -P[e];
        } else {
            const double v = S[(size_t)(SW_T * i + a) * SW_T + b];
            if (i > k) A[(size_t)(SW_T * i + a) * ld + SW_T * k + b] = v;
            else A[(size_t)(SW_T * k + b) * ld + SW_T * i + a] = v;
        }
    }
}

// Einv = -A / (d d^T) on [0, nc)^2 (A the swept scaled matrix): the lower tiles in place, then (k_coarse_mirror)
// the upper tiles from them
__global__ void k_coarse_unscale(int nc, int ld, double *A, const double *dsc) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (int64_t)nc * ld) return;
    const int i = (int)(e / ld), j = (int)(e - (int64_t)i * ld);
    if (j >= nc || i / SW_T < j / SW_T) return;
    A[e] = -A[e] / (dsc[i] * dsc[j]);
}

__global__ void k_coarse_mirror(int nc, int ld, double *A) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (int64_t)nc * ld) return;
    const int i = (int)(e / ld), j = (int)(e - (int64_t)i * ld);
    if (j >= nc || i / SW_T >= j / SW_T) return;
    A[e] = A[(size_t)j * ld + i];
}

__global__ void k_gba_brow(int nk, const int32_t *rowptr, int32_t *brow) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= nk) return;
    for (int i = rowptr[p]; i < rowptr[p + 1]; i++) brow[i] = p;
}

// ------------------------------------------------------------------------------- order statistics (device) --
// The k-th smallest of n keys (k_to_keys) by an 8-pass radix select over their bytes.  sel[0] = prefix, sel[1] =
// remaining rank, sel[2] = result.
__global__ void k_sel_hist(int n, const uint64_t *v, int pass, const uint64_t *sel, uint32_t *hist) {
    __shared__ uint32_t sh[256];
    for (int i = threadIdx.x; i < 256; i += blockDim.x) sh[i] = 0;
    __syncthreads();
    const int shift = 56 - 8 * pass;
    const uint64_t pre = sel[0];
    const uint64_t mask = pass == 0 ? 0ull : (~0ull << (64 - 8 * pass));
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const uint64_t key = v[i];
        if ((key & mask) == (pre & mask)) atomicAdd(&sh[(key >> shift) & 255], 1u);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < 256; i += blockDim.x)
        if (sh[i]) atomicAdd(&hist[i], sh[i]);
}

__global__ void k_sel_pick(int pass, uint64_t *sel, uint32_t *hist) {
    if (threadIdx.x) return;
    uint64_t r = sel[1];
    int d = 0;
    for (; d < 255; d++) {
        if (r < hist[d]) break;
        r -= hist[d];
    }
    sel[0] |= (uint64_t)d << (56 - 8 * pass);
    sel[1] = r;
    for (int i = 0; i < 256; i++) hist[i] = 0;
    if (pass == 7) sel[2] = sel[0];
}

// |e - m| for the median absolute deviation; rows with keep = e <= thr
__global__ void k_absdev(int n, const double *e, const double *m, double *out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = fabs(e[i] - *m);
}

// (rho - rho0) invsig of the valid priors and its magnitude (+inf for the others), and their count
__global__ void k_prior_res(Dev G, double *sgn, double *mag, int32_t *cnt) {
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= G.nl) return;
    if (G.pval[l] > 0.5) {
        const double v = (G.rho[l] - G.prho0[l]) * G.pisg[l];
        sgn[l] = v;
        mag[l] = fabs(v);
        atomicAdd(cnt, 1);
    } else {
        sgn[l] = mag[l] = __longlong_as_double(0x7ff0000000000000ll);
    }
}

// order-preserving 64-bit keys of doubles (negative values bit-inverted, the others with the sign bit set)
__global__ void k_to_keys(int n, const double *v, uint64_t *key) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint64_t b = (uint64_t)__double_as_longlong(v[i]);
    key[i] = (b >> 63) ? ~b : (b | (1ull << 63));
}

// cnt += the entries of v above thr
__global__ void k_count_above(int n, const double *v, double thr, int32_t *cnt) {
    int c = 0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) c += v[i] > thr;
    for (int d = 32; d > 0; d >>= 1) c += __shfl_xor(c, d, XBA_WAVE);
    if ((threadIdx.x & (XBA_WAVE - 1)) == 0 && c) atomicAdd(cnt, c);
}

__global__ void k_scale(int n, const double *a, double f, double *out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = a[i] * f;
}

__global__ void k_keep(int n, const double *e, double thr, uint8_t *keep, int32_t *nkeep) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int k = e[i] <= thr;
    keep[i] = (uint8_t)k;
    if (k) atomicAdd(nkeep, 1);
}

}  // namespace gba

// ================================================================================================ host driver ==
namespace gba {

struct HipFail {
    hipError_t e;
};
struct RangeFail {};
#define GCK(x)                                                                           \
    do {                                                                                 \
        hipError_t e_ = (x);                                                             \
        if (e_ != hipSuccess) {                                                          \
            fprintf(stderr, "xba gba: %s failed: %s\n", #x, hipGetErrorString(e_));      \
            throw HipFail{e_};                                                           \
        }                                                                                \
    } while (0)

struct Arena {
    std::vector<void *> ptrs;
    template <typename T>
    T *get(size_t n) {
        void *p = nullptr;
        GCK(hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T)));
        ptrs.push_back(p);
        return (T *)p;
    }
    void release(void *p) {
        for (auto &q : ptrs)
            if (q == p) {
                (void)hipFree(q);
                q = nullptr;
            }
    }
    ~Arena() {
        for (void *p : ptrs)
            if (p) (void)hipFree(p);
    }
};

inline int nblk(int64_t n, int bs) { return (int)std::max<int64_t>(1, (n + bs - 1) / bs); }
inline double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

struct Pass {  // the rows of one LM pass and the structures derived from them
    int M = 0;
    int32_t *lstart = nullptr, *rkf = nullptr, *rlm = nullptr, *vstart = nullptr, *visit = nullptr;
    double *ruv = nullptr, *rd = nullptr;
    int32_t *rowptr = nullptr, *col = nullptr, *brow = nullptr;
    int64_t nnzb = 0;
    int4 *chunks = nullptr;
    int nchunks = 0;
    size_t lds = 0;
    double *Hb = nullptr, *bp = nullptr;
    const double *pisg = nullptr;  // the pass's prior whitening (re-weighted in the second pass)
};

struct PcgInfo {
    int reason = PR_SETUP, iters = 0, matvecs = 0, restarts = 0;
    double relres = NAN;
    bool converged = false;
};

struct PassStats {
    double chi0v = 0, chi0s = 0, chibest = 0;
    int acc = 0, best = 0, iters_run = 0;
};

class Solver {
  public:
    hipStream_t s;
    const xba_gba_problem &q;
    const xba_gba_options &o;
    Arena A;
    int nk, nl, n;
    double *T[3], *RH[3], *MB[3];  // state slots; cur / trial / best index into them
    int cur = 0, tri = 1, bst = 2;
    double *Hlli, *bl, *whs, *rrec = nullptr;
    size_t rrec_n = 0;
    // segments and the deflated PCG
    int m = 0, nseg = 0, ldEmax = 0, seg_max = 0;
    std::vector<int32_t> segk_h;
    int32_t *segk, *seg_of, *rcount;
    int64_t *boff;
    double *Binv, *Zg, *E, *dsc, *Pt, *St;
    double *x, *r, *z, *pv, *w, *rt, *xb, *hz, *yc, *cc;
    double *part, *csum, *cpart = nullptr, *e = nullptr, *tmpv = nullptr, *tmpv2 = nullptr;
    uint64_t *keys = nullptr, *sel;
    uint32_t *hist;
    Pcg *st;
    int32_t *flag;
    // totals and the phase timings (HIP events, read at the end)
    int solves = 0, pcg_iters = 0, unconv = 0;
    std::vector<hipEvent_t> ev;
    std::vector<std::pair<int, int>> asm_ev, setup_ev, pcg_ev;  // event index pairs
    int mark() {
        if (!o.profile) return -1;
        hipEvent_t e;
        GCK(hipEventCreate(&e));
        GCK(hipEventRecord(e, s));
        ev.push_back(e);
        return (int)ev.size() - 1;
    }
    double mean_ms(const std::vector<std::pair<int, int>> &v) {
        if (v.empty()) return NAN;
        double t = 0.0;
        for (auto &pr : v) {
            float ms = 0.0f;
            GCK(hipEventElapsedTime(&ms, ev[pr.first], ev[pr.second]));
            t += ms;
        }
        return t / v.size();
    }
    ~Solver() {
        for (auto e : ev) (void)hipEventDestroy(e);
    }
    std::vector<double> histo;

    Solver(const xba_gba_problem &q_, const xba_gba_options &o_, hipStream_t s_) : s(s_), q(q_), o(o_) {
        nk = q.nk;
        nl = q.nl;
        n = 6 * nk;
        for (int i = 0; i < 3; i++) {
            T[i] = A.get<double>(16 * (size_t)nk);
            RH[i] = A.get<double>(nl);
            MB[i] = A.get<double>(3 * (size_t)nl);
        }
        xba::k_np_inv_poses<<<nblk(nk, 128), 128, 0, s>>>(nk, q.poses, T[0]);
        GCK(hipMemcpyAsync(RH[0], q.rho, sizeof(double) * nl, hipMemcpyDeviceToDevice, s));
        GCK(hipMemcpyAsync(MB[0], q.mbear, sizeof(double) * 3 * nl, hipMemcpyDeviceToDevice, s));
        Hlli = A.get<double>(6 * (size_t)nl);
        whs = A.get<double>(18 * (size_t)nl);
        bl = A.get<double>(3 * (size_t)nl);
        // segments (gba_linsolve.py:147-154), capped at GBA_SEG_MAX keyframes
        if (o.seg_len > 0) m = o.seg_len;
        else m = std::max(2, (int)std::nearbyint(std::cbrt(2.0 * 49.0 * nk / 36.0)));  // Python round: half to even
        m = std::min(m, GBA_SEG_MAX);
        // gba_linsolve.py:152-154: round(nk / m) segments at rounded linspace bounds; more segments while the
        // longest one exceeds GBA_SEG_MAX keyframes
        std::vector<int32_t> b;
        int maxlen = 0;
        for (int ns = std::max(1, (int)std::nearbyint((double)nk / m));; ns++) {
            b.clear();
            for (int i = 0; i <= ns; i++) {
                const int v = (int)std::nearbyint((double)nk * i / ns);  // numpy rounds half to even
                if (b.empty() || v != b.back()) b.push_back(v);
            }
            maxlen = 0;
            for (size_t i = 0; i + 1 < b.size(); i++) maxlen = std::max(maxlen, b[i + 1] - b[i]);
            if (maxlen <= GBA_SEG_MAX) break;
        }
        segk_h = b;
        nseg = (int)b.size() - 1;
        seg_max = maxlen;
        std::vector<int32_t> so(nk);
        std::vector<int64_t> bo(nseg + 1, 0);
        for (int i = 0; i < nseg; i++) {
            for (int k = b[i]; k < b[i + 1]; k++) so[k] = i;
            const int64_t d = 6 * (int64_t)(b[i + 1] - b[i]);
            bo[i + 1] = bo[i] + d * d;
        }
        segk = A.get<int32_t>(nseg + 1);
        seg_of = A.get<int32_t>(nk);
        boff = A.get<int64_t>(nseg + 1);
        GCK(hipMemcpyAsync(segk, b.data(), sizeof(int32_t) * (nseg + 1), hipMemcpyHostToDevice, s));
        GCK(hipMemcpyAsync(seg_of, so.data(), sizeof(int32_t) * nk, hipMemcpyHostToDevice, s));
        GCK(hipMemcpyAsync(boff, bo.data(), sizeof(int64_t) * (nseg + 1), hipMemcpyHostToDevice, s));
        Binv = A.get<double>(bo[nseg]);
        Zg = A.get<double>(7 * (size_t)n);
        rcount = A.get<int32_t>(nseg + 1);
        ldEmax = ((7 * nseg + SW_T - 1) / SW_T) * SW_T;
        E = A.get<double>((size_t)ldEmax * ldEmax);
        dsc = A.get<double>(ldEmax);
        Pt = A.get<double>(SW_T * SW_T);
        St = A.get<double>((size_t)ldEmax * SW_T);
        for (double **v : {&x, &r, &z, &pv, &w, &rt, &xb, &hz}) *v = A.get<double>(n);
        yc = A.get<double>(ldEmax);
        cc = A.get<double>(ldEmax);
        part = A.get<double>(std::max(GBA_RED, nseg));
        csum = A.get<double>(4);
        st = A.get<Pcg>(1);
        flag = A.get<int32_t>(8);
        sel = A.get<uint64_t>(4);
        hist = A.get<uint32_t>(256);
        GCK(hipMemsetAsync(hist, 0, 256 * sizeof(uint32_t), s));
        GCK(hipFuncSetAttribute((const void *)k_seg_inv, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)(sizeof(double) * 36 * GBA_SEG_MAX * GBA_SEG_MAX)));
        GCK(hipFuncSetAttribute((const void *)k_gba_assemble, hipFuncAttributeMaxDynamicSharedMemorySize,
                                GBA_LDS_CAP));
    }

    Dev dev(const Pass &P, int slot, double lam) const {
        Dev G{};
        G.nk = nk;
        G.nl = nl;
        G.N = P.M;
        G.Tcw = T[slot];
        G.rho = RH[slot];
        G.mb = MB[slot];
        G.host = q.host;
        G.lstart = P.lstart;
        G.rkf = P.rkf;
        G.rlm = P.rlm;
        G.ruv = P.ruv;
        G.rd = P.rd;
        G.vtw = q.vt_w;
        G.free_mask = q.free_mask;
        if (q.pr_m0) {
            G.pm0 = q.pr_m0;
            G.prho0 = q.pr_rho0;
            G.pwb = q.pr_wbear;
            G.pisg = P.pisg;
            G.pval = q.pr_valid;
        }
        G.fx = q.fx;
        G.cx = q.cx;
        G.cy = q.cy;
        G.fB = q.fB;
        G.isg_s = 1.0 / std::max(q.rho_sigma > 0 ? q.rho_sigma : 1e-9, 1e-12);
        G.huber_k = 1.345;  // core/derive.py:47
        G.lam = lam;
        G.vstart = P.vstart;
        G.visit = P.visit;
        G.rowptr = P.rowptr;
        G.col = P.col;
        G.Hb = P.Hb;
        G.bp = P.bp;
        G.Hlli = Hlli;
        G.bl = bl;
        return G;
    }

    template <typename V>
    V read1(const V *d) {
        V h;
        GCK(hipMemcpyAsync(&h, d, sizeof(V), hipMemcpyDeviceToHost, s));
        GCK(hipStreamSynchronize(s));
        return h;
    }

    // rows (okf, olm, ouv[, odisp]) of Nin inputs, minus those at their own host and those keep_in rejects
    void build(Pass &P, const int32_t *okf, const int32_t *olm, const double *ouv, const double *odisp, int Nin,
               const uint8_t *keep_in) {
        int32_t *err = flag + 4;
        GCK(hipMemsetAsync(err, 0, sizeof(int32_t), s));
        P.lstart = A.get<int32_t>(nl + 1);
        GCK(hipMemsetAsync(P.lstart, 0, sizeof(int32_t) * (nl + 1), s));
        k_gba_check_host<<<nblk(nl, 256), 256, 0, s>>>(nl, nk, q.host, err);
        if (q.n_rot) k_gba_check_pairs<<<nblk(q.n_rot, 128), 128, 0, s>>>(q.n_rot, nk, q.rot_ab, err);
        if (q.n_weld) k_gba_check_pairs<<<nblk(q.n_weld, 128), 128, 0, s>>>(q.n_weld, nk, q.weld_ab, err);
        GCK(hipGetLastError());
        if (read1(err)) throw RangeFail{};
        if (Nin) k_gba_count<<<nblk(Nin, 256), 256, 0, s>>>(Nin, nk, nl, okf, olm, q.host, keep_in, P.lstart, err);
        k_gba_scan<<<1, 1024, 0, s>>>(nl, P.lstart);
        GCK(hipGetLastError());
        if (read1(err)) throw RangeFail{};
        P.M = read1(P.lstart + nl);
        const int M = P.M;
        int32_t *fill = A.get<int32_t>(nl + 1), *src = A.get<int32_t>(M);
        GCK(hipMemcpyAsync(fill, P.lstart, sizeof(int32_t) * nl, hipMemcpyDeviceToDevice, s));
        if (Nin) k_gba_scatter<<<nblk(Nin, 256), 256, 0, s>>>(Nin, okf, olm, q.host, keep_in, fill, src);
        k_gba_sort_rows<<<nblk(nl, 256), 256, 0, s>>>(nl, P.lstart, src);
        P.rkf = A.get<int32_t>(M);
        P.rlm = A.get<int32_t>(M);
        P.ruv = A.get<double>(2 * (size_t)M);
        P.rd = odisp ? A.get<double>(M) : nullptr;
        if (M) k_gba_gather<<<nblk(M, 256), 256, 0, s>>>(M, src, okf, olm, ouv, odisp, P.rkf, P.rlm, P.ruv, P.rd, nullptr);
        // visits
        P.vstart = A.get<int32_t>(nk + 1);
        GCK(hipMemsetAsync(P.vstart, 0, sizeof(int32_t) * (nk + 1), s));
        k_gba_visit_count<<<nblk((int64_t)M + nl, 256), 256, 0, s>>>(M, nl, P.rkf, q.host, P.vstart);
        k_gba_scan<<<1, 1024, 0, s>>>(nk, P.vstart);
        int32_t *vfill = A.get<int32_t>(nk + 1);
        GCK(hipMemcpyAsync(vfill, P.vstart, sizeof(int32_t) * nk, hipMemcpyDeviceToDevice, s));
        P.visit = A.get<int32_t>((size_t)M + nl);
        k_gba_visit_fill<<<nblk((int64_t)M + nl, 256), 256, 0, s>>>(M, nl, P.rkf, q.host, vfill, P.visit);
        // block pattern
        const int64_t wpr = (nk + 31) / 32;
        uint32_t *bits = A.get<uint32_t>((size_t)nk * wpr);
        GCK(hipMemsetAsync(bits, 0, sizeof(uint32_t) * nk * wpr, s));
        k_gba_pattern<<<nblk(std::max(nk, nl), 256), 256, 0, s>>>(nk, nl, P.lstart, P.rkf, q.host, bits, wpr);
        if (q.n_rot) k_gba_pattern_pairs<<<nblk(q.n_rot, 128), 128, 0, s>>>(q.n_rot, q.rot_ab, bits, wpr);
        if (q.n_weld) k_gba_pattern_pairs<<<nblk(q.n_weld, 128), 128, 0, s>>>(q.n_weld, q.weld_ab, bits, wpr);
        P.rowptr = A.get<int32_t>(nk + 1);
        k_gba_rowcount<<<nblk(nk, 4), 256, 0, s>>>(nk, bits, wpr, P.rowptr);
        k_gba_scan<<<1, 1024, 0, s>>>(nk, P.rowptr);
        GCK(hipGetLastError());
        std::vector<int32_t> rp(nk + 1);
        GCK(hipMemcpyAsync(rp.data(), P.rowptr, sizeof(int32_t) * (nk + 1), hipMemcpyDeviceToHost, s));
        GCK(hipStreamSynchronize(s));
        P.nnzb = rp[nk];
        P.col = A.get<int32_t>(P.nnzb);
        k_gba_rowfill<<<nblk(nk, 128), 128, 0, s>>>(nk, bits, wpr, P.rowptr, P.col);
        P.brow = A.get<int32_t>(P.nnzb);
        k_gba_brow<<<nblk(nk, 256), 256, 0, s>>>(nk, P.rowptr, P.brow);
        GCK(hipGetLastError());
        GCK(hipStreamSynchronize(s));
        A.release(bits);
        // chunks of block rows that fit the assembly kernel's LDS
        const int cap = (int)((GBA_LDS_CAP - 6 * sizeof(double)) / (36 * sizeof(double) + sizeof(int32_t)));
        std::vector<int4> ch;
        size_t lmax = 0;
        for (int p = 0; p < nk; p++) {
            const int deg = rp[p + 1] - rp[p];
            for (int c0 = 0; c0 < std::max(deg, 1); c0 += cap) {
                const int c1 = std::min(deg, c0 + cap);
                ch.push_back(make_int4(p, c0, c1, c0 == 0));
                lmax = std::max(lmax, (size_t)(c1 - c0) * (36 * sizeof(double) + sizeof(int32_t)) + 6 * sizeof(double));
            }
        }
        P.nchunks = (int)ch.size();
        P.lds = lmax;
        P.chunks = A.get<int4>(ch.size());
        GCK(hipMemcpyAsync(P.chunks, ch.data(), sizeof(int4) * ch.size(), hipMemcpyHostToDevice, s));
        P.Hb = A.get<double>(36 * (size_t)P.nnzb);
        P.bp = A.get<double>(n);
        GCK(hipStreamSynchronize(s));
    }

    // (vision, stereo) chi at a state slot; e (row residual norms) when asked
    void cost(const Pass &P, int slot, double *out2, double *erow = nullptr) {
        Dev G = dev(P, slot, 0.0);
        const int nb = nblk(nl, 256);
        if ((size_t)nb > cpart_n) {
            cpart = A.get<double>(2 * (size_t)nb);
            cpart_n = nb;
        }
        k_gba_cost<<<nb, 256, 0, s>>>(G, cpart, erow);
        k_gba_sum<<<1, 1024, 0, s>>>(cpart, nb, 2, csum);
        GCK(hipGetLastError());
        GCK(hipMemcpyAsync(out2, csum, 2 * sizeof(double), hipMemcpyDeviceToHost, s));
        GCK(hipStreamSynchronize(s));
    }
    size_t cpart_n = 0;

    void sum_parts(int np, int stage) { k_pcg_ctl<<<1, 1024, 0, s>>>(st, part, np, stage); }

    // the PCG set-up of the assembled system (segment inverses, deflation basis, coarse inverse); false on a set-up
    // failure (a block not numerically positive definite)
    int set_nc = 0, set_ldE = 0;
    bool setup(const Pass &P) {
        const int es0 = mark();
        GCK(hipMemsetAsync(flag, 0, sizeof(int32_t), s));
        k_seg_inv<<<nseg, 1024, sizeof(double) * 36 * (size_t)seg_max * seg_max, s>>>(segk, boff, P.rowptr, P.col, P.brow,
                                                                                P.Hb, Binv, flag);
        k_seg_z<<<nseg, XBA_WAVE, 0, s>>>(segk, T[cur], q.free_mask, Zg, rcount);
        k_gba_scan<<<1, 1024, 0, s>>>(nseg, rcount);
        GCK(hipGetLastError());
        const int nc = read1(rcount + nseg);
        const int ldE = std::max(SW_T, ((nc + SW_T - 1) / SW_T) * SW_T);
        GCK(hipMemsetAsync(E, 0, sizeof(double) * ldE * ldE, s));
        k_coarse_E<<<nblk(nk, 4), 256, 0, s>>>(nk, P.rowptr, P.col, P.Hb, seg_of, rcount, Zg, E, ldE);
        k_coarse_diag<<<nblk(nc, 256), 256, 0, s>>>(nc, ldE, E, dsc);
        k_coarse_scale<<<nblk((int64_t)ldE * ldE, 256), 256, 0, s>>>(nc, ldE, E, dsc);
        const int ntl = ldE / SW_T;
        for (int k = 0; k < ntl; k++) {
            k_sw_pivot<<<1, 1024, 0, s>>>(E, ldE, k, Pt, flag);
            k_sw_strip<<<ntl, 256, 0, s>>>(E, ldE, k, Pt, St, flag);
            k_sw_update<<<ntl * (ntl + 1) / 2, 256, 0, s>>>(E, ldE, k, St, flag);
            k_sw_finish<<<ntl, 256, 0, s>>>(E, ldE, k, Pt, St, flag);
        }
        k_coarse_unscale<<<nblk((int64_t)nc * ldE, 256), 256, 0, s>>>(nc, ldE, E, dsc);
        k_coarse_mirror<<<nblk((int64_t)nc * ldE, 256), 256, 0, s>>>(nc, ldE, E);
        GCK(hipGetLastError());
        const int es1 = mark();
        if (es0 >= 0) setup_ev.push_back({es0, es1});
        set_nc = nc;
        set_ldE = ldE;
        return read1(flag) == 0;
    }

    // the DPCG solve of the assembled, gauged, damped system P.Hb / P.bp into x (gba_linsolve.py:447-478)
    PcgInfo dpcg(const Pass &P, double rtol) {
        if (!setup(P)) {
            PcgInfo inf;
            inf.reason = PR_SETUP;
            return inf;
        }
        return recur(P, P.bp, rtol);
    }

    // the recursion (gba_linsolve.py:353-444) for right-hand side b with the current set-up, into x
    PcgInfo recur(const Pass &P, const double *b, double rtol) {
        PcgInfo inf;
        const int nred = std::min(GBA_RED, nblk(n, 256));
        const int nsp = std::min(GBA_RED, nblk(nk, 4));
        const int nc = set_nc, ldE = set_ldE;
        const int es1 = mark();
        k_norm2<<<nred, 256, 0, s>>>(n, b, part);
        Pcg h{};
        h.maxiter = o.maxiter;
        GCK(hipMemcpyAsync(st, &h, sizeof(Pcg), hipMemcpyHostToDevice, s));
        k_gba_sum<<<1, 1024, 0, s>>>(part, nred, 1, csum);
        double nb2 = 0;
        GCK(hipMemcpyAsync(&nb2, csum, sizeof(double), hipMemcpyDeviceToHost, s));
        GCK(hipStreamSynchronize(s));
        const double nb = sqrt(nb2);
        if (nb == 0.0 || !std::isfinite(nb)) {
            GCK(hipMemsetAsync(x, 0, sizeof(double) * n, s));
            inf.converged = nb == 0.0;
            inf.reason = nb == 0.0 ? PR_ZERO_RHS : PR_NONFINITE_RHS;
            inf.relres = nb == 0.0 ? 0.0 : NAN;
            return inf;
        }
        h.nb = nb;
        h.tol = rtol * nb;
        GCK(hipMemcpyAsync(st, &h, sizeof(Pcg), hipMemcpyHostToDevice, s));
        // x0 = Z E^-1 Z^T b, r = b - H x0
        k_zt<<<nseg, XBA_WAVE, 0, s>>>(segk, rcount, Zg, b, yc, st, 1);
        k_einv<<<nblk(nc, 4), 256, 0, s>>>(nc, ldE, E, yc, cc, st, 1);
        k_zapply<<<nseg, 256, 0, s>>>(segk, rcount, Zg, cc, nullptr, x, st, 1, 1, 1);
        k_spmv<<<nsp, 256, 0, s>>>(nk, P.rowptr, P.col, P.Hb, x, r, b, 1, 1, st, part);
        sum_parts(nsp, CT_INIT);
        GCK(hipMemcpyAsync(xb, x, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
        restart_tail(P, nc, ldE, nsp, 1);
        GCK(hipGetLastError());
        const int K = 10;
        for (;;) {
            GCK(hipMemcpyAsync(&h, st, sizeof(Pcg), hipMemcpyDeviceToHost, s));
            GCK(hipStreamSynchronize(s));
            if (h.done) break;
            if (h.paused) {  // residual replacement: r = b - H x after x += Z E^-1 Z^T (b - H x)
                k_zt<<<nseg, XBA_WAVE, 0, s>>>(segk, rcount, Zg, rt, yc, st, 1);
                k_einv<<<nblk(nc, 4), 256, 0, s>>>(nc, ldE, E, yc, cc, st, 1);
                k_zapply<<<nseg, 256, 0, s>>>(segk, rcount, Zg, cc, nullptr, x, st, 2, 1, 1);
                k_spmv<<<nsp, 256, 0, s>>>(nk, P.rowptr, P.col, P.Hb, x, r, b, 1, 1, st, part);
                h.paused = 0;
                h.matvecs += 1;
                GCK(hipMemcpyAsync(st, &h, sizeof(Pcg), hipMemcpyHostToDevice, s));
                restart_tail(P, nc, ldE, nsp, 1);
                continue;
            }
            for (int i = 0; i < K; i++) iteration(P, b, nc, ldE, nsp, nred);
            GCK(hipGetLastError());
        }
        const int es2 = mark();
        if (es1 >= 0) pcg_ev.push_back({es1, es2});
        inf.reason = h.reason;
        inf.iters = h.it;
        inf.matvecs = h.matvecs;
        inf.restarts = h.restarts;
        inf.converged = h.reason == PR_TOL;
        inf.relres = (inf.converged ? h.truev : h.best) / nb;
        if (h.use_best) GCK(hipMemcpyAsync(x, xb, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
        return inf;
    }

    // z = M^-1 r, rz = r.z, p = P z (the start of the recursion and of a restart)
    void restart_tail(const Pass &P, int nc, int ldE, int nsp, int force) {
        k_minv<<<nseg, 256, 0, s>>>(segk, boff, Binv, r, z, st, force, part);
        sum_parts(nseg, CT_RZ);
        k_spmv<<<nsp, 256, 0, s>>>(nk, P.rowptr, P.col, P.Hb, z, hz, nullptr, 0, force, st, part);
        k_zt<<<nseg, XBA_WAVE, 0, s>>>(segk, rcount, Zg, hz, yc, st, force);
        k_einv<<<nblk(nc, 4), 256, 0, s>>>(nc, ldE, E, yc, cc, st, force);
        k_zapply<<<nseg, 256, 0, s>>>(segk, rcount, Zg, cc, z, pv, st, 0, 1, force);
    }

    // one CG iteration of gba_linsolve.py:395-444 as a fixed launch sequence
    void iteration(const Pass &P, const double *b, int nc, int ldE, int nsp, int nred) {
        sum_parts(1, CT_START);  // part[0] unused by START
        k_spmv<<<nsp, 256, 0, s>>>(nk, P.rowptr, P.col, P.Hb, pv, w, nullptr, 0, 0, st, part);
        sum_parts(nsp, CT_ALPHA);
        k_axpy<<<nred, 256, 0, s>>>(n, st, x, r, pv, w, part);
        sum_parts(nred, CT_CHECK);
        k_spmv<<<nsp, 256, 0, s>>>(nk, P.rowptr, P.col, P.Hb, x, rt, b, 1, 0, st, part);
        sum_parts(nsp, CT_TRUE);
        k_copybest<<<nred, 256, 0, s>>>(n, st, x, xb);
        k_minv<<<nseg, 256, 0, s>>>(segk, boff, Binv, r, z, st, 0, part);
        sum_parts(nseg, CT_BETA);
        k_spmv<<<nsp, 256, 0, s>>>(nk, P.rowptr, P.col, P.Hb, z, hz, nullptr, 0, 0, st, part);
        k_zt<<<nseg, XBA_WAVE, 0, s>>>(segk, rcount, Zg, hz, yc, st, 0);
        k_einv<<<nblk(nc, 4), 256, 0, s>>>(nc, ldE, E, yc, cc, st, 0);
        k_zapply<<<nseg, 256, 0, s>>>(segk, rcount, Zg, cc, z, pv, st, 0, 0, 0);
    }

    // global_ba.py:2080-2126: the 6x6 diagonal blocks of the inverse of the last damped reduced system for want
    // free keyframes sampled evenly (every step-th, step = max(1, free / want)), column by column with the PCG
    int marginal_cov(const Pass &P, int want, int32_t *ids_out, double *cov_out) {
        if (want <= 0 || n > XBA_GBA_DIRECT_MAX) return 0;
        std::vector<uint8_t> fm(nk);
        GCK(hipMemcpyAsync(fm.data(), q.free_mask, nk, hipMemcpyDeviceToHost, s));
        GCK(hipStreamSynchronize(s));
        std::vector<int> ids;
        for (int i = 0; i < nk; i++)
            if (fm[i]) ids.push_back(i);
        if (ids.empty()) return 0;
        const int step = std::max<int>(1, (int)ids.size() / std::max(want, 1));
        std::vector<int> pick;
        for (size_t i = 0; i < ids.size() && (int)pick.size() < want; i += step) pick.push_back(ids[i]);
        if (!setup(P)) return 0;
        double *e = A.get<double>(n);
        const double one = 1.0;
        int cnt = 0;
        for (int i : pick) {
            double Cm[36];
            bool ok = true;
            for (int c = 0; c < 6 && ok; c++) {
                GCK(hipMemsetAsync(e, 0, sizeof(double) * n, s));
                GCK(hipMemcpyAsync(e + 6 * (size_t)i + c, &one, sizeof(double), hipMemcpyHostToDevice, s));
                const PcgInfo inf = recur(P, e, o.rtol_direct);
                const bool usable = inf.converged || (inf.reason != PR_SETUP && std::isfinite(inf.relres) &&
                                                      inf.relres <= o.rtol_direct_floor);
                if (!usable) {
                    ok = false;
                    break;
                }
                double col[6];
                GCK(hipMemcpyAsync(col, x + 6 * (size_t)i, sizeof(col), hipMemcpyDeviceToHost, s));
                GCK(hipStreamSynchronize(s));
                for (int r = 0; r < 6; r++) Cm[6 * r + c] = col[r];
            }
            for (int v = 0; v < 36 && ok; v++) ok = std::isfinite(Cm[v]);
            if (!ok) continue;
            ids_out[cnt] = i;
            for (int r = 0; r < 6; r++)
                for (int c = 0; c < 6; c++) cov_out[36 * (size_t)cnt + 6 * r + c] = 0.5 * (Cm[6 * r + c] + Cm[6 * c + r]);
            cnt++;
        }
        return cnt;
    }

    double kth(const double *v, int cnt, int64_t k) {
        if (!keys_n || keys_n < (size_t)cnt) {
            keys = A.get<uint64_t>(cnt);
            keys_n = cnt;
        }
        k_to_keys<<<nblk(cnt, 256), 256, 0, s>>>(cnt, v, keys);
        const uint64_t init[3] = {0ull, (uint64_t)k, 0ull};
        GCK(hipMemcpyAsync(sel, init, sizeof(init), hipMemcpyHostToDevice, s));
        for (int ps = 0; ps < 8; ps++) {
            k_sel_hist<<<std::min(nblk(cnt, 256), 1024), 256, 0, s>>>(cnt, keys, ps, sel, hist);
            k_sel_pick<<<1, 64, 0, s>>>(ps, sel, hist);
        }
        GCK(hipGetLastError());
        const uint64_t kk = read1(sel + 2);
        const uint64_t bits = (kk >> 63) ? (kk & ~(1ull << 63)) : ~kk;
        double d;
        memcpy(&d, &bits, sizeof(d));
        return d;
    }
    size_t keys_n = 0;

    // numpy's median of the first cnt order statistics of v (n entries; the others sort after them)
    double median(const double *v, int cnt) {
        if (cnt <= 0) return NAN;
        const double a = kth(v, cnt, (cnt - 1) / 2);
        if (cnt % 2) return a;
        const double b = kth(v, cnt, cnt / 2);
        return 0.5 * (a + b);
    }
    double median_n(const double *v, int n_all, int cnt) {  // the cnt smallest of n_all
        if (cnt <= 0) return NAN;
        const double a = kth(v, n_all, (cnt - 1) / 2);
        if (cnt % 2) return a;
        const double b = kth(v, n_all, cnt / 2);
        return 0.5 * (a + b);
    }

    // one LM pass (global_ba.py:999-1500 for the default path); the best state ends in slot cur
    void lm_pass(const Pass &P, int iters, double lam0, double t0, double budget, int pass, PassStats &ps,
                 double *hist_out, int hist_cap, int &nh) {
        double c2[2];
        cost(P, cur, c2);
        ps.chi0v = c2[0];
        ps.chi0s = c2[1];
        double chi = c2[0] + c2[1];
        double best = c2[0];  // the best state is seeded with the vision chi (global_ba.py:958-960)
        GCK(hipMemcpyAsync(T[bst], T[cur], sizeof(double) * 16 * nk, hipMemcpyDeviceToDevice, s));
        GCK(hipMemcpyAsync(RH[bst], RH[cur], sizeof(double) * nl, hipMemcpyDeviceToDevice, s));
        GCK(hipMemcpyAsync(MB[bst], MB[cur], sizeof(double) * 3 * nl, hipMemcpyDeviceToDevice, s));
        double lam = lam0;
        const double rtol = n > XBA_GBA_DIRECT_MAX ? o.rtol : o.rtol_direct;
        for (int it = 0; it < iters; it++) {
            if (budget > 0 && now_s() - t0 > budget) break;
            ps.iters_run++;
            Dev G = dev(P, cur, lam);
            if ((size_t)P.M > rrec_n) {
                if (rrec) A.release(rrec);
                rrec = A.get<double>(RR_N * (size_t)P.M);
                rrec_n = P.M;
            }
            const int e0 = mark();
            k_gba_rows<<<nblk(nl, 128), 128, 0, s>>>(G, rrec, whs);
            k_gba_assemble<<<P.nchunks, GBA_WG, P.lds, s>>>(G, P.chunks, rrec, whs);
            k_gba_mirror<<<nblk(P.nnzb, 256), 256, 0, s>>>(P.nnzb, P.brow, P.rowptr, P.col, P.Hb);
            if (q.n_rot + q.n_weld)
                k_gba_pose_factors<<<nblk(q.n_rot + q.n_weld, 64), 64, 0, s>>>(G, q.n_rot, q.rot_ab, q.rot_M,
                                                                              q.rot_sig, q.rot_huber, q.n_weld,
                                                                              q.weld_ab, q.weld_Z, q.weld_w);
            const int e1 = mark();
            if (e0 >= 0) asm_ev.push_back({e0, e1});
            k_gba_gauge<<<nblk(std::max<int64_t>(P.nnzb, nk), 256), 256, 0, s>>>(G, P.brow, P.nnzb);
            GCK(hipGetLastError());
            const PcgInfo inf = dpcg(P, rtol);
            solves++;
            pcg_iters += inf.iters;
            double rec[XBA_GBA_HIST] = {(double)pass, (double)it, lam, chi, NAN, 0.0, (double)inf.iters, inf.relres,
                                        (double)inf.reason, (double)inf.restarts};
            // at or below XBA_GBA_DIRECT_MAX unknowns the reference solves directly: a PCG that stops at the
            // attainable accuracy of its recursion (stagnation, divergence, maxiter) within rtol_direct_floor of
            // |b| stands in for that solve with its best iterate
            const bool direct = n <= XBA_GBA_DIRECT_MAX;
            const bool usable = inf.converged || (direct && inf.reason != PR_SETUP && std::isfinite(inf.relres) &&
                                                  inf.relres <= o.rtol_direct_floor);
            if (!usable) {
                unconv++;
                lam = std::min(lam * 10.0, 1e3);
                if (hist_out && nh < hist_cap) memcpy(hist_out + XBA_GBA_HIST * nh, rec, sizeof(rec));
                nh++;
                continue;
            }
            GCK(hipMemsetAsync(flag + 1, 0, sizeof(int32_t), s));
            k_gba_trial<<<nblk(nk, 128), 128, 0, s>>>(nk, T[cur], q.free_mask, x, T[tri], flag + 1);
            k_gba_backsub<<<nblk(nl, 128), 128, 0, s>>>(G, x, q.rho_floor, RH[tri], MB[tri], flag + 2);
            GCK(hipGetLastError());
            if (read1(flag + 1)) {
                lam = std::min(lam * 10.0, 1e3);
                if (hist_out && nh < hist_cap) memcpy(hist_out + XBA_GBA_HIST * nh, rec, sizeof(rec));
                nh++;
                continue;
            }
            cost(P, tri, c2);
            const double chi2 = c2[0] + c2[1];
            rec[4] = chi2;
            if (chi2 <= chi) {
                rec[5] = 1.0;
                ps.acc++;
                std::swap(cur, tri);
                chi = chi2;
                lam = std::max(lam * 0.4, 1e-7);
                if (chi2 < best) {
                    ps.best++;
                    best = chi2;
                    GCK(hipMemcpyAsync(T[bst], T[cur], sizeof(double) * 16 * nk, hipMemcpyDeviceToDevice, s));
                    GCK(hipMemcpyAsync(RH[bst], RH[cur], sizeof(double) * nl, hipMemcpyDeviceToDevice, s));
                    GCK(hipMemcpyAsync(MB[bst], MB[cur], sizeof(double) * 3 * nl, hipMemcpyDeviceToDevice, s));
                }
            } else {
                lam = std::min(lam * 5.0, 1e3);
            }
            if (hist_out && nh < hist_cap) memcpy(hist_out + XBA_GBA_HIST * nh, rec, sizeof(rec));
            nh++;
        }
        ps.chibest = best;
        std::swap(cur, bst);  // the returned state is the best one
    }
};

}  // namespace gba

extern "C" int xba_gba_solve(const xba_gba_problem *q, const xba_gba_options *o, double *stats, double *hist,
                             int32_t hist_cap, void *stream) {
    using namespace gba;
    if (!q || !o || !stats) return XBA_E_ARG;
    if (q->nk <= 0 || q->nl < 0 || q->N < 0 || q->nk > XBA_GBA_MAX_NK) return XBA_E_ARG;
    if (!q->poses || (q->nl && (!q->rho || !q->mbear || !q->host)) || !q->free_mask) return XBA_E_ARG;
    if (q->N && (!q->okf || !q->olm || !q->ouv)) return XBA_E_ARG;
    const bool pr = q->pr_m0 != nullptr;
    if (pr && (!q->pr_rho0 || !q->pr_wbear || !q->pr_invsig || !q->pr_valid)) return XBA_E_ARG;
    if (o->iters < 0 || o->passes < 1 || o->maxiter < 1 || o->seg_len < 0 || o->seg_len > GBA_SEG_MAX) return XBA_E_ARG;
    if (q->n_rot < 0 || q->n_weld < 0 || (q->n_rot && (!q->rot_ab || !q->rot_M || !q->rot_sig)) ||
        (q->n_weld && (!q->weld_ab || !q->weld_Z || !q->weld_w)))
        return XBA_E_ARG;
    hipStream_t s = (hipStream_t)stream;
    for (int i = 0; i < XBA_GBA_NSTATS; i++) stats[i] = NAN;
    try {
        const double t0 = now_s();
        Solver S(*q, *o, s);
        Pass P1;
        P1.pisg = q->pr_invsig;
        S.build(P1, q->okf, q->olm, q->ouv, q->odisp && q->fB > 0 ? q->odisp : nullptr, q->N, nullptr);
        double c2[2];
        S.cost(P1, S.cur, c2);
        const int N1 = P1.M;
        stats[XBA_GBA_ST_RMS0] = N1 ? sqrt(c2[0] / N1) : NAN;
        stats[XBA_GBA_ST_OBS] = N1;
        stats[XBA_GBA_ST_NSEG] = S.nseg;
        stats[XBA_GBA_ST_SEGLEN] = S.m;
        stats[XBA_GBA_ST_NNZB] = (double)P1.nnzb;
        stats[XBA_GBA_ST_STEREO] = 0;
        if (P1.rd && N1) {  // disparity rows: d > 1e-6, counted on the device
            GCK(hipMemsetAsync(S.flag + 6, 0, sizeof(int32_t), s));
            k_count_above<<<std::min(nblk(N1, 256), 1024), 256, 0, s>>>(N1, P1.rd, 1e-6, S.flag + 6);
            stats[XBA_GBA_ST_STEREO] = S.read1(S.flag + 6);
        }
        int nh = 0;
        PassStats ps1;
        const double budget = o->time_budget_s > 0 ? o->time_budget_s : 0.0;
        S.lm_pass(P1, o->iters, o->lam0, t0, budget, 1, ps1, hist, hist_cap, nh);
        stats[XBA_GBA_ST_P1_CHI0V] = ps1.chi0v;
        stats[XBA_GBA_ST_P1_CHI0S] = ps1.chi0s;
        stats[XBA_GBA_ST_P1_CHIBEST] = ps1.chibest;
        stats[XBA_GBA_ST_P1_ACC] = ps1.acc;
        stats[XBA_GBA_ST_P1_BEST] = ps1.best;
        PassStats last = ps1;
        const Pass *Pf = &P1;
        Pass P2;
        int passes_run = 1, n_rej = 0, iters_run = ps1.iters_run;
        double thr = NAN, first_sigma = NAN;
        if (o->passes > 1 && (budget <= 0 || now_s() - t0 < 0.6 * budget) && N1) {
            S.e = S.A.get<double>(N1);
            S.cost(P1, S.cur, c2, S.e);
            const double med = S.median(S.e, N1);
            S.tmpv = S.A.get<double>(std::max(N1, S.nl));
            double hm = med;
            GCK(hipMemcpyAsync(S.csum + 2, &hm, sizeof(double), hipMemcpyHostToDevice, s));
            k_absdev<<<nblk(N1, 256), 256, 0, s>>>(N1, S.e, S.csum + 2, S.tmpv);
            const double mad = S.median(S.tmpv, N1) * 1.4826;
            thr = med + o->reject_sigma * std::max(mad, 1e-6);
            uint8_t *keep = S.A.get<uint8_t>(N1);
            GCK(hipMemsetAsync(S.flag + 3, 0, sizeof(int32_t), s));
            k_keep<<<nblk(N1, 256), 256, 0, s>>>(N1, S.e, thr, keep, S.flag + 3);
            const int nkeep = S.read1(S.flag + 3);
            n_rej = N1 - nkeep;
            const double *pisg2 = q->pr_invsig;
            if (pr) {
                Dev G = S.dev(P1, S.cur, 0.0);
                S.tmpv2 = S.A.get<double>(S.nl);
                GCK(hipMemsetAsync(S.flag + 5, 0, sizeof(int32_t), s));
                k_prior_res<<<nblk(S.nl, 256), 256, 0, s>>>(G, S.tmpv2, S.tmpv, S.flag + 5);
                const int cnt = S.read1(S.flag + 5);
                if (cnt > 50) {
                    const double spread = S.median_n(S.tmpv, S.nl, cnt) * 1.4826;
                    first_sigma = spread * 0.5;  // core/derive.py:60 DISPARITY_NOISE_PX
                    if (0.05 < spread && spread < 20.0) {
                        double *rw = S.A.get<double>(S.nl);
                        k_scale<<<nblk(S.nl, 256), 256, 0, s>>>(S.nl, q->pr_invsig, 1.0 / spread, rw);
                        pisg2 = rw;
                    }
                }
            }
            if (n_rej && nkeep > 0.5 * N1) {
                P2.pisg = pisg2;
                S.build(P2, P1.rkf, P1.rlm, P1.ruv, nullptr, N1, keep);
                PassStats ps2;
                const double rem = budget > 0 ? budget - (now_s() - t0) : 0.0;
                const double t2 = now_s();
                S.lm_pass(P2, o->iters, o->lam0, t2, budget > 0 ? rem : 0.0, 2, ps2, hist, hist_cap, nh);
                last = ps2;
                Pf = &P2;
                passes_run = 2;
                iters_run += ps2.iters_run;
            }
        }
        // the marginal covariances from the last LM iteration's system of the last pass (before the projection,
        // which does not touch it)
        stats[XBA_GBA_ST_NCOV] = 0;
        if (o->want_cov > 0 && o->cov_ids && o->cov && last.iters_run > 0)
            stats[XBA_GBA_ST_NCOV] = S.marginal_cov(*Pf, o->want_cov, o->cov_ids, o->cov);
        // the final SO(3) projection and the statistics of the returned state
        xba::k_np_project<<<nblk(S.nk, 128), 128, 0, s>>>(S.nk, S.T[S.cur]);
        GCK(hipGetLastError());
        S.cost(*Pf, S.cur, c2);
        stats[XBA_GBA_ST_RMS1] = Pf->M ? sqrt(c2[0] / Pf->M) : NAN;
        stats[XBA_GBA_ST_REJECTED] = n_rej;
        stats[XBA_GBA_ST_THR] = thr;
        stats[XBA_GBA_ST_CHI0V] = last.chi0v;
        stats[XBA_GBA_ST_CHI0S] = last.chi0s;
        stats[XBA_GBA_ST_CHIBEST] = last.chibest;
        stats[XBA_GBA_ST_ACC] = last.acc;
        stats[XBA_GBA_ST_BEST] = last.best;
        stats[XBA_GBA_ST_PASSES] = passes_run;
        stats[XBA_GBA_ST_ITERS_RUN] = iters_run;
        if (pr) {
            Dev G = S.dev(*Pf, S.cur, 0.0);
            double *sg = S.A.get<double>(S.nl), *mg = S.A.get<double>(S.nl);
            GCK(hipMemsetAsync(S.flag + 5, 0, sizeof(int32_t), s));
            k_prior_res<<<nblk(S.nl, 256), 256, 0, s>>>(G, sg, mg, S.flag + 5);
            const int cnt = S.read1(S.flag + 5);
            if (cnt > 0) {
                stats[XBA_GBA_ST_SIGMA] = S.median_n(mg, S.nl, cnt) * 1.4826 * 0.5;
                stats[XBA_GBA_ST_BIAS] = S.median_n(sg, S.nl, cnt) * 0.5;
            }
            if (passes_run == 2 && std::isfinite(first_sigma)) stats[XBA_GBA_ST_SIGMA] = first_sigma;
        }
        stats[XBA_GBA_ST_SIGMA1] = first_sigma;
        stats[XBA_GBA_ST_SOLVES] = S.solves;
        stats[XBA_GBA_ST_PCG_IT] = S.pcg_iters;
        stats[XBA_GBA_ST_UNCONV] = S.unconv;
        stats[XBA_GBA_ST_ROWS] = Pf->M;
        xba::k_np_inv_poses<<<nblk(S.nk, 128), 128, 0, s>>>(S.nk, S.T[S.cur], q->poses);
        if (S.nl) {
            GCK(hipMemcpyAsync(q->rho, S.RH[S.cur], sizeof(double) * S.nl, hipMemcpyDeviceToDevice, s));
            GCK(hipMemcpyAsync(q->mbear, S.MB[S.cur], sizeof(double) * 3 * S.nl, hipMemcpyDeviceToDevice, s));
        }
        GCK(hipStreamSynchronize(s));
        stats[XBA_GBA_ST_SECONDS] = now_s() - t0;
        stats[XBA_GBA_ST_ASM_MS] = S.mean_ms(S.asm_ev);
        stats[XBA_GBA_ST_SETUP_MS] = S.mean_ms(S.setup_ev);
        stats[XBA_GBA_ST_PCG_MS] = S.mean_ms(S.pcg_ev);
        (void)nh;
        return XBA_OK;
    } catch (const RangeFail &) {
        (void)hipStreamSynchronize(s);
        return XBA_E_RANGE;
    } catch (const HipFail &) {
        (void)hipStreamSynchronize(s);
        return XBA_E_HIP;
    } catch (const std::bad_alloc &) {
        return XBA_E_HIP;
    }
}
