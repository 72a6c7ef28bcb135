// Marginalisation algebra of the square-root prior, FP64, batched over matrices (vio/marg.py; paths relative to the
// reference's src/xclean/):
//   xba_marg_schur      marg.py:165-177 schur: the kept block's Schur complement, H_mm inverted through its
//                       eigen-decomposition with the rank floor (a rank-deficient block marginalises its observable
//                       part)
//   xba_prior_from_info marg.py:73-84 Prior.from_info (S = diag(sqrt l) V^T over eigenvalues above the floor,
//                       r = diag(1 / sqrt l) V^T b), with, optionally, combine's re-expression of b at the blocks'
//                       first linearisation points (marg.py:212-242: b_x0 = b_cur - H d, d = x_cur [-] x0)
//
// One workgroup owns one matrix.  The symmetric eigen-decomposition is a two-sided cyclic Jacobi with the
// round-robin parallel ordering: each round rotates n/2 disjoint index pairs at once (rotation angles from the
// current 2x2 blocks, then the row pass, then the column pass and the eigenvector update, each spread over the whole
// workgroup); a pair is rotated only while its off-diagonal entry exceeds both 1e-15 of the geometric mean of its
// diagonal entries and 1e-15 of the matrix's Frobenius norm, and sweeps repeat until one rotates nothing (or 40).
// Square roots (xba_prior_from_info) of matrices of order up to MG_PACK_N keep both the matrix (its upper triangle,
// packed) and the eigenvectors in LDS and apply a round as one pass over the 2x2 blocks of the round's pairs (each
// block B <- G_k B G_l^T read and written once, the eigenvector columns beside them); larger ones keep the matrix in
// LDS when the batch's leading dimension is at most MG_LDS_LD (row pass, then column pass), else in the global
// workspace, and the eigenvectors in the workspace.
#include "xba_internal.h"

namespace xba {

#define MG_THREADS 1024
#define MG_EIG_FLOOR 1e-12  // marg.py:32 EIG_FLOOR, relative to the largest eigenvalue
#define MG_MAX_SWEEPS 40
#define MG_LDS_LD 128  // the largest leading dimension whose FP64 matrix (128 KiB) is kept in LDS
// the largest (even) order whose packed upper triangle and eigenvectors share the LDS (1.5 n^2 + 0.5 n doubles)
#define MG_PACK_N 112
#define MG_PACK_DOUBLES(n2) ((n2) * ((n2) + 1) / 2 + (n2) * (n2))
static_assert(MG_THREADS >= XBA_MAX_STATE, "the column pass gives every row at least one thread");

__device__ inline double mg_block_sum(double v, double *red) {
    const int tid = threadIdx.x;
    for (int d = XBA_WAVE / 2; d >= 1; d >>= 1) v += __shfl_xor(v, d, XBA_WAVE);
    __syncthreads();
    if (tid % XBA_WAVE == 0) red[tid / XBA_WAVE] = v;
    __syncthreads();
    double s = 0.0;
    for (int w = 0; w < (int)blockDim.x / XBA_WAVE; w++) s += red[w];
    __syncthreads();
    return s;
}

// Eigen-decomposition of the symmetric n x n matrix A (leading dimension ld, overwritten: its diagonal ends as the
// eigenvalues), eigenvectors into the columns of V.  cs, pqs: LDS for n/2 + 1 (c, s) and (p, q) pairs; rotated: an
// LDS flag; red: LDS for the sums.  The
// workgroup has at least n threads (XBA_MAX_STATE <= MG_THREADS), so every row has a thread in the column pass.
// Always inlined, so that every call site's pointers keep their address space (LDS or global) and the passes use
// LDS or global instructions rather than generic (flat) ones.
__device__ __attribute__((always_inline)) inline void mg_jacobi(double *A, double *V, int n, int ld, double *cs,
                                                                int *pqs, int &rotated, double *red) {
    const int tid = threadIdx.x, nt = blockDim.x;
    for (int e = tid; e < n * n; e += nt) V[(e / n) * ld + e % n] = (e / n == e % n) ? 1.0 : 0.0;
    const int N = n + (n & 1);  // players of the round-robin; N - 1 is a dummy when n is odd
    const int np = N / 2;
    double fro = 0.0;
    for (int e = tid; e < n * n; e += nt) {
        const double v = A[(e / n) * ld + e % n];
        fro += v * v;
    }
    const double tiny = 1e-15 * sqrt(mg_block_sum(fro, red));
    for (int sweep = 0; sweep < MG_MAX_SWEEPS; sweep++) {
        // a pair is left alone when its off-diagonal entry is below 1e-15 of the geometric mean of its two diagonal
        // entries (the relative criterion) or below 1e-15 of the matrix's Frobenius norm (the rounding level of the
        // rotations, so that sweeps terminate on semi-definite matrices); a sweep that rotates nothing ends it
        if (tid == 0) rotated = 0;
        __syncthreads();
        for (int rd = 0; rd < N - 1; rd++) {
            // pairs of this round: (N - 1, rd) and ((rd + k) mod (N - 1), (rd - k) mod (N - 1)), k = 1 .. np - 1
            for (int k = tid; k < np; k += nt) {
                int p = k == 0 ? N - 1 : (rd + k) % (N - 1), q = k == 0 ? rd : (rd - k + N - 1) % (N - 1);
                if (p > q) { const int x = p; p = q; q = x; }
                double c = 1.0, s = 0.0;
                if (q < n) {
                    const double apq = A[p * ld + q], app = A[p * ld + p], aqq = A[q * ld + q];
                    if (fabs(apq) > 1e-15 * sqrt(fabs(app * aqq)) && fabs(apq) > tiny) {
                        const double th = (aqq - app) / (2.0 * apq);
                        const double t = (th >= 0.0 ? 1.0 : -1.0) / (fabs(th) + sqrt(th * th + 1.0));
                        c = 1.0 / sqrt(t * t + 1.0);
                        s = t * c;
                        rotated = 1;
                    }
                } else {
                    p = q = 0;  // the dummy player's pair: rotates nothing
                }
                cs[2 * k] = c;
                cs[2 * k + 1] = s;
                pqs[2 * k] = p;
                pqs[2 * k + 1] = q;
            }
            __syncthreads();
            // rows: A <- J^T A; `per` threads per pair, each striding over the columns with its pair fixed
            {
                const int per = nt / np, k = tid / per, j0 = tid % per;
                if (k < np) {
                    const int p = pqs[2 * k], q = pqs[2 * k + 1];
                    const double c = cs[2 * k], s = cs[2 * k + 1];
                    if (s != 0.0)
                        for (int j = j0; j < n; j += per) {
                            const double ap = A[p * ld + j], aq = A[q * ld + j];
                            A[p * ld + j] = c * ap - s * aq;
                            A[q * ld + j] = s * ap + c * aq;
                        }
                }
            }
            __syncthreads();
            // columns: A <- A J, V <- V J; `pr` threads per row, each striding over the pairs
            {
                const int pr = nt / n, i = tid / pr, k0 = tid % pr;
                if (i < n)
                    for (int k = k0; k < np; k += pr) {
                        const double s = cs[2 * k + 1];
                        if (s == 0.0) continue;
                        const int p = pqs[2 * k], q = pqs[2 * k + 1];
                        const double c = cs[2 * k];
                        double ap = A[i * ld + p], aq = A[i * ld + q];
                        A[i * ld + p] = c * ap - s * aq;
                        A[i * ld + q] = s * ap + c * aq;
                        ap = V[i * ld + p];
                        aq = V[i * ld + q];
                        V[i * ld + p] = c * ap - s * aq;
                        V[i * ld + q] = s * ap + c * aq;
                    }
            }
            __syncthreads();
        }
        if (!rotated) break;
        __syncthreads();
    }
}

// Packed upper triangle: entry (i, j) of an order-n2 symmetric matrix.
__device__ inline int mg_pk(int i, int j, int n2) {
    if (i > j) {
        const int t = i;
        i = j;
        j = t;
    }
    return i * (2 * n2 - i + 1) / 2 + (j - i);
}

// The same cyclic Jacobi on an even order n2 with everything in LDS: P the packed upper triangle (overwritten; its
// diagonal ends as the eigenvalues), V (n2 x n2, leading dimension n2) the eigenvectors.  A round computes the
// rotations G_k = [[c, -s], [s, c]] of its pairs, then, in one pass, every 2x2 block B of pairs (k, l), k <= l, as
// G_k B G_l^T and every row of V's pair columns as V G_k^T.  An odd order is padded by the caller with a decoupled
// zero row and column, whose off-diagonal entries stay exactly zero, so it is never rotated.
__device__ __attribute__((always_inline)) inline void mg_jacobi_lds(double *P, double *V, int n2, double *cs,
                                                                    int *pqs, int &rotated, double *red) {
    const int tid = threadIdx.x, nt = blockDim.x, np = n2 / 2;
    for (int e = tid; e < n2 * n2; e += nt) V[e] = (e / n2 == e % n2) ? 1.0 : 0.0;
    double fro = 0.0;
    for (int i = tid; i < n2; i += nt)
        for (int j = i; j < n2; j++) {
            const double v = P[mg_pk(i, j, n2)];
            fro += (i == j ? 1.0 : 2.0) * v * v;
        }
    const double tiny = 1e-15 * sqrt(mg_block_sum(fro, red));
    const int nblk = np * (np + 1) / 2, nwork = nblk + n2 * np;
    // this thread's work items, the same every round: block (k, l) of pairs, or (row i, pair k) of V (the launch has
    // MG_THREADS threads, so MAXI items per thread cover every order up to MG_PACK_N)
    constexpr int MAXI = (MG_PACK_N / 2 * (MG_PACK_N / 2 + 1) / 2 + MG_PACK_N * MG_PACK_N / 2 + MG_THREADS - 1) /
                         MG_THREADS;
    int ia[MAXI], ib[MAXI];
#pragma unroll
    for (int it = 0; it < MAXI; it++) {
        const int b = tid + it * nt;
        ia[it] = -1;
        ib[it] = 0;
        if (b < nblk) {
            int l = (int)((sqrtf(8.0f * (float)b + 1.0f) - 1.0f) * 0.5f);
            while ((l + 1) * (l + 2) / 2 <= b) l++;
            while (l * (l + 1) / 2 > b) l--;
            ia[it] = b - l * (l + 1) / 2;  // k
            ib[it] = l;
        } else if (b < nwork) {
            const int e = b - nblk;
            ia[it] = np + e / np;  // np + row
            ib[it] = e % np;       // pair
        }
    }
    for (int sweep = 0; sweep < MG_MAX_SWEEPS; sweep++) {
        if (tid == 0) rotated = 0;
        __syncthreads();
        for (int rd = 0; rd < n2 - 1; rd++) {
            // pairs of this round: (n2 - 1, rd) and ((rd + k) mod (n2 - 1), (rd - k) mod (n2 - 1)), k = 1 .. np - 1
            for (int k = tid; k < np; k += nt) {
                int p = k == 0 ? n2 - 1 : (rd + k) % (n2 - 1), q = k == 0 ? rd : (rd - k + n2 - 1) % (n2 - 1);
                if (p > q) {
                    const int x = p;
                    p = q;
                    q = x;
                }
                double c = 1.0, sn = 0.0;
                const double apq = P[mg_pk(p, q, n2)], app = P[mg_pk(p, p, n2)], aqq = P[mg_pk(q, q, n2)];
                if (fabs(apq) > 1e-15 * sqrt(fabs(app * aqq)) && fabs(apq) > tiny) {
                    const double th = (aqq - app) / (2.0 * apq);
                    const double t = (th >= 0.0 ? 1.0 : -1.0) / (fabs(th) + sqrt(th * th + 1.0));
                    c = 1.0 / sqrt(t * t + 1.0);
                    sn = t * c;
                    rotated = 1;
                }
                cs[2 * k] = c;
                cs[2 * k + 1] = sn;
                pqs[2 * k] = p;
                pqs[2 * k + 1] = q;
            }
            __syncthreads();
#pragma unroll
            for (int it = 0; it < MAXI; it++) {
                const int a = ia[it];
                if (a < 0) continue;
                if (a < np) {  // block (k, l) = (a, ib)
                    const int k = a, l = ib[it];
                    const double ck = cs[2 * k], sk = cs[2 * k + 1], cl = cs[2 * l], sl = cs[2 * l + 1];
                    if (sk == 0.0 && sl == 0.0) continue;
                    const int pk = pqs[2 * k], qk = pqs[2 * k + 1], pl = pqs[2 * l], ql = pqs[2 * l + 1];
                    const int i00 = mg_pk(pk, pl, n2), i01 = mg_pk(pk, ql, n2), i10 = mg_pk(qk, pl, n2),
                              i11 = mg_pk(qk, ql, n2);
                    const double b00 = P[i00], b01 = P[i01], b10 = P[i10], b11 = P[i11];
                    const double l00 = ck * b00 - sk * b10, l01 = ck * b01 - sk * b11;  // G_k B
                    const double l10 = sk * b00 + ck * b10, l11 = sk * b01 + ck * b11;
                    const double n00 = cl * l00 - sl * l01, n01 = sl * l00 + cl * l01;  // (G_k B) G_l^T
                    const double n10 = cl * l10 - sl * l11, n11 = sl * l10 + cl * l11;
                    if (k == l) {  // the diagonal block: (p, q) and (q, p) share one entry
                        P[i00] = n00;
                        P[i01] = 0.5 * (n01 + n10);
                        P[i11] = n11;
                    } else {
                        P[i00] = n00;
                        P[i01] = n01;
                        P[i10] = n10;
                        P[i11] = n11;
                    }
                } else {  // row a - np of V, pair ib
                    const int i = a - np, k = ib[it];
                    const double sk = cs[2 * k + 1];
                    if (sk == 0.0) continue;
                    const double ck = cs[2 * k];
                    const int p = pqs[2 * k], q = pqs[2 * k + 1];
                    const double vp = V[i * n2 + p], vq = V[i * n2 + q];
                    V[i * n2 + p] = ck * vp - sk * vq;
                    V[i * n2 + q] = sk * vp + ck * vq;
                }
            }
            __syncthreads();
        }
        if (!rotated) break;
        __syncthreads();
    }
}

// The rank of eigenvalue i in ascending order (ties by index).
__device__ inline int mg_rank(const double *A, int ld, int n, int i) {
    const double li = A[i * ld + i];
    int r = 0;
    for (int j = 0; j < n; j++) {
        const double lj = A[j * ld + j];
        r += (lj < li || (lj == li && j < i)) ? 1 : 0;
    }
    return r;
}

// marg.py:165-177.  keep (nk) and marg (m) index lists; H2 (nk x nk, leading dimension ld) and b2 over keep's order.
// Workspace per matrix (MG_WORK): slab 0 holds H_mm (rows 0..m-1) and T = H_km V (rows m..m+nk-1), slab 1 the
// eigenvectors V, slab 2 Wm = T diag(1/l) (0 for the dropped directions), then V^T b_m.
#define MG_WORK(ld) (3 * (size_t)(ld) * (ld) + 2 * (size_t)(ld))
__global__ __launch_bounds__(MG_THREADS) void k_marg_schur(int ld, const int32_t *nkeep, const int32_t *nmarg,
                                                            const int32_t *keep, const int32_t *marg, const double *H,
                                                            const double *b, double *H2, double *b2, double *work) {
    __shared__ double cs[2 * (XBA_MAX_STATE / 2 + 1)];
    __shared__ int pqs[2 * (XBA_MAX_STATE / 2 + 1)];
    __shared__ int rotated;
    __shared__ double red[MG_THREADS / XBA_WAVE];
    const int w = blockIdx.x, tid = threadIdx.x, nt = blockDim.x;
    const int nk = nkeep[w], m = nmarg[w];
    const size_t L2 = (size_t)ld * ld;
    const double *Hw = H + L2 * w, *bw = b + (size_t)ld * w;
    const int32_t *kp = keep + (size_t)ld * w, *mp = marg + (size_t)ld * w;
    double *A = work + MG_WORK(ld) * w, *V = A + L2, *Wm = V + L2, *vb = Wm + L2, *T = A + (size_t)m * ld;
    double *Ho = H2 + L2 * w, *bo = b2 + (size_t)ld * w;
    for (int e = tid; e < m * m; e += nt) {  // H_mm, symmetrised
        const int i = e / m, j = e % m;
        A[i * ld + j] = 0.5 * (Hw[mp[i] * ld + mp[j]] + Hw[mp[j] * ld + mp[i]]);
    }
    __syncthreads();
    if (m > 0) mg_jacobi(A, V, m, ld, cs, pqs, rotated, red);
    // Hinv = V_ok diag(1 / l) V_ok^T over the eigenvalues above the floor relative to the largest one
    double lmax = -INFINITY;
    for (int i = 0; i < m; i++) lmax = fmax(lmax, A[i * ld + i]);
    const double floor_ = MG_EIG_FLOOR * fmax(lmax, 1e-300);
    for (int e = tid; e < nk * m; e += nt) {
        const int i = e / m, j = e % m;
        double s = 0.0;
        for (int q = 0; q < m; q++) s += Hw[kp[i] * ld + mp[q]] * V[q * ld + j];
        T[i * ld + j] = s;
    }
    for (int i = tid; i < m; i += nt) {
        double s = 0.0;
        for (int q = 0; q < m; q++) s += V[q * ld + i] * bw[mp[q]];
        vb[i] = s;
    }
    __syncthreads();
    for (int e = tid; e < nk * m; e += nt) {
        const int i = e / m, j = e % m;
        const double l = A[j * ld + j];
        Wm[i * ld + j] = (l > floor_) ? T[i * ld + j] / l : 0.0;
    }
    __syncthreads();
    // H2 = H_kk - Wm T^T ; b2 = b_k - Wm (V^T b_m)
    for (int e = tid; e < nk * nk; e += nt) {
        const int i = e / nk, j = e % nk;
        double s = 0.0;
        for (int q = 0; q < m; q++) s += Wm[i * ld + q] * T[j * ld + q];
        Ho[i * ld + j] = Hw[kp[i] * ld + kp[j]] - s;
    }
    for (int i = tid; i < nk; i += nt) {
        double s = 0.0;
        for (int q = 0; q < m; q++) s += Wm[i * ld + q] * vb[q];
        bo[i] = bw[kp[i]] - s;
    }
    __syncthreads();
    for (int e = tid; e < nk * nk; e += nt) {  // symmetrise (each pair once)
        const int i = e / nk, j = e % nk;
        if (i < j) {
            const double v = 0.5 * (Ho[i * ld + j] + Ho[j * ld + i]);
            Ho[i * ld + j] = v;
            Ho[j * ld + i] = v;
        }
    }
}

// marg.py:73-84 (and 212-242 when blocks are given): S, r of (H, b) over eigenvalues above the floor.
// blocks (optional): per matrix a range [blk0[w], blk0[w+1]) of blocks, each with its offset and dims in the matrix
// and the current state and first linearisation point (R 9, p, v, bg, ba: 21 FP64 each).
__global__ __launch_bounds__(MG_THREADS) void k_prior_from_info(int ld, const int32_t *nn, const double *H,
                                                                 const double *b, const int32_t *blk0,
                                                                 const int32_t *boff, const int32_t *bdim,
                                                                 const double *bcur, const double *bx0, double *S,
                                                                 double *r, int32_t *rank, double *lam,
                                                                 double *bx_out, double *work) {
    __shared__ double cs[2 * (XBA_MAX_STATE / 2 + 1)];
    __shared__ int pqs[2 * (XBA_MAX_STATE / 2 + 1)];
    __shared__ int rotated;
    __shared__ double red[MG_THREADS / XBA_WAVE];
    extern __shared__ double mg_dyn[];
    const int w = blockIdx.x, tid = threadIdx.x, nt = blockDim.x;
    const int n = nn[w];
    const size_t L2 = (size_t)ld * ld;
    const double *Hw = H + L2 * w, *bw = b + (size_t)ld * w;
    double *Ag = work + MG_WORK(ld) * w, *V = Ag + L2, *bx = V + 2 * L2, *d = bx + ld;  // bx: b at x0
    const int n2 = n + (n & 1);
    // per matrix: the launch holds at least MG_PACK_DOUBLES(min(ld rounded up to even, MG_PACK_N)) doubles of LDS
    const bool packed = n2 <= MG_PACK_N;
    double *A = ld <= MG_LDS_LD ? mg_dyn : Ag;  // the matrix being diagonalised (unpacked paths)
    double *Pk = mg_dyn, *Vl = mg_dyn + n2 * (n2 + 1) / 2;  // the packed path
    double *Sw = S + L2 * w, *rw = r + (size_t)ld * w;
    if (packed) {
        for (int e = tid; e < n2 * n2; e += nt) {
            const int i = e / n2, j = e % n2;
            if (i > j) continue;
            Pk[mg_pk(i, j, n2)] = (i < n && j < n) ? 0.5 * (Hw[i * ld + j] + Hw[j * ld + i]) : 0.0;
        }
    } else {
        for (int e = tid; e < n * n; e += nt) {
            const int i = e / n, j = e % n;
            A[i * ld + j] = 0.5 * (Hw[i * ld + j] + Hw[j * ld + i]);
        }
    }
    // b at the first linearisation points: b - H d, d = x_cur [-] x0 per block (Log(R0^T R), p - p0, ...)
    for (int i = tid; i < n; i += nt) d[i] = 0.0;
    __syncthreads();
    if (blk0)
        for (int k = blk0[w] + tid; k < blk0[w + 1]; k += nt) {
            const double *xc = bcur + 21 * (size_t)k, *x0 = bx0 + 21 * (size_t)k;
            double M[9];
            for (int i = 0; i < 3; i++)
                for (int j = 0; j < 3; j++)
                    M[3 * i + j] = x0[i] * xc[j] + x0[3 + i] * xc[3 + j] + x0[6 + i] * xc[6 + j];  // R0^T R
            double *dk = d + boff[k];
            so3_log_nb(M, dk);
            for (int q = 3; q < bdim[k]; q++) dk[q] = xc[6 + q] - x0[6 + q];  // p, v, bg, ba follow R in the record
        }
    __syncthreads();
    for (int i = tid; i < n; i += nt) {
        double s = 0.0;
        for (int j = 0; j < n; j++) s += Hw[i * ld + j] * d[j];
        bx[i] = bw[i] - s;
        if (bx_out) bx_out[(size_t)ld * w + i] = bx[i];
    }
    __syncthreads();
    // eigenvalues into d (free again), eigenvectors in the columns of Ve (leading dimension ldv)
    const double *Ve = packed ? Vl : V;
    const int ldv = packed ? n2 : ld;
    if (n > 0) {
        if (packed)
            mg_jacobi_lds(Pk, Vl, n2, cs, pqs, rotated, red);
        else if (ld <= MG_LDS_LD)
            mg_jacobi(mg_dyn, V, n, ld, cs, pqs, rotated, red);
        else
            mg_jacobi(Ag, V, n, ld, cs, pqs, rotated, red);
    }
    for (int i = tid; i < n; i += nt) d[i] = packed ? Pk[mg_pk(i, i, n2)] : A[i * ld + i];
    __syncthreads();
    double lmax = -INFINITY;
    for (int i = 0; i < n; i++) lmax = fmax(lmax, d[i]);
    const double floor_ = MG_EIG_FLOOR * fmax(lmax, 1e-300);
    int nkeep = 0;
    for (int i = 0; i < n; i++) nkeep += d[i] > floor_ ? 1 : 0;
    // rows of S in ascending eigenvalue order among the kept ones (eigh's order); the rest of S and r zero
    for (int e = tid; e < n * n; e += nt) Sw[(e / n) * ld + e % n] = 0.0;
    for (int i = tid; i < n; i += nt) rw[i] = 0.0;
    __syncthreads();
    for (int i = tid; i < n; i += nt) {
        const double l = d[i];
        int rk = 0;  // ascending rank, ties by index
        for (int j = 0; j < n; j++) rk += (d[j] < l || (d[j] == l && j < i)) ? 1 : 0;
        if (lam) lam[(size_t)ld * w + rk] = l;
        if (!(l > floor_)) continue;
        const int row = rk - (n - nkeep);  // the kept eigenvalues are the nkeep largest
        const double sl = sqrt(l);
        double vb = 0.0;
        for (int j = 0; j < n; j++) {
            Sw[row * ld + j] = sl * Ve[j * ldv + i];
            vb += Ve[j * ldv + i] * bx[j];
        }
        rw[row] = vb / sl;
    }
    if (tid == 0) rank[w] = nkeep;
}

}  // namespace xba
