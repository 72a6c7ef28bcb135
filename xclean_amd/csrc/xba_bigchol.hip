// Tiled multi-workgroup Cholesky solve of large reduced systems (windows above the one-workgroup size limit),
// right-looking over 64 x 64 tiles of the damped, padded matrix:
//
//   k_big_prep    Hd = H + diag(lam diag(H) + 1e-9) (vio/window.py:933-936), Jacobi-scaled in mixed precision,
//                 copied into the padded layout (identity on the padding), lower tiles only
//   k_big_panel   step k: L(i,k) = A(i,k) L(k,k)^-T for the tiles below the diagonal tile and L(k,k)^-1 into a
//                 side buffer, one per wavefront with a row per lane, each wavefront factoring the diagonal tile
//                 in its registers fused with its triangular solve
//   k_big_update  step k: A(i,j) -= L(i,k) L(j,k)^T over the trailing lower tiles on the matrix cores (FP32 or
//                 FP64 16x16x4 MFMA), one tile per workgroup, 32 x 32 per wavefront
//   k_big_trsv    one solve L L^T y = s of the scaled residual with the tiled factor (one workgroup per window:
//                 tile-row matrix-vector products and the diagonal-tile inverses, no serial column loop), x += y
//   k_big_resid   r = b - Hd x in FP64 over many workgroups, with per-workgroup maxima for the stopping test
//                 |r| <= 1e-15 |b| (mixed precision refines until it holds or the step budget ends)
//   k_big_finish  dx = -x (vio/window.py:937-941) and the trial state
//
// The diagonal tiles of the factor are only needed as their inverses, which the solves apply; the matrix keeps
// A(k,k) (the other workgroups of the same panel launch still read it).
#include "xba_internal.h"

namespace xba {

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x4 __attribute__((ext_vector_type(4)));

template <typename T>
struct Mfma;
// D = A B + C on one wavefront: A 16x4 (lane l supplies A[l % 16][l / 16]), B 4x16 (lane l supplies
// B[l / 16][l % 16]), C/D 16x16 with four elements per lane in column l % 16; row(l / 16, q) gives the row of element
// q, which differs between the two instructions (FP32: four consecutive rows per lane group; FP64: rows interleaved
// with stride 4).
template <>
struct Mfma<float> {
    typedef f32x4 V;
    __device__ static V fma(float a, float b, V c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
    __device__ static int row(int g, int q) { return 4 * g + q; }
};
template <>
struct Mfma<double> {
    typedef f64x4 V;
    __device__ static V fma(double a, double b, V c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
    __device__ static int row(int g, int q) { return g + 4 * q; }
};

template <typename T>
__device__ inline T *big_base(const DArgs &D) {
    if constexpr (sizeof(T) == 4)
        return (T *)D.bigF32;
    else
        return (T *)D.bigF64;
}

// Offset of the window's matrix and of its diagonal-tile side buffer inside the tiled-factor buffer.
__host__ __device__ inline int64_t big_matrix_entries(int np) { return (int64_t)np * np + (int64_t)np * XBA_TILE; }

// ---------------------------------------------------------------------------------------------- k_big_prep --
// grid (max np, nbig): one padded row per workgroup.  gn = 1: lambda = 0 (the Gauss-Newton step).
template <typename T>
__global__ __launch_bounds__(256) void k_big_prep(DArgs D, int gn) {
    __shared__ double ss[XBA_MAX_STATE];
    const BigWin B = D.bigs[blockIdx.y];
    const int i = blockIdx.x, tid = threadIdx.x;
    if (i >= B.np) return;
    const WinInfo W = D.wins[B.win];
    const int n = B.n, np = B.np;
    double *ws = D.wstate + (size_t)NWS * B.win;
    const double lam = gn ? 0.0 : ws[WS_LAM];
    const double *H = D.H + W.h0;
    for (int j = tid; j < n; j += blockDim.x) {
        const double h = H[(size_t)j * n + j], hd = h + (lam * h + 1e-9);
        ss[j] = (sizeof(T) == 8) ? 1.0 : (hd > 0.0 ? 1.0 / sqrt(hd) : 1.0);
    }
    __syncthreads();
    T *F = big_base<T>(D) + B.f0 + (size_t)i * np;
    const double si = i < n ? ss[i] : 1.0;
    const int jend = (i / XBA_TILE + 1) * XBA_TILE;  // through the diagonal tile: nothing above it is read
    for (int j = tid; j < jend; j += blockDim.x) {
        double v = 0.0;
        if (i < n && j < n) {
            double h = H[(size_t)i * n + j];
            if (i == j) h += lam * h + 1e-9;
            v = h * si * ss[j];
        } else if (i == j) {
            v = 1.0;
        }
        F[j] = (T)v;
    }
    if (tid == 0) {
        D.bigS[B.s0 + i] = si;
        if (i == 0) {
            ws[WS_BIGOK] = 1.0;
            ws[WS_BIGCONV] = 0.0;
        }
    }
}

// Broadcast of lane jj's value (jj uniform).
__device__ inline float big_bcast(float v, int jj) {
    return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), jj));
}
__device__ inline double big_bcast(double v, int jj) {
    const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)u, jj);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(u >> 32), jj);
    return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}

// A tile row (XBA_TILE consecutive entries, 16-byte aligned: rows start at multiples of XBA_TILE entries of the
// padded matrix) moved in 16-byte pieces, so that a wavefront whose lanes each move their own row touches each
// cache line of it once per piece rather than once per entry.
template <typename T>
__device__ inline void big_row_load(const T *p, T (&v)[XBA_TILE]) {
    typedef T V4 __attribute__((ext_vector_type(16 / sizeof(T))));
    constexpr int W = 16 / sizeof(T);
#pragma unroll
    for (int q = 0; q < XBA_TILE / W; q++) {
        const V4 x = *(const V4 *)(p + W * q);
#pragma unroll
        for (int e = 0; e < W; e++) v[W * q + e] = x[e];
    }
}
template <typename T>
__device__ inline void big_row_store(T *p, const T (&v)[XBA_TILE]) {
    typedef T V4 __attribute__((ext_vector_type(16 / sizeof(T))));
    constexpr int W = 16 / sizeof(T);
#pragma unroll
    for (int q = 0; q < XBA_TILE / W; q++) {
        V4 x;
#pragma unroll
        for (int e = 0; e < W; e++) x[e] = v[W * q + e];
        *(V4 *)(p + W * q) = x;
    }
}

// --------------------------------------------------------------------------------------------- k_big_panel --
// One column step J of the diagonal tile's factor fused with the same step of a triangular solve x L^T = v, on one
// wavefront: lane r holds row r of the tile in a (entries above the diagonal zero) and row r of the right-hand side
// in v.  The pivot's inverse square root scales column J of L (lane m's a[J], m > J), x_J = v_J / L(J,J), and every
// later entry takes the rank-one updates a[m] -= L(r,J) L(m,J), v[m] -= x_J L(m,J), L(m,J) broadcast from lane m's
// register.  The broadcasts are issued in groups ahead of their multiply-adds (a broadcast's scalar result read by
// the next instruction stalls the wavefront), each group at most BIG_BC_GROUP values (scalar registers: one per
// FP32, two per FP64 value).  The pivots' inverse square roots come from the hardware approximation (FP32: within an
// ulp; FP64: refined by two Newton steps to within an ulp); the pivot is d times it.  ok = 0 and the remaining steps skipped at a pivot that is not positive and finite (uniform: every lane
// reads the same pivot).
#define BIG_BC_GROUP 8
template <typename T, int J, int M0>
__device__ inline void big_rank1x2(T (&a)[XBA_TILE], T (&v)[XBA_TILE], T lij, T xj) {
    constexpr int N = (XBA_TILE - M0 < BIG_BC_GROUP) ? XBA_TILE - M0 : BIG_BC_GROUP;
    T bc[N];
#pragma unroll
    for (int e = 0; e < N; e++) bc[e] = big_bcast(a[J], M0 + e);
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int e = 0; e < N; e++) {
        a[M0 + e] -= lij * bc[e];
        v[M0 + e] -= xj * bc[e];
        // the updates happen here: otherwise the compiler defers each entry's update to its use at step M0 + e and
        // keeps the broadcasts live in scalar registers until then
        asm volatile("" : "+v"(a[M0 + e]), "+v"(v[M0 + e]));
    }
    if constexpr (M0 + N < XBA_TILE) big_rank1x2<T, J, M0 + N>(a, v, lij, xj);
}

template <typename T, int J>
struct BigPanelStep {
    __device__ static inline void step(T (&a)[XBA_TILE], T (&v)[XBA_TILE], int lane, int &ok) {
        const T d = big_bcast(a[J], J);
        if (!(d > (T)0) || !isfinite(d)) {
            ok = 0;
            return;
        }
        T sd, inv;
        if constexpr (sizeof(T) == 4) {
            inv = __builtin_amdgcn_rsqf(d);
        } else {
            // the hardware approximation refined by two Newton steps y <- y + y (1/2 - (d/2) y^2): within an ulp
            const double h = 0.5 * d;
            inv = __builtin_amdgcn_rsq(d);
            inv = fma(inv, fma(-h * inv, inv, 0.5), inv);
            inv = fma(inv, fma(-h * inv, inv, 0.5), inv);
        }
        sd = d * inv;
        a[J] = (lane == J) ? sd : a[J] * inv;  // L[lane][J] below the diagonal
        const T xj = v[J] * inv;
        v[J] = xj;
        if constexpr (J + 1 < XBA_TILE) {
            // lanes up to J change only entries of a that are not used
            big_rank1x2<T, J, J + 1>(a, v, (lane > J) ? a[J] : (T)0, xj);
            BigPanelStep<T, J + 1>::step(a, v, lane, ok);
        }
    }
};

// grid (ceil((max nb - k) / 4), nbig), 256 threads.  Wavefront w of workgroup x takes task t = 4 x + w of the
// window's nb - k tasks: t = 0 the inverse of the diagonal tile's factor L(k,k) (right-hand side: the identity rows,
// so lane j solves column j of L(k,k)^-1) into the side buffer, t >= 1 the tile row i = k + t: L(i,k) = A(i,k)
// L(k,k)^-T.  Every wavefront factors the diagonal tile itself, in registers, fused with its solve: no LDS and no
// barriers.  The wavefronts' factors are identical (the same instructions on the same values).
template <typename T>
__global__ __launch_bounds__(256) void k_big_panel(DArgs D, int k) {
    __shared__ int sok;
    const BigWin B = D.bigs[blockIdx.y];
    const int ntask = B.nb - k;
    if ((int)blockIdx.x * 4 >= ntask) return;  // (uniform per workgroup)
    double *ws = D.wstate + (size_t)NWS * B.win;
    const int tid = threadIdx.x, lane = tid & 63, t = 4 * (int)blockIdx.x + (tid >> 6), np = B.np;
    // the flag is read once per workgroup: another workgroup of this launch may clear it meanwhile
    if (tid == 0) sok = ws[WS_BIGOK] != 0.0;
    __syncthreads();
    if (!sok || t >= ntask) return;
    T *F = big_base<T>(D) + B.f0;
    T a[XBA_TILE], v[XBA_TILE];
    big_row_load(F + (size_t)(k * XBA_TILE + lane) * np + k * XBA_TILE, a);
#pragma unroll
    for (int m = 0; m < XBA_TILE; m++) a[m] = (m <= lane) ? a[m] : (T)0;
    T *Xr = F + (size_t)((k + t) * XBA_TILE + lane) * np + k * XBA_TILE;  // (t >= 1)
    if (t == 0) {
#pragma unroll
        for (int m = 0; m < XBA_TILE; m++) v[m] = (m == lane) ? (T)1 : (T)0;
    } else {
        big_row_load(Xr, v);
    }
    int ok = 1;
    BigPanelStep<T, 0>::step(a, v, lane, ok);
    if (!ok) {
        if (lane == 0) ws[WS_BIGOK] = 0.0;
        return;
    }
    if (t == 0) {
        // lane j holds column j of L(k,k)^-1: row m of the side buffer (row-major, 64 wide) gets entry j from lane j
        T *Dc = F + (size_t)np * np + (size_t)(k * XBA_TILE) * XBA_TILE + lane;
#pragma unroll
        for (int m = 0; m < XBA_TILE; m++) Dc[(size_t)m * XBA_TILE] = v[m];
    } else {
        big_row_store(Xr, v);
    }
}

// -------------------------------------------------------------------------------------------- k_big_update --
// grid (m (m + 1) / 2 for m = max nb - k - 1, nbig): workgroup x updates the trailing lower tile (i, j),
// i = k + 1 + ti, j = k + 1 + tj, x = ti (ti + 1) / 2 + tj.  Wavefront w computes the 32 x 32 quadrant
// (rows 32 (w / 2), cols 32 (w % 2)) as 2 x 2 MFMA blocks of 16 x 16, over the tile's k-range staged in LDS whole
// (both tiles read 16 bytes at a time, every load of the workgroup in flight together).
template <typename T>
__global__ __launch_bounds__(256) void k_big_update(DArgs D, int k) {
    constexpr int LD = XBA_TILE + (sizeof(T) == 4 ? 4 : 2), W = 16 / sizeof(T);  // (rows stay 16-byte aligned)
    typedef T VT __attribute__((ext_vector_type(16 / sizeof(T))));
    __shared__ __align__(16) T sI[XBA_TILE][LD];
    __shared__ __align__(16) T sJ[XBA_TILE][LD];
    typedef typename Mfma<T>::V V;
    const BigWin B = D.bigs[blockIdx.y];
    const int m = B.nb - k - 1, x = blockIdx.x;
    int ti = (int)((sqrtf(8.0f * (float)x + 1.0f) - 1.0f) * 0.5f);
    while ((ti + 1) * (ti + 2) / 2 <= x) ti++;
    while (ti * (ti + 1) / 2 > x) ti--;
    const int tj = x - ti * (ti + 1) / 2;
    if (ti >= m) return;
    __shared__ int sgo;
    if (threadIdx.x == 0) sgo = D.wstate[(size_t)NWS * B.win + WS_BIGOK] != 0.0;
    __syncthreads();
    if (!sgo) return;
    const int i = k + 1 + ti, j = k + 1 + tj, np = B.np;
    T *F = big_base<T>(D) + B.f0;
    const T *Li = F + (size_t)(i * XBA_TILE) * np + k * XBA_TILE;
    const T *Lj = F + (size_t)(j * XBA_TILE) * np + k * XBA_TILE;
    T *C = F + (size_t)(i * XBA_TILE) * np + j * XBA_TILE;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int r0 = 32 * (wv >> 1), c0 = 32 * (wv & 1), lr = lane & 15, lk = lane >> 4;
    V acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++)
#pragma unroll
            for (int q = 0; q < 4; q++)
                acc[a][b][q] = C[(size_t)(r0 + 16 * a + Mfma<T>::row(lk, q)) * np + c0 + 16 * b + lr];
    {
        constexpr int RW = XBA_TILE / W, NV = XBA_TILE * RW / 256;  // vectors per row; per thread and tile
        VT va[NV], vb[NV];
#pragma unroll
        for (int u = 0; u < NV; u++) {
            const int e = tid + 256 * u, r = e / RW, c = (e % RW) * W;
            va[u] = *(const VT *)(Li + (size_t)r * np + c);
            vb[u] = *(const VT *)(Lj + (size_t)r * np + c);
        }
#pragma unroll
        for (int u = 0; u < NV; u++) {
            const int e = tid + 256 * u, r = e / RW, c = (e % RW) * W;
            *(VT *)&sI[r][c] = va[u];
            *(VT *)&sJ[r][c] = vb[u];
        }
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < XBA_TILE / 4; kk++) {
        T av[2], bv[2];
#pragma unroll
        for (int a = 0; a < 2; a++) av[a] = -sI[r0 + 16 * a + lr][4 * kk + lk];
#pragma unroll
        for (int b = 0; b < 2; b++) bv[b] = sJ[c0 + 16 * b + lr][4 * kk + lk];
#pragma unroll
        for (int a = 0; a < 2; a++)
#pragma unroll
            for (int b = 0; b < 2; b++) acc[a][b] = Mfma<T>::fma(av[a], bv[b], acc[a][b]);
    }
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++)
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const int row = r0 + 16 * a + Mfma<T>::row(lk, q), col = c0 + 16 * b + lr;
                if (i != j || row >= col) C[(size_t)row * np + col] = acc[a][b][q];
            }
}

// ---------------------------------------------------------------------------------------------- k_big_trsv --
// grid (nbig), 1024 threads.  Refinement step `it`: it = 0 solves for the right-hand side b from x = 0; a later step
// first tests the residual of the previous one (|r| <= 1e-15 |b|: converged, nothing more to do), then solves for r.
// The solve runs in the factor's precision on s = scale * rhs and adds scale * y to x (FP64).
// (sum over each aligned group of 16 lanes, by DPP row rotations: every lane of the group ends with its total)
template <int CTRL>
__device__ inline float big_ror(float v) {
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), CTRL, 0xf, 0xf, false));
}
template <int CTRL>
__device__ inline double big_ror(double v) {
    const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
    const unsigned lo = (unsigned)__builtin_amdgcn_update_dpp(0, (int)(unsigned)u, CTRL, 0xf, 0xf, false);
    const unsigned hi = (unsigned)__builtin_amdgcn_update_dpp(0, (int)(unsigned)(u >> 32), CTRL, 0xf, 0xf, false);
    return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}
template <typename T>
__device__ inline T big_red16(T v) {
    v += big_ror<0x128>(v);  // row_ror:8
    v += big_ror<0x124>(v);  // row_ror:4
    v += big_ror<0x122>(v);  // row_ror:2
    v += big_ror<0x121>(v);  // row_ror:1
    return v;
}

#define TRSV_BATCH 8  // loads a thread of the triangular solve issues together
template <typename T>
__global__ __launch_bounds__(1024) void k_big_trsv(DArgs D, int it) {
    __shared__ __align__(16) T sv[XBA_MAX_STATE];
    __shared__ T st[XBA_TILE];
    __shared__ T spart[16][XBA_TILE];
    __shared__ double sred[32];
    __shared__ int sgo;
    const BigWin B = D.bigs[blockIdx.x];
    const WinInfo W = D.wins[B.win];
    const int n = B.n, np = B.np, nb = B.nb, tid = threadIdx.x, nt = blockDim.x;
    double *ws = D.wstate + (size_t)NWS * B.win;
    double *x = D.bigV + B.v0, *r = x + np, *pm = r + np;
    const double *ss = D.bigS + B.s0, *bv = D.b + W.st0;
    if (tid == 0) sgo = ws[WS_BIGOK] != 0.0 && ws[WS_BIGCONV] == 0.0;
    __syncthreads();
    if (!sgo) return;
    if (it > 0) {
        const int nres = (n + XBA_RES_ROWS - 1) / XBA_RES_ROWS;
        double rn = 0.0, bn = 0.0;
        for (int q = tid; q < nres; q += nt) {
            rn = fmax(rn, pm[2 * q]);
            bn = fmax(bn, pm[2 * q + 1]);
        }
        rn = dwave_max(rn);
        bn = dwave_max(bn);
        if ((tid & 63) == 0) {
            sred[2 * (tid >> 6)] = rn;
            sred[2 * (tid >> 6) + 1] = bn;
        }
        __syncthreads();
        if (tid == 0) {
            double a = 0.0, c = 0.0;
            for (int q = 0; q < nt / 64; q++) {
                a = fmax(a, sred[2 * q]);
                c = fmax(c, sred[2 * q + 1]);
            }
            sgo = !(a <= 1e-15 * c);
            if (!sgo) ws[WS_BIGCONV] = 1.0;
        }
        __syncthreads();
        if (!sgo) return;
    }
    for (int i = tid; i < np; i += nt) {
        const double rhs = i < n ? (it == 0 ? bv[i] : r[i]) : 0.0;
        sv[i] = (T)(rhs * ss[i]);
    }
    __syncthreads();
    const T *F = big_base<T>(D) + B.f0, *Dinv = F + (size_t)np * np;
    // forward: y_I = L(I,I)^-1 (s_I - sum_{J<I} L(I,J) y_J), 16 threads per row, each reading 16 bytes of every
    // 16-vector-wide column chunk of the row, TRSV_BATCH chunks' loads issued together.  Each stage's loads of its
    // diagonal-tile inverse row are issued at its start, and the next stage's first batch of row chunks before its
    // barriers, so that they arrive while the stage waits.
    {
        constexpr int V = 16 / sizeof(T), CW = 16 * V;
        typedef T VT __attribute__((ext_vector_type(V)));
        const int i = tid >> 4, q = tid & 15;
        VT rv[TRSV_BATCH];  // the current stage's first batch
        for (int I = 0; I < nb; I++) {
            const int nch = I * XBA_TILE / CW;
            const T *Dr = Dinv + (size_t)(I * XBA_TILE + i) * XBA_TILE;
            T dv[XBA_TILE / 16];
#pragma unroll
            for (int u = 0; u < XBA_TILE / 16; u++) dv[u] = Dr[q + 16 * u];
            const T *row = F + (size_t)(I * XBA_TILE + i) * np;
            T acc = (T)0;
            for (int k0 = 0; k0 < nch; k0 += TRSV_BATCH) {
                if (k0 > 0) {
#pragma unroll
                    for (int u = 0; u < TRSV_BATCH; u++)
                        if (k0 + u < nch) rv[u] = *(const VT *)(row + (k0 + u) * CW + V * q);
                }
#pragma unroll
                for (int u = 0; u < TRSV_BATCH; u++)
                    if (k0 + u < nch) {
                        const VT sx = *(const VT *)(sv + (k0 + u) * CW + V * q);
#pragma unroll
                        for (int e = 0; e < V; e++) acc += rv[u][e] * sx[e];
                    }
            }
            if (I + 1 < nb) {
                const T *nrow = F + (size_t)((I + 1) * XBA_TILE + i) * np;
                const int nn = (I + 1) * XBA_TILE / CW;
#pragma unroll
                for (int u = 0; u < TRSV_BATCH; u++)
                    if (u < nn) rv[u] = *(const VT *)(nrow + u * CW + V * q);
            }
            acc = big_red16(acc);
            if (q == 0) st[i] = sv[I * XBA_TILE + i] - acc;
            __syncthreads();
            T y = (T)0;
#pragma unroll
            for (int u = 0; u < XBA_TILE / 16; u++) {
                const int c = q + 16 * u;
                if (c <= i) y += dv[u] * st[c];
            }
            y = big_red16(y);
            if (q == 0) sv[I * XBA_TILE + i] = y;
            __syncthreads();
        }
    }
    // backward: x_I = L(I,I)^-T (y_I - sum_{J>I} L(J,I)^T x_J), a column per lane, 16 row groups, TRSV_BATCH rows'
    // loads issued together; the diagonal-tile inverse entries at the stage's start and the next stage's first
    // batch before its barriers, as above
    {
        const int c = tid & 63, g = tid >> 6;
        T fv[TRSV_BATCH];  // the current stage's first batch
        {
            const int I = nb - 1;
#pragma unroll
            for (int u = 0; u < TRSV_BATCH; u++) {
                const int rr = (I + 1) * XBA_TILE + g + 16 * u;
                fv[u] = rr < np ? F[(size_t)rr * np + I * XBA_TILE + c] : (T)0;
            }
        }
        for (int I = nb - 1; I >= 0; I--) {
            T dv[XBA_TILE / 16];
#pragma unroll
            for (int u = 0; u < XBA_TILE / 16; u++) dv[u] = Dinv[(size_t)(I * XBA_TILE + g + 16 * u) * XBA_TILE + c];
            T acc = (T)0;
            for (int r0 = (I + 1) * XBA_TILE + g; r0 < np; r0 += 16 * TRSV_BATCH) {
                if (r0 > (I + 1) * XBA_TILE + g) {
#pragma unroll
                    for (int u = 0; u < TRSV_BATCH; u++) {
                        const int rr = r0 + 16 * u;
                        fv[u] = rr < np ? F[(size_t)rr * np + I * XBA_TILE + c] : (T)0;
                    }
                }
#pragma unroll
                for (int u = 0; u < TRSV_BATCH; u++) {
                    const int rr = r0 + 16 * u;
                    if (rr < np) acc += fv[u] * sv[rr];
                }
            }
            if (I > 0) {
#pragma unroll
                for (int u = 0; u < TRSV_BATCH; u++) {
                    const int rr = I * XBA_TILE + g + 16 * u;
                    fv[u] = rr < np ? F[(size_t)rr * np + (I - 1) * XBA_TILE + c] : (T)0;
                }
            }
            spart[g][c] = acc;
            __syncthreads();
            if (g == 0) {
                T t = sv[I * XBA_TILE + c];
                for (int q = 0; q < 16; q++) t -= spart[q][c];
                st[c] = t;
            }
            __syncthreads();
            T xv = (T)0;
#pragma unroll
            for (int u = 0; u < XBA_TILE / 16; u++) {
                const int i = g + 16 * u;
                if (i >= c) xv += dv[u] * st[i];
            }
            spart[g][c] = xv;
            __syncthreads();
            if (g == 0) {
                T t = (T)0;
                for (int q = 0; q < 16; q++) t += spart[q][c];
                sv[I * XBA_TILE + c] = t;
            }
            __syncthreads();
        }
    }
    double cmax = 0.0, xmax = 0.0;
    for (int i = tid; i < n; i += nt) {
        const double c = (double)sv[i] * ss[i], xi = (it == 0 ? 0.0 : x[i]) + c;
        x[i] = xi;
        cmax = fmax(cmax, fabs(c));
        xmax = fmax(xmax, fabs(xi));
    }
    if (sizeof(T) == 4) {  // refinement: converged once the correction is small against the solution
        cmax = dwave_max(cmax);
        xmax = dwave_max(xmax);
        __syncthreads();  // sred is free again (its last readers were before the solves)
        if ((tid & 63) == 0) {
            sred[2 * (tid >> 6)] = cmax;
            sred[2 * (tid >> 6) + 1] = xmax;
        }
        __syncthreads();
        if (tid == 0) {
            double a = 0.0, c = 0.0;
            for (int q = 0; q < nt / 64; q++) {
                a = fmax(a, sred[2 * q]);
                c = fmax(c, sred[2 * q + 1]);
            }
            if (a <= XBA_REFINE_TOL * c) ws[WS_BIGCONV] = 1.0;
        }
    }
    if (tid == 0) ws[WS_NREF] = it + 1;
}

// --------------------------------------------------------------------------------------------- k_big_resid --
// grid (ceil(max n / XBA_RES_ROWS), nbig), 256 threads: one row per wavefront, r_i = b_i - (Hd x)_i in FP64, and
// the workgroup's max |r_i| and max |b_i|.
__global__ __launch_bounds__(256) void k_big_resid(DArgs D) {
    __shared__ double sm[8];
    const BigWin B = D.bigs[blockIdx.y];
    const int n = B.n, np = B.np, lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if ((int)blockIdx.x * XBA_RES_ROWS >= n) return;
    const double *ws = D.wstate + (size_t)NWS * B.win;
    if (ws[WS_BIGOK] == 0.0 || ws[WS_BIGCONV] != 0.0) return;  // constant during this launch
    const WinInfo W = D.wins[B.win];
    const double lam = ws[WS_LAM];
    const double *H = D.H + W.h0, *bv = D.b + W.st0;
    double *x = D.bigV + B.v0, *r = x + np, *pm = r + np;
    const int i = blockIdx.x * XBA_RES_ROWS + wv;
    double ra = 0.0, ba = 0.0;
    if (i < n) {
        const double *Hi = H + (size_t)i * n;
        double acc = 0.0;
        for (int j = lane; j < n; j += 64) acc += Hi[j] * x[j];
        acc = dwave_sum(acc);
        const double ri = bv[i] - acc - (lam * Hi[i] + 1e-9) * x[i];
        if (lane == 0) r[i] = ri;
        ra = fabs(ri);
        ba = fabs(bv[i]);
    }
    if (lane == 0) {
        sm[2 * wv] = ra;
        sm[2 * wv + 1] = ba;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double a = 0.0, c = 0.0;
        for (int q = 0; q < XBA_RES_ROWS; q++) {
            a = fmax(a, sm[2 * q]);
            c = fmax(c, sm[2 * q + 1]);
        }
        pm[2 * blockIdx.x] = a;
        pm[2 * blockIdx.x + 1] = c;
    }
}

// -------------------------------------------------------------------------------------------- k_big_finish --
// grid (nbig), 256 threads: the update dx = -x and the trial state (NaN update when the factorisation failed).
__global__ __launch_bounds__(256) void k_big_finish(DArgs D) {
    const BigWin B = D.bigs[blockIdx.x];
    const WinInfo W = D.wins[B.win];
    const int n = B.n, tid = threadIdx.x, nt = blockDim.x;
    double *ws = D.wstate + (size_t)NWS * B.win;
    const bool ok = ws[WS_BIGOK] != 0.0;
    // mixed: converged by the residual or the correction test; an FP64 factor (one solve, no refinement): always
    const bool conv = !D.fallback_on || D.big64 || ws[WS_BIGCONV] != 0.0;
    const bool fb = D.fallback_on && !(ok && conv);  // left to the FP64 fallback (k_solve<double>, fallback mode)
    const double *x = D.bigV + B.v0;
    if (!fb) {
        for (int i = tid; i < n; i += nt) D.dx[W.st0 + i] = ok ? -x[i] : NAN;
        __syncthreads();
        if (ok) trial_state(D, W, tid, nt);
    }
    if (tid == 0) {
        ws[WS_CHOLOK] = ok ? 1.0 : 0.0;
        if (!ok) ws[WS_NREF] = 0.0;
        ws[WS_CONV] = (ok && conv) ? 1.0 : 0.0;
        ws[WS_FALLBACK] = fb ? 1.0 : 0.0;
        if (fb) ws[WS_NFALLBACK] += 1.0;
    }
}

template __global__ void k_big_prep<float>(DArgs, int);
template __global__ void k_big_prep<double>(DArgs, int);
template __global__ void k_big_panel<float>(DArgs, int);
template __global__ void k_big_panel<double>(DArgs, int);
template __global__ void k_big_update<float>(DArgs, int);
template __global__ void k_big_update<double>(DArgs, int);
template __global__ void k_big_trsv<float>(DArgs, int);
template __global__ void k_big_trsv<double>(DArgs, int);

}  // namespace xba
