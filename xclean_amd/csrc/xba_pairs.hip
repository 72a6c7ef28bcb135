// Ordered observation pairs inside landmark segments (backend/ba.py:902-927 build_pairs), on the device.
//
// Input: the landmark index of every observation, sorted ascending (the observation list grouped by landmark).
// Output: offs[l] = first observation of landmark l (nl + 1 entries), and for every landmark with c observations
// its c^2 ordered pairs (a, b), a-major then b, landmarks in ascending order, as two int32 lists.
//
// offs comes from a binary search per landmark (the list is sorted, so no counting and no atomics); the pair
// offsets base[l] = sum_{m<l} c_m^2 from a three-kernel int64 scan; the expansion is one thread per pair, which
// finds its landmark by a binary search over base and writes both lists coalesced.
#pragma once
#include "xba_internal.h"

#define PAIRS_BS 1024

// offs[l] = lower_bound(olm, l) for l in [0, nl]; every observation's landmark in range and non-decreasing (err 1
// otherwise).
__global__ void k_pairs_offs(const int32_t *olm, int64_t n, int32_t nl, int32_t *offs, int32_t *err) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n || i <= nl; i += stride) {
        if (i < n) {
            const int32_t v = olm[i];
            if (v < 0 || v >= nl || (i + 1 < n && olm[i + 1] < v)) atomicOr(err, 1);
        }
        if (i <= nl) {
            int64_t lo = 0, hi = n;
            while (lo < hi) {
                const int64_t mid = (lo + hi) >> 1;
                if (olm[mid] < (int32_t)i) lo = mid + 1;
                else hi = mid;
            }
            offs[i] = (int32_t)lo;
        }
    }
}

// Inclusive scan of one int64 per thread over a block of PAIRS_BS threads (wavefront shuffles, then the wave
// totals).
__device__ inline int64_t pairs_block_scan(int64_t v, int64_t *swave) {
    const int lane = threadIdx.x & (XBA_WAVE - 1), w = threadIdx.x / XBA_WAVE;
    for (int d = 1; d < XBA_WAVE; d <<= 1) {
        const int64_t u = __shfl_up(v, d, XBA_WAVE);
        if (lane >= d) v += u;
    }
    if (lane == XBA_WAVE - 1) swave[w] = v;
    __syncthreads();
    if (w == 0) {
        const int nw = blockDim.x / XBA_WAVE;
        int64_t t = lane < nw ? swave[lane] : 0;
        for (int d = 1; d < XBA_WAVE; d <<= 1) {
            const int64_t u = __shfl_up(t, d, XBA_WAVE);
            if (lane >= d) t += u;
        }
        if (lane < nw) swave[lane] = t;
    }
    __syncthreads();
    if (w > 0) v += swave[w - 1];
    return v;
}

__device__ inline int64_t pairs_count2(const int32_t *offs, int32_t l) {
    const int64_t c = offs[l + 1] - offs[l];
    return c * c;
}

// bsum[b] = sum of c_l^2 over the landmarks of block b.
__global__ __launch_bounds__(PAIRS_BS) void k_pairs_bsum(const int32_t *offs, int32_t nl, int64_t *bsum,
                                                         const int32_t *err) {
    __shared__ int64_t swave[PAIRS_BS / XBA_WAVE];
    if (*err) return;
    const int32_t l = blockIdx.x * PAIRS_BS + threadIdx.x;
    const int64_t s = pairs_block_scan(l < nl ? pairs_count2(offs, l) : 0, swave);
    if (threadIdx.x == PAIRS_BS - 1) bsum[blockIdx.x] = s;
}

// Exclusive scan of the block sums in place (one block; nb entries), the total at bsum[nb].
__global__ __launch_bounds__(PAIRS_BS) void k_pairs_bscan(int64_t *bsum, int32_t nb, const int32_t *err) {
    __shared__ int64_t swave[PAIRS_BS / XBA_WAVE];
    __shared__ int64_t carry;
    if (*err) return;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (int32_t b0 = 0; b0 < nb; b0 += PAIRS_BS) {
        const int32_t b = b0 + threadIdx.x;
        const int64_t v = b < nb ? bsum[b] : 0;
        const int64_t inc = pairs_block_scan(v, swave);
        const int64_t c0 = carry;
        if (b < nb) bsum[b] = c0 + inc - v;
        __syncthreads();
        if (threadIdx.x == PAIRS_BS - 1) carry = c0 + inc;
        __syncthreads();
    }
    if (threadIdx.x == 0) bsum[nb] = carry;
}

// base[l] = sum_{m<l} c_m^2 for l in [0, nl], base[nl] = the pair count.
__global__ __launch_bounds__(PAIRS_BS) void k_pairs_base(const int32_t *offs, int32_t nl, const int64_t *bsum,
                                                         int64_t *base, const int32_t *err) {
    __shared__ int64_t swave[PAIRS_BS / XBA_WAVE];
    if (*err) return;
    const int32_t l = blockIdx.x * PAIRS_BS + threadIdx.x;
    const int64_t v = l < nl ? pairs_count2(offs, l) : 0;
    const int64_t inc = pairs_block_scan(v, swave);
    if (l < nl) base[l] = bsum[blockIdx.x] + inc - v;
    if (l == nl) base[nl] = bsum[gridDim.x];   // the grid covers l == nl; bsum[gridDim.x] is the total
}

// One thread per pair t: its landmark l (the last l with base[l] <= t, a landmark with c > 0), then a and b.
__global__ void k_pairs_expand(const int32_t *offs, const int64_t *base, int32_t nl, int64_t npair, int32_t *pa,
                               int32_t *pb) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < npair; t += stride) {
        int32_t lo = 0, hi = nl;   // invariant: base[lo] <= t < base[hi]
        while (hi - lo > 1) {
            const int32_t mid = (lo + hi) >> 1;
            if (base[mid] <= t) lo = mid;
            else hi = mid;
        }
        const int64_t loc = t - base[lo];
        const int32_t o = offs[lo], c = offs[lo + 1] - o;
        pa[t] = o + (int32_t)(loc / c);
        pb[t] = o + (int32_t)(loc % c);
    }
}
