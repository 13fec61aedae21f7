// Exclusive prefix sums in three launches: block totals, a one-workgroup scan of the totals, then
// each block's items scanned in LDS and offset. No workgroup waits on another (a decoupled look-back
// scan does, and beside a persistent kernel that holds most of the device its waits stretch).
#pragma once
#include "fmmhip_internal.hpp"

namespace fmmhip {

constexpr int SCAN_BLOCK = 256;
constexpr int SCAN_ITEMS = 8;   // items per thread: a block covers 2048 items

// exclusive scan of one value per thread of a 256-thread block; sh holds 4 entries; total of the block
template <class O>
__device__ inline O block_excl_scan(O v, O& total, O* sh) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    O x = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const O y = __shfl_up(x, d, 64);
        if (lane >= d) x += y;
    }
    if (lane == 63) sh[w] = x;
    __syncthreads();
    O before = 0;
    for (int k = 0; k < w; ++k) before += sh[k];
    total = sh[0] + sh[1] + sh[2] + sh[3];
    __syncthreads();
    return before + x - v;
}

template <class T, class O>
__global__ __launch_bounds__(SCAN_BLOCK) void k_scan_totals(const T* __restrict__ in, int64_t n, O* __restrict__ tot) {
    __shared__ O sh[4];
    const int64_t base = (int64_t)blockIdx.x * SCAN_BLOCK * SCAN_ITEMS;
    O s = 0;
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; ++k) {
        const int64_t i = base + (int64_t)k * SCAN_BLOCK + threadIdx.x;
        if (i < n) s += (O)in[i];
    }
    O total;
    (void)block_excl_scan<O>(s, total, sh);
    if (threadIdx.x == 0) tot[blockIdx.x] = total;
}

// exclusive scan of nb block totals in place, by one workgroup; tot[nb] = the sum of all
template <class O>
__global__ __launch_bounds__(SCAN_BLOCK) void k_scan_offsets(O* __restrict__ tot, int64_t nb) {
    __shared__ O sh[4];
    O carry = 0;
    for (int64_t b0 = 0; b0 < nb; b0 += SCAN_BLOCK) {
        const int64_t b = b0 + threadIdx.x;
        const O v = b < nb ? tot[b] : O(0);
        O total;
        const O e = block_excl_scan<O>(v, total, sh);
        if (b < nb) tot[b] = carry + e;
        carry += total;
    }
    if (threadIdx.x == 0) tot[nb] = carry;
}

template <class T, class O>
__global__ __launch_bounds__(SCAN_BLOCK) void k_scan_apply(const T* __restrict__ in, int64_t n, const O* __restrict__ tot,
                                                           O* __restrict__ out) {
    __shared__ O sh[4];
    const int64_t base = (int64_t)blockIdx.x * SCAN_BLOCK * SCAN_ITEMS;
    O carry = tot[blockIdx.x];
    for (int k = 0; k < SCAN_ITEMS; ++k) {
        const int64_t i = base + (int64_t)k * SCAN_BLOCK + threadIdx.x;
        const O v = i < n ? (O)in[i] : O(0);
        O total;
        const O e = block_excl_scan<O>(v, total, sh);
        if (i < n) out[i] = carry + e;
        carry += total;
    }
}

// out[i] = in[0] + ... + in[i - 1] for i < n, on the context's stream; the context's temp buffer holds the
// block totals (so a scan's result must be consumed before the next scan on the same context)
template <class T, class O>
inline void excl_scan3(Context& C, const T* in, O* out, int64_t n) {
    if (n <= 0) return;
    const int64_t per = (int64_t)SCAN_BLOCK * SCAN_ITEMS;
    const int64_t nb = (n + per - 1) / per;
    C.temp.resize((size_t)(nb + 1) * sizeof(O) + 16);
    O* tot = (O*)C.temp.data();
    k_scan_totals<T, O><<<(unsigned)nb, SCAN_BLOCK, 0, C.stream>>>(in, n, tot);
    LAUNCH_CHECK();
    k_scan_offsets<O><<<1, SCAN_BLOCK, 0, C.stream>>>(tot, nb);
    LAUNCH_CHECK();
    k_scan_apply<T, O><<<(unsigned)nb, SCAN_BLOCK, 0, C.stream>>>(in, n, tot, out);
    LAUNCH_CHECK();
}

}  // namespace fmmhip
