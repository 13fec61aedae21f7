// Stable LSD radix sort of (key, value) pairs or keys alone, 8-bit digits, one pass per digit in three
// steps: per-workgroup digit counts, an exclusive scan of the counts laid out digit-major (scan3.hpp, so the
// scan gives every (digit, workgroup) its output offset directly), and a scatter that ranks each tile's
// keys in LDS. The scatter walks its 2048-key tile in 8 rounds of 256 consecutive keys; a key's rank among
// equal digits inside its wavefront comes from 8 ballots (one per digit bit), its rank across the
// wavefronts of the round from per-wavefront digit counts in LDS, and the rounds add up in a running count
// per digit, so equal digits keep their input order. No workgroup waits on another.
//
// Segmented sorts (each range seg[s] .. seg[s + 1] - 1 sorted on its own) add passes over the segment
// index of each pair's value after the key passes, most significant last. That equals a per-range sort
// when every value lies in the range of its pair's input position, which holds for the callers here: the
// values are input positions, or a permutation that an earlier segmented sort kept inside the ranges.
#pragma once
#include "fmmhip_internal.hpp"
#include "scan3.hpp"
#include <vector>

namespace fmmhip {

constexpr int RS_BLOCK = 256;
constexpr int RS_ROUNDS = 8;
constexpr int RS_TILE = RS_BLOCK * RS_ROUNDS;   // keys per workgroup

// keys as unsigned integers in their sort order (floats: the sign bit flipped for positive values, all
// bits for negative ones)
template <class K> struct RadixBits;
template <> struct RadixBits<uint32_t> {
    static __device__ inline uint32_t get(uint32_t k) { return k; }
};
template <> struct RadixBits<uint64_t> {
    static __device__ inline uint64_t get(uint64_t k) { return k; }
};
template <> struct RadixBits<float> {
    static __device__ inline uint32_t get(float k) {
        const uint32_t u = __float_as_uint(k);
        return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
    }
};

struct RadixPass {
    int seg;     // 1: a digit of the value's segment index; 0: a digit of the key
    int shift;   // lowest bit of the digit
    int width;   // bits of the digit, 1..8
    int desc;    // key digits in descending order
};

template <class K>
__device__ inline unsigned radix_digit(K k, uint32_t v, RadixPass P, const int64_t* __restrict__ seg, int n_seg) {
    const unsigned mask = (1u << P.width) - 1u;
    if (P.seg) {   // the segment holding value v: the largest s with seg[s] <= v
        int lo = 0, hi = n_seg - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (seg[mid] <= (int64_t)v) lo = mid;
            else hi = mid - 1;
        }
        return ((unsigned)lo >> P.shift) & mask;
    }
    const unsigned d = (unsigned)(RadixBits<K>::get(k) >> P.shift) & mask;
    return P.desc ? mask - d : d;
}

// digit counts of each tile, stored digit-major: table[d * nblk + tile]
template <class K, bool HAS_V>
__global__ __launch_bounds__(RS_BLOCK) void k_radix_count(const K* __restrict__ k, const uint32_t* __restrict__ v, int64_t n,
                                                          RadixPass P, const int64_t* __restrict__ seg, int n_seg,
                                                          int32_t* __restrict__ table, int64_t nblk) {
    __shared__ int32_t h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    const int64_t base = (int64_t)blockIdx.x * RS_TILE;
    for (int r = 0; r < RS_ROUNDS; ++r) {
        const int64_t i = base + (int64_t)r * RS_BLOCK + threadIdx.x;
        if (i < n) atomicAdd(&h[radix_digit<K>(k[i], HAS_V ? v[i] : 0u, P, seg, n_seg)], 1);
    }
    __syncthreads();
    table[(int64_t)threadIdx.x * nblk + blockIdx.x] = h[threadIdx.x];
}

// each tile's pairs to off[d * nblk + tile] + their stable rank among the tile's pairs of digit d
template <class K, bool HAS_V>
__global__ __launch_bounds__(RS_BLOCK) void k_radix_scatter(const K* __restrict__ kin, const uint32_t* __restrict__ vin, int64_t n,
                                                            RadixPass P, const int64_t* __restrict__ seg, int n_seg,
                                                            const int32_t* __restrict__ off, int64_t nblk, K* __restrict__ kout,
                                                            uint32_t* __restrict__ vout) {
    __shared__ int32_t cnt[RS_BLOCK / 64][256];   // per wavefront: digit counts, then their prefix over wavefronts
    __shared__ int32_t run[256];                 // next output position of each digit
    __shared__ int32_t tot[256];                 // keys of each digit in the round
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    run[tid] = off[(int64_t)tid * nblk + blockIdx.x];
#pragma unroll
    for (int q = 0; q < RS_BLOCK / 64; ++q) cnt[q][tid] = 0;
    __syncthreads();
    const unsigned long long below = (1ull << lane) - 1ull;
    const int64_t base = (int64_t)blockIdx.x * RS_TILE;
    for (int r = 0; r < RS_ROUNDS; ++r) {
        const int64_t i = base + (int64_t)r * RS_BLOCK + tid;
        const bool valid = i < n;
        K key = K();
        uint32_t val = 0u;
        unsigned d = 0u;
        if (valid) {
            key = kin[i];
            if (HAS_V) val = vin[i];
            d = radix_digit<K>(key, val, P, seg, n_seg);
        }
        unsigned long long peers = __ballot(valid);   // lanes of the wavefront with this lane's digit
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const bool bit = (d >> b) & 1u;
            const unsigned long long m = __ballot(bit);
            peers &= bit ? m : ~m;
        }
        const int rank = __popcll(peers & below);
        if (valid && rank == 0) cnt[w][d] = __popcll(peers);
        __syncthreads();
        {
            int s = 0;
#pragma unroll
            for (int q = 0; q < RS_BLOCK / 64; ++q) {
                const int c = cnt[q][tid];
                cnt[q][tid] = s;
                s += c;
            }
            tot[tid] = s;
        }
        __syncthreads();
        if (valid) {
            const int64_t pos = (int64_t)run[d] + cnt[w][d] + rank;
            kout[pos] = key;
            if (HAS_V) vout[pos] = val;
        }
        __syncthreads();
        run[tid] += tot[tid];
#pragma unroll
        for (int q = 0; q < RS_BLOCK / 64; ++q) cnt[q][tid] = 0;
        __syncthreads();
    }
}

// Every pass of a sort of at most RS_TILE pairs in one workgroup, the pairs in LDS between passes (the
// same digits and stable ranking as the passes above, the tile's own digit counts giving the offsets).
constexpr int RS_MAX_PASSES = 16;
struct RadixPasses {
    RadixPass p[RS_MAX_PASSES];
    int n;
};

template <class K, bool HAS_V>
__global__ __launch_bounds__(RS_BLOCK) void k_radix_small(const K* __restrict__ kin, const uint32_t* __restrict__ vin, int n,
                                                          RadixPasses passes, const int64_t* __restrict__ seg, int n_seg,
                                                          K* __restrict__ kout, uint32_t* __restrict__ vout) {
    __shared__ K ks[2][RS_TILE];
    __shared__ uint32_t vs[HAS_V ? 2 : 1][HAS_V ? RS_TILE : 1];
    __shared__ int32_t cnt[RS_BLOCK / 64][256];
    __shared__ int32_t run[256];
    __shared__ int32_t tot[256];
    __shared__ int64_t sh[4];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const unsigned long long below = (1ull << lane) - 1ull;
    for (int i = tid; i < n; i += RS_BLOCK) {
        ks[0][i] = kin[i];
        if (HAS_V) vs[0][i] = vin[i];
    }
    int cur = 0;
    for (int pi = 0; pi < passes.n; ++pi) {
        const RadixPass P = passes.p[pi];
        tot[tid] = 0;
#pragma unroll
        for (int q = 0; q < RS_BLOCK / 64; ++q) cnt[q][tid] = 0;
        __syncthreads();
        for (int i = tid; i < n; i += RS_BLOCK)
            atomicAdd(&tot[radix_digit<K>(ks[cur][i], HAS_V ? vs[cur][i] : 0u, P, seg, n_seg)], 1);
        __syncthreads();
        {   // each digit's first output position
            int64_t total;
            run[tid] = (int32_t)block_excl_scan<int64_t>((int64_t)tot[tid], total, sh);
        }
        __syncthreads();
        for (int r = 0; r < RS_ROUNDS; ++r) {
            const int i = r * RS_BLOCK + tid;
            const bool valid = i < n;
            K key = K();
            uint32_t val = 0u;
            unsigned d = 0u;
            if (valid) {
                key = ks[cur][i];
                if (HAS_V) val = vs[cur][i];
                d = radix_digit<K>(key, val, P, seg, n_seg);
            }
            unsigned long long peers = __ballot(valid);
#pragma unroll
            for (int b = 0; b < 8; ++b) {
                const bool bit = (d >> b) & 1u;
                const unsigned long long m = __ballot(bit);
                peers &= bit ? m : ~m;
            }
            const int rank = __popcll(peers & below);
            if (valid && rank == 0) cnt[w][d] = __popcll(peers);
            __syncthreads();
            {
                int s = 0;
#pragma unroll
                for (int q = 0; q < RS_BLOCK / 64; ++q) {
                    const int c = cnt[q][tid];
                    cnt[q][tid] = s;
                    s += c;
                }
                tot[tid] = s;
            }
            __syncthreads();
            if (valid) {
                const int pos = run[d] + cnt[w][d] + rank;
                ks[cur ^ 1][pos] = key;
                if (HAS_V) vs[cur ^ 1][pos] = val;
            }
            __syncthreads();
            run[tid] += tot[tid];
#pragma unroll
            for (int q = 0; q < RS_BLOCK / 64; ++q) cnt[q][tid] = 0;
            __syncthreads();
        }
        cur ^= 1;
    }
    for (int i = tid; i < n; i += RS_BLOCK) {
        kout[i] = ks[cur][i];
        if (HAS_V) vout[i] = vs[cur][i];
    }
}

// Sorts n pairs by bits 0 .. key_bits - 1 of the key (descending when desc), stably, on the context's
// stream; vin/vout may be null (keys alone). With seg (device, n_seg + 1 offsets, seg[0] = 0 and
// seg[n_seg] = n), within each range seg[s] .. seg[s + 1] - 1 (see the file's header for what the values
// must satisfy). kin and kout must differ; the inputs are not changed. Uses the context's sort_work and temp.
template <class K>
void radix_sort(Context& C, const K* kin, K* kout, const uint32_t* vin, uint32_t* vout, int64_t n, int key_bits,
                bool desc = false, const int64_t* seg = nullptr, int n_seg = 1) {
    if (n <= 0) return;
    if (n >= ((int64_t)1 << 31)) throw Error(FMMHIP_EINVAL, "radix sort of 2^31 or more keys");
    const bool has_v = vin != nullptr;
    if (seg && n_seg > 1 && !has_v) throw Error(FMMHIP_EINVAL, "a segmented radix sort needs values");
    std::vector<RadixPass> passes;
    for (int s = 0; s < key_bits; s += 8) passes.push_back({0, s, std::min(8, key_bits - s), desc ? 1 : 0});
    if (seg && n_seg > 1) {
        int sb = 0;
        while (((int64_t)1 << sb) < n_seg) ++sb;   // segment indices 0 .. n_seg - 1
        for (int s = 0; s < sb; s += 8) passes.push_back({1, s, std::min(8, sb - s), 0});
    }
    if (passes.empty()) {
        HIP_CHECK(hipMemcpyAsync(kout, kin, sizeof(K) * n, hipMemcpyDeviceToDevice, C.stream));
        if (has_v) HIP_CHECK(hipMemcpyAsync(vout, vin, sizeof(uint32_t) * n, hipMemcpyDeviceToDevice, C.stream));
        return;
    }
    if (n <= RS_TILE && (int)passes.size() <= RS_MAX_PASSES) {   // one workgroup, one launch
        RadixPasses rp;
        rp.n = (int)passes.size();
        for (int i = 0; i < rp.n; ++i) rp.p[i] = passes[i];
        if (has_v) k_radix_small<K, true><<<1, RS_BLOCK, 0, C.stream>>>(kin, vin, (int)n, rp, seg, n_seg, kout, vout);
        else k_radix_small<K, false><<<1, RS_BLOCK, 0, C.stream>>>(kin, nullptr, (int)n, rp, seg, n_seg, kout, nullptr);
        LAUNCH_CHECK();
        return;
    }
    const int64_t nblk = (n + RS_TILE - 1) / RS_TILE;
    const int64_t nt = 256 * nblk;
    auto up = [](size_t b) { return (b + 255) & ~(size_t)255; };
    const size_t b_tab = up(sizeof(int32_t) * nt), b_key = up(sizeof(K) * n), b_val = has_v ? up(sizeof(uint32_t) * n) : 0;
    C.sort_work.resize(2 * b_tab + b_key + b_val);
    unsigned char* w = C.sort_work.data();
    int32_t* table = (int32_t*)w;
    int32_t* off = (int32_t*)(w + b_tab);
    K* tk = (K*)(w + 2 * b_tab);
    uint32_t* tv = has_v ? (uint32_t*)(w + 2 * b_tab + b_key) : nullptr;
    const int P = (int)passes.size();
    const K* sk = kin;
    const uint32_t* sv = vin;
    for (int p = 0; p < P; ++p) {
        // the last pass writes the caller's output; earlier ones alternate so that it does
        const bool to_out = ((P - 1 - p) & 1) == 0;
        K* dk = to_out ? kout : tk;
        uint32_t* dv = to_out ? vout : tv;
        if (has_v) {
            k_radix_count<K, true><<<(unsigned)nblk, RS_BLOCK, 0, C.stream>>>(sk, sv, n, passes[p], seg, n_seg, table, nblk);
            LAUNCH_CHECK();
            excl_scan3<int32_t, int32_t>(C, table, off, nt);
            k_radix_scatter<K, true><<<(unsigned)nblk, RS_BLOCK, 0, C.stream>>>(sk, sv, n, passes[p], seg, n_seg, off, nblk, dk, dv);
        } else {
            k_radix_count<K, false><<<(unsigned)nblk, RS_BLOCK, 0, C.stream>>>(sk, nullptr, n, passes[p], seg, n_seg, table, nblk);
            LAUNCH_CHECK();
            excl_scan3<int32_t, int32_t>(C, table, off, nt);
            k_radix_scatter<K, false><<<(unsigned)nblk, RS_BLOCK, 0, C.stream>>>(sk, nullptr, n, passes[p], seg, n_seg, off, nblk,
                                                                                 dk, nullptr);
        }
        LAUNCH_CHECK();
        sk = dk;
        sv = dv;
    }
}

}  // namespace fmmhip
