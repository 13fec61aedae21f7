// Interaction lists on the GPU, identical (set and order) to FastMultipole's
// build_interaction_lists + sort_by_target (src/interaction_list.jl:3-213).
//
// The reference recurses depth-first from (root, root); each call either emits
// one pair (direct or M2L), drops it, or splits the target or the source into
// its children in branch order. Here the recursion tree is built breadth-first,
// one level per step, and only pending pairs are kept: level k + 1 holds the
// children of the split pairs of level k, in order (exclusive scan of child
// counts). A backward pass then counts the M2L and direct pairs below every node
// (two counters packed in one 64-bit word), and a forward pass hands every node
// the position of its first emitted pair in depth-first order, so each emitted
// pair is written straight to its slot of the reference's discovery order. A
// stable radix sort by target then gives sort_by_target's output. Compiled with
// -ffp-contract=off (the MAC must match).
#include "fmmhip_internal.hpp"
#include "scan3.hpp"
#include "radix_sort.hpp"
#include <algorithm>
#include <memory>
#include <vector>

namespace fmmhip {

enum : int32_t { D_DROP = 0, D_EMIT_M2L = 1, D_EMIT_DIRECT = 2, D_SPLIT_T = 3, D_SPLIT_S = 4 };

// packed (M2L, direct) pair counts: M2L in the high word
static constexpr uint64_t ONE_M2L = 1ull << 32, ONE_DIRECT = 1ull;

struct ListParams {
    int64_t ls;
    double mac;
    int method;
    bool ff, nf, si;
};

__device__ inline bool mac_accept(const Branch& t, const Branch& s, double mac) {
    double dx = t.target_center[0] - s.source_center[0];
    double dy = t.target_center[1] - s.source_center[1];
    double dz = t.target_center[2] - s.source_center[2];
    double sep2 = dx * dx + dy * dy + dz * dz;
    double sr = s.source_radius + t.target_radius;
    return sep2 * mac * mac > sr * sr;
}

// one step of build_interaction_lists! for one pair (:23-166): its decision and number of children (0
// unless split)
__device__ inline int32_t decide_pair(int32_t it, int32_t js, const Branch* __restrict__ tb, const Branch* __restrict__ sb,
                                      const ListParams& P, int32_t& n_children) {
    const Branch& t = tb[it];
    const Branch& s = sb[js];
    const bool both_leaves = (s.n_branches == 0 && t.n_branches == 0);
    int32_t d;
    if (P.method == FMMHIP_SELFTUNING) {
        double fraction = 0.0;
        fraction += (double)s.n_bodies / (double)(P.ls * P.ls);
        fraction *= (double)t.n_bodies;
        if (fraction < 1.0 || both_leaves) d = (P.nf && (it != js || P.si)) ? D_EMIT_DIRECT : D_DROP;
        else if (mac_accept(t, s, P.mac)) d = P.ff ? D_EMIT_M2L : D_DROP;
        else d = (s.n_branches == 0 || (t.n_bodies >= s.n_bodies && t.n_branches != 0)) ? D_SPLIT_T : D_SPLIT_S;
    } else {
        const bool treestop = P.method == FMMHIP_SELFTUNING_TREESTOP;
        const bool direct_ok = (P.nf || (it == js && P.si)) && (it != js || P.si);
        if (treestop && both_leaves) d = direct_ok ? D_EMIT_DIRECT : D_DROP;
        else if (mac_accept(t, s, P.mac)) d = P.ff ? D_EMIT_M2L : D_DROP;
        else if (!treestop && both_leaves) d = direct_ok ? D_EMIT_DIRECT : D_DROP;
        else {
            bool split_t = treestop ? (s.n_branches == 0 || (t.n_bodies >= s.n_bodies && t.n_branches != 0))
                                    : (s.n_branches == 0 || (t.target_radius >= s.source_radius && t.n_branches != 0));
            d = split_t ? D_SPLIT_T : D_SPLIT_S;
        }
    }
    n_children = (d == D_SPLIT_T) ? (int32_t)t.n_branches : (d == D_SPLIT_S) ? (int32_t)s.n_branches : 0;
    return d;
}

// The pairs of a level in tiles of scan3's size (SCAN_BLOCK threads x SCAN_ITEMS pairs, pair i of a tile at
// round i / SCAN_BLOCK): each pair's decision and child count, and each tile's total count. Entry n counts
// 0, so that the exclusive scan of the counts ends with the size of the next level.
__global__ __launch_bounds__(SCAN_BLOCK) void k_list_decide(const int32_t* __restrict__ ft, const int32_t* __restrict__ fs,
                                                            int64_t n, const Branch* __restrict__ tb,
                                                            const Branch* __restrict__ sb, ListParams P,
                                                            int8_t* __restrict__ dec, int32_t* __restrict__ cnt,
                                                            int64_t* __restrict__ tile_total) {
    __shared__ int64_t sh[4];
    const int64_t base = (int64_t)blockIdx.x * SCAN_BLOCK * SCAN_ITEMS;
    int64_t sum = 0;
    for (int r = 0; r < SCAN_ITEMS; ++r) {
        const int64_t i = base + (int64_t)r * SCAN_BLOCK + threadIdx.x;
        if (i < n) {
            int32_t c;
            dec[i] = (int8_t)decide_pair(ft[i], fs[i], tb, sb, P, c);
            cnt[i] = c;
            sum += c;
        } else if (i == n) cnt[i] = 0;
    }
    int64_t total;
    (void)block_excl_scan<int64_t>(sum, total, sh);
    if (threadIdx.x == 0) tile_total[blockIdx.x] = total;
}

// With the tiles' offsets (the exclusive scan of their totals): each pair's position in the next level
// (pos, n + 1 entries, the scan of the counts), and the children of the split pairs, in branch order, there.
__global__ __launch_bounds__(SCAN_BLOCK) void k_list_expand(const int32_t* __restrict__ ft, const int32_t* __restrict__ fs,
                                                            int64_t n, const Branch* __restrict__ tb,
                                                            const Branch* __restrict__ sb, const int8_t* __restrict__ dec,
                                                            const int32_t* __restrict__ cnt,
                                                            const int64_t* __restrict__ tile_off, int64_t* __restrict__ pos,
                                                            int32_t* __restrict__ ot, int32_t* __restrict__ os) {
    __shared__ int64_t sh[4];
    const int64_t base = (int64_t)blockIdx.x * SCAN_BLOCK * SCAN_ITEMS;
    int64_t carry = tile_off[blockIdx.x];
    for (int r = 0; r < SCAN_ITEMS; ++r) {
        const int64_t i = base + (int64_t)r * SCAN_BLOCK + threadIdx.x;
        const int64_t v = i <= n ? (int64_t)cnt[i] : 0;
        int64_t total;
        const int64_t o = carry + block_excl_scan<int64_t>(v, total, sh);
        carry += total;
        if (i > n) continue;
        pos[i] = o;
        if (i == n || v == 0) continue;
        const int32_t it = ft[i], js = fs[i];
        if (dec[i] == D_SPLIT_T) {
            const Branch& t = tb[it];
            for (int64_t k = 0; k < t.n_branches; ++k) { ot[o + k] = (int32_t)(t.child_first + k); os[o + k] = js; }
        } else {
            const Branch& s = sb[js];
            for (int64_t k = 0; k < s.n_branches; ++k) { ot[o + k] = it; os[o + k] = (int32_t)(s.child_first + k); }
        }
    }
}

// packed (M2L, direct) counts of the subtree of every pair of a level
__global__ void k_list_sizes(int64_t n, const int8_t* __restrict__ dec, const int64_t* __restrict__ pos,
                             const uint64_t* __restrict__ child_size, uint64_t* __restrict__ size) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int d = dec[i];
    uint64_t v = 0;
    if (d == D_EMIT_M2L) v = ONE_M2L;
    else if (d == D_EMIT_DIRECT) v = ONE_DIRECT;
    else if (d == D_SPLIT_T || d == D_SPLIT_S)
        for (int64_t c = pos[i]; c < pos[i + 1]; ++c) v += child_size[c];
    size[i] = v;
}

// depth-first positions: an emitted pair goes to its slot of the M2L or direct
// list; a split pair passes running offsets to its children
__global__ void k_list_place(const int32_t* __restrict__ ft, const int32_t* __restrict__ fs, int64_t n,
                             const int8_t* __restrict__ dec, const int64_t* __restrict__ pos, const uint64_t* __restrict__ off,
                             const uint64_t* __restrict__ child_size, uint64_t* __restrict__ child_off,
                             uint32_t* __restrict__ m2l_t, int32_t* __restrict__ m2l_s, uint32_t* __restrict__ dir_t,
                             int32_t* __restrict__ dir_s) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int d = dec[i];
    const uint64_t o = off[i];
    if (d == D_EMIT_M2L) {
        const uint64_t k = o >> 32;
        m2l_t[k] = (uint32_t)ft[i]; m2l_s[k] = fs[i];
    } else if (d == D_EMIT_DIRECT) {
        const uint64_t k = o & 0xffffffffull;
        dir_t[k] = (uint32_t)ft[i]; dir_s[k] = fs[i];
    } else if (d == D_SPLIT_T || d == D_SPLIT_S) {
        uint64_t r = o;
        for (int64_t c = pos[i]; c < pos[i + 1]; ++c) { child_off[c] = r; r += child_size[c]; }
    }
}

__global__ void k_target_starts(const uint32_t* __restrict__ t, int64_t n, int32_t* __restrict__ flag) {
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < n) flag[i] = (i == 0 || t[i] != t[i - 1]) ? 1 : 0;
}

__global__ void k_csr(const uint32_t* __restrict__ t, const int32_t* __restrict__ flag, const int32_t* __restrict__ rank, int64_t n,
                      int32_t* __restrict__ targets, int64_t* __restrict__ offsets, int32_t* __restrict__ out_t) {
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    out_t[i] = (int32_t)t[i];
    if (flag[i]) { targets[rank[i]] = (int32_t)t[i]; offsets[rank[i]] = i; }
}

static inline unsigned grid_for(int64_t n, int block) { return (unsigned)std::max<int64_t>(1, (n + block - 1) / block); }

template <class T, class O>
static void excl_scan(Context& C, const T* in, O* out, int64_t n) {
    excl_scan3<T, O>(C, in, out, n);
}

// number of body-body pairs of the direct list (nearfield work): grid-stride
// partial sums, one atomic per workgroup
__global__ __launch_bounds__(256) void k_count_pairs(const int32_t* __restrict__ t, const int32_t* __restrict__ s, int64_t n,
                                                     const Branch* __restrict__ tb, const Branch* __restrict__ sb,
                                                     unsigned long long* __restrict__ out) {
    __shared__ unsigned long long part[4];
    unsigned long long v = 0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        v += (unsigned long long)tb[t[i]].n_bodies * (unsigned long long)sb[s[i]].n_bodies;
    for (int w = 32; w > 0; w >>= 1) v += __shfl_xor(v, w, 64);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        const unsigned long long tot = part[0] + part[1] + part[2] + part[3];
        if (tot) atomicAdd(out, tot);
    }
}

// source leaf size of every direct-list entry (input of the flattened source stream of the P2P)
__global__ void k_dir_counts(const int32_t* __restrict__ s, int64_t n, const Branch* __restrict__ sb, int64_t* __restrict__ cnt) {
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < n) cnt[i] = sb[s[i]].n_bodies;
    else if (i == n) cnt[i] = 0;
}

// the recursion starts from (root, root) of every system; roots are branches 0 .. n - 1 of both trees
__global__ void k_init_roots(int32_t* ft, int32_t* fs, int64_t n) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < n) { ft[i] = (int32_t)i; fs[i] = (int32_t)i; }
}

__global__ void k_zero_counter(unsigned long long* c) { *c = 0; }

// one level of the recursion tree
struct ListLevel {
    DBuf<int32_t> ft, fs;
    DBuf<int8_t> dec;
    DBuf<int64_t> pos;
    DBuf<uint64_t> size, off;
};

struct ListScratch {
    std::vector<std::unique_ptr<ListLevel>> levels;
    DBuf<int32_t> cnt, flag, rank;
    DBuf<int64_t> pos, tile;
    DBuf<unsigned long long> counter;
    DBuf<uint32_t> kt_m2l, kt_dir, kt_sorted;
    DBuf<int32_t> vs_m2l, vs_dir;
    ListLevel& level(size_t k) {
        while (levels.size() <= k) levels.emplace_back(new ListLevel());
        return *levels[k];
    }
};

static ListScratch& lscratch(Context& C) {   // per context, like the tree build's
    if (!C.list_scratch) C.list_scratch = new ListScratch();
    return *C.list_scratch;
}

void free_list_scratch(ListScratch* s) { delete s; }

// stable sort by target (== sort_by_target, src/interaction_list.jl:183-213) of a
// list in discovery order, then CSR over the unique targets
static void finish_list(Context& C, ListScratch& S, uint32_t* kt, int32_t* vs, int64_t n_out, int target_bits,
                        DBuf<int32_t>& out_t, DBuf<int32_t>& out_s, DBuf<int32_t>& targets, DBuf<int64_t>& offsets,
                        int64_t& n_targets) {
    out_t.resize(std::max<int64_t>(1, n_out));
    out_s.resize(std::max<int64_t>(1, n_out));
    n_targets = 0;
    if (n_out == 0) { offsets.resize(1); targets.resize(1); HIP_CHECK(hipMemsetAsync(offsets.data(), 0, 8, C.stream)); return; }
    S.kt_sorted.resize(n_out);
    radix_sort<uint32_t>(C, kt, S.kt_sorted.data(), (const uint32_t*)vs, (uint32_t*)out_s.data(), n_out, target_bits);
    S.flag.resize(n_out + 1);
    S.rank.resize(n_out + 1);
    k_target_starts<<<grid_for(n_out, 256), 256, 0, C.stream>>>(S.kt_sorted.data(), n_out, S.flag.data());
    LAUNCH_CHECK();
    excl_scan(C, S.flag.data(), S.rank.data(), n_out);
    n_targets = read_back_sum(C, S.rank.data() + (n_out - 1), S.flag.data() + (n_out - 1));
    targets.resize(n_targets);
    offsets.resize(n_targets + 1);
    k_csr<<<grid_for(n_out, 256), 256, 0, C.stream>>>(S.kt_sorted.data(), S.flag.data(), S.rank.data(), n_out, targets.data(),
                                                     offsets.data(), out_t.data());
    LAUNCH_CHECK();
    HIP_CHECK(hipMemcpyAsync(offsets.data() + n_targets, &n_out, sizeof(int64_t), hipMemcpyHostToDevice, C.stream));
    HIP_CHECK(hipStreamSynchronize(C.stream));
}

void build_lists(Context& C, const Tree& tt, const Tree& st, Lists& L, int64_t ls, double mac, int method, bool ff, bool nf,
                 bool si, bool count_pairs) {
    L.n_m2l = L.n_direct = L.n_pairs = 0;
    L.n_direct_targets = L.n_m2l_targets = 0;
    if (tt.n_branches == 0 || st.n_branches == 0) return;
    ListScratch& S = lscratch(C);
    ListParams P{ls, mac, method, ff, nf, si};
    // a forest of independent systems (one root per system) pairs each system with itself only
    const int64_t n_roots = tt.seg.size() > 2 ? (int64_t)tt.seg.size() - 1 : 1;
    if ((st.seg.size() > 2 ? (int64_t)st.seg.size() - 1 : 1) != n_roots)
        throw Error(FMMHIP_EINVAL, "target and source forests hold different numbers of systems");
    // forward: levels of pending pairs
    std::vector<int64_t> sizes;
    {
        ListLevel& L0 = S.level(0);
        L0.ft.resize(n_roots); L0.fs.resize(n_roots); L0.off.resize(n_roots);
        k_init_roots<<<grid_for(n_roots, 256), 256, 0, C.stream>>>(L0.ft.data(), L0.fs.data(), n_roots);
        LAUNCH_CHECK();
    }
    int64_t n = n_roots;
    for (size_t k = 0;; ++k) {
        if (k >= 4096) throw Error(FMMHIP_EDEPTH, "interaction-list recursion deeper than 4096 levels");
        ListLevel& A = S.level(k);
        sizes.push_back(n);
        A.dec.resize(n);
        A.pos.resize(n + 1);
        S.cnt.resize(n + 1);
        // decisions with per-tile totals, the tiles' offsets, then positions and children: three launches
        const int64_t tiles = (n + 1 + SCAN_BLOCK * SCAN_ITEMS - 1) / (SCAN_BLOCK * SCAN_ITEMS);
        S.tile.resize(tiles + 1);
        k_list_decide<<<(unsigned)tiles, SCAN_BLOCK, 0, C.stream>>>(A.ft.data(), A.fs.data(), n, tt.branches.data(),
                                                                    st.branches.data(), P, A.dec.data(), S.cnt.data(),
                                                                    S.tile.data());
        LAUNCH_CHECK();
        k_scan_offsets<int64_t><<<1, SCAN_BLOCK, 0, C.stream>>>(S.tile.data(), tiles);
        LAUNCH_CHECK();
        const int64_t next = read_back(C, S.tile.data() + tiles);
        ListLevel& B = S.level(k + 1);
        if (next) { B.ft.resize(next); B.fs.resize(next); }
        k_list_expand<<<(unsigned)tiles, SCAN_BLOCK, 0, C.stream>>>(A.ft.data(), A.fs.data(), n, tt.branches.data(),
                                                                    st.branches.data(), A.dec.data(), S.cnt.data(),
                                                                    S.tile.data(), A.pos.data(), B.ft.data(), B.fs.data());
        LAUNCH_CHECK();
        if (next == 0) break;
        n = next;
    }
    const int K = (int)sizes.size();
    // backward: subtree counts
    for (int k = K - 1; k >= 0; --k) {
        ListLevel& A = S.level(k);
        A.size.resize(sizes[k]);
        const uint64_t* child = (k + 1 < K) ? S.level(k + 1).size.data() : nullptr;
        k_list_sizes<<<grid_for(sizes[k], 256), 256, 0, C.stream>>>(sizes[k], A.dec.data(), A.pos.data(), child, A.size.data());
        LAUNCH_CHECK();
    }
    // depth-first positions of the roots' subtrees: exclusive scan of their packed counts
    ListLevel& L0 = S.level(0);
    excl_scan(C, L0.size.data(), L0.off.data(), n_roots);
    const uint64_t total = (uint64_t)read_back_sum(C, L0.off.data() + (n_roots - 1), L0.size.data() + (n_roots - 1));
    L.n_m2l = (int64_t)(total >> 32);
    L.n_direct = (int64_t)(total & 0xffffffffull);
    S.kt_m2l.resize(std::max<int64_t>(1, L.n_m2l)); S.vs_m2l.resize(std::max<int64_t>(1, L.n_m2l));
    S.kt_dir.resize(std::max<int64_t>(1, L.n_direct)); S.vs_dir.resize(std::max<int64_t>(1, L.n_direct));
    // forward: depth-first placement
    for (int k = 0; k < K; ++k) {
        ListLevel& A = S.level(k);
        uint64_t* child_off = nullptr;
        const uint64_t* child_size = nullptr;
        if (k + 1 < K) {
            ListLevel& B = S.level(k + 1);
            B.off.resize(sizes[k + 1]);
            child_off = B.off.data();
            child_size = B.size.data();
        }
        k_list_place<<<grid_for(sizes[k], 256), 256, 0, C.stream>>>(A.ft.data(), A.fs.data(), sizes[k], A.dec.data(), A.pos.data(),
                                                                   A.off.data(), child_size, child_off, S.kt_m2l.data(),
                                                                   S.vs_m2l.data(), S.kt_dir.data(), S.vs_dir.data());
        LAUNCH_CHECK();
    }
    if (L.keep_discovery) {
        L.dfs_t.resize(std::max<int64_t>(1, L.n_direct));
        L.dfs_s.resize(std::max<int64_t>(1, L.n_direct));
        if (L.n_direct) {
            HIP_CHECK(hipMemcpyAsync(L.dfs_t.data(), S.kt_dir.data(), sizeof(uint32_t) * L.n_direct, hipMemcpyDeviceToDevice, C.stream));
            HIP_CHECK(hipMemcpyAsync(L.dfs_s.data(), S.vs_dir.data(), sizeof(int32_t) * L.n_direct, hipMemcpyDeviceToDevice, C.stream));
        }
    }
    int bits = 1;
    while (bits < 32 && (int64_t(1) << bits) < tt.n_branches) ++bits;
    finish_list(C, S, S.kt_m2l.data(), S.vs_m2l.data(), L.n_m2l, bits, L.m2l_t, L.m2l_s, L.m2l_targets, L.m2l_offsets,
                L.n_m2l_targets);
    finish_list(C, S, S.kt_dir.data(), S.vs_dir.data(), L.n_direct, bits, L.dir_t, L.dir_s, L.dir_targets, L.dir_offsets,
                L.n_direct_targets);
    L.dir_level_split.clear();
    if (method == FMMHIP_SELFTUNING && L.n_direct_targets > 0) {
        // SelfTuning emits direct pairs between branches that are not leaves (src/interaction_list.jl:84-88),
        // so one direct target's bodies can lie inside another's. The nearfield then runs one launch per
        // level of targets (targets of one level are disjoint), ancestors first: the order of the
        // reference's target-sorted list.
        std::vector<int32_t> t(L.n_direct_targets);
        HIP_CHECK(hipMemcpyAsync(t.data(), L.dir_targets.data(), t.size() * sizeof(int32_t), hipMemcpyDeviceToHost, C.stream));
        HIP_CHECK(hipStreamSynchronize(C.stream));
        for (size_t l = 0; l < tt.level_first.size(); ++l)
            L.dir_level_split.push_back(std::lower_bound(t.begin(), t.end(), (int32_t)tt.level_first[l]) - t.begin());
        L.dir_level_split.push_back(L.n_direct_targets);
    }
    L.m2l_order.resize(std::max<int64_t>(1, L.n_m2l));
    if (L.n_direct > 0) {
        S.pos.resize(L.n_direct + 1);
        L.dir_cum.resize(L.n_direct + 1);
        k_dir_counts<<<grid_for(L.n_direct + 1, 256), 256, 0, C.stream>>>(L.dir_s.data(), L.n_direct, st.branches.data(), S.pos.data());
        LAUNCH_CHECK();
        excl_scan(C, S.pos.data(), L.dir_cum.data(), L.n_direct + 1);
    }
    L.n_pairs = 0;
    if (count_pairs && L.n_direct > 0) {
        S.counter.resize(1);
        k_zero_counter<<<1, 1, 0, C.stream>>>(S.counter.data());
        k_count_pairs<<<std::min(1024u, grid_for(L.n_direct, 256)), 256, 0, C.stream>>>(L.dir_t.data(), L.dir_s.data(), L.n_direct,
                                                                       tt.branches.data(), st.branches.data(), S.counter.data());
        LAUNCH_CHECK();
        L.n_pairs = (int64_t)read_back(C, S.counter.data());
    }
}

}  // namespace fmmhip
