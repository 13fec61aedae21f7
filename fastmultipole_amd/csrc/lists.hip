// Interaction lists on the GPU, identical (set and order) to FastMultipole's
// build_interaction_lists + sort_by_target (src/interaction_list.jl:3-213).
//
// The reference recurses depth-first from (root, root); each call either emits
// one pair (direct or M2L), drops it, or splits the target or the source into
// its children in branch order. Here the recursion runs breadth-first over a
// frontier that is kept in depth-first order: every step replaces each pending
// pair by its outputs in place (exclusive scan of output counts), and emitted
// pairs are carried along unchanged. When nothing is pending the frontier is
// the reference's discovery order; a stable radix sort by target then gives
// sort_by_target's output. Compiled with -ffp-contract=off (the MAC must match).
#include "fmmhip_internal.hpp"
#include <cstring>   // host memset for rocprim headers
#include <rocprim/rocprim.hpp>
#include <algorithm>

namespace fmmhip {

enum : int32_t { ST_PENDING = 0, ST_M2L = 1, ST_DIRECT = 2 };
enum : int32_t { D_CARRY = 0, D_DROP = 1, D_EMIT_M2L = 2, D_EMIT_DIRECT = 3, D_SPLIT_T = 4, D_SPLIT_S = 5 };

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

// one step of build_interaction_lists! for every pending pair (:23-166)
__global__ void k_list_decide(const int32_t* __restrict__ ft, const int32_t* __restrict__ fs, const int32_t* __restrict__ fst,
                              int64_t n, const Branch* __restrict__ tb, const Branch* __restrict__ sb, ListParams P,
                              int32_t* __restrict__ dec, int32_t* __restrict__ cnt, unsigned long long* __restrict__ n_pending) {
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (fst[i] != ST_PENDING) { dec[i] = D_CARRY; cnt[i] = 1; return; }
    const int32_t it = ft[i], js = fs[i];
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
    dec[i] = d;
    int32_t c = (d == D_DROP) ? 0 : (d == D_SPLIT_T) ? (int32_t)t.n_branches : (d == D_SPLIT_S) ? (int32_t)s.n_branches : 1;
    cnt[i] = c;
    if (d == D_SPLIT_T || d == D_SPLIT_S) atomicAdd(n_pending, (unsigned long long)c);
}

__global__ void k_list_write(const int32_t* __restrict__ ft, const int32_t* __restrict__ fs, const int32_t* __restrict__ fst,
                             int64_t n, const Branch* __restrict__ tb, const Branch* __restrict__ sb,
                             const int32_t* __restrict__ dec, const int64_t* __restrict__ pos, int32_t* __restrict__ ot,
                             int32_t* __restrict__ os, int32_t* __restrict__ ost) {
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int64_t o = pos[i];
    const int32_t it = ft[i], js = fs[i];
    switch (dec[i]) {
        case D_CARRY: ot[o] = it; os[o] = js; ost[o] = fst[i]; break;
        case D_EMIT_M2L: ot[o] = it; os[o] = js; ost[o] = ST_M2L; break;
        case D_EMIT_DIRECT: ot[o] = it; os[o] = js; ost[o] = ST_DIRECT; break;
        case D_SPLIT_T: {
            const Branch& t = tb[it];
            for (int64_t k = 0; k < t.n_branches; ++k) { ot[o + k] = (int32_t)(t.child_first + k); os[o + k] = js; ost[o + k] = ST_PENDING; }
            break;
        }
        case D_SPLIT_S: {
            const Branch& s = sb[js];
            for (int64_t k = 0; k < s.n_branches; ++k) { ot[o + k] = it; os[o + k] = (int32_t)(s.child_first + k); ost[o + k] = ST_PENDING; }
            break;
        }
        default: break;
    }
}

__global__ void k_state_flags(const int32_t* __restrict__ st, int64_t n, int32_t want, int32_t* __restrict__ flag) {
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < n) flag[i] = st[i] == want ? 1 : 0;
}

__global__ void k_compact(const int32_t* __restrict__ ft, const int32_t* __restrict__ fs, const int32_t* __restrict__ flag,
                          const int32_t* __restrict__ rank, int64_t n, uint32_t* __restrict__ kt, int32_t* __restrict__ vs) {
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < n && flag[i]) { kt[rank[i]] = (uint32_t)ft[i]; vs[rank[i]] = fs[i]; }
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
    size_t bytes = 0;
    HIP_CHECK(rocprim::exclusive_scan(nullptr, bytes, in, out, O(0), (size_t)n, rocprim::plus<O>(), C.stream));
    C.temp.resize(bytes + 16);
    HIP_CHECK(rocprim::exclusive_scan(C.temp.data(), bytes, in, out, O(0), (size_t)n, rocprim::plus<O>(), C.stream));
}

template <class T>
static T read_back(Context& C, const T* dptr) {
    T v;
    HIP_CHECK(hipMemcpyAsync(&v, dptr, sizeof(T), hipMemcpyDeviceToHost, C.stream));
    HIP_CHECK(hipStreamSynchronize(C.stream));
    return v;
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

__global__ void k_init_frontier(int32_t* ft, int32_t* fs, int32_t* fst, unsigned long long* counter) {
    ft[0] = 0; fs[0] = 0; fst[0] = ST_PENDING;
    *counter = 0;
}

__global__ void k_zero_counter(unsigned long long* c) { *c = 0; }

struct ListScratch {
    DBuf<int32_t> ft[2], fs[2], fst[2];
    DBuf<int32_t> dec, cnt, flag, rank;
    DBuf<int64_t> pos;
    DBuf<unsigned long long> counter;
    DBuf<uint32_t> kt, kt_sorted;
    DBuf<int32_t> vs, vs_sorted;
};

static ListScratch& lscratch() {
    static ListScratch* s = new ListScratch();   // process lifetime, see tree.hip
    return *s;
}

// sorted-by-target list of one kind; fills CSR over unique targets
static void extract_list(Context& C, ListScratch& S, int cur, int64_t n, int32_t state, DBuf<int32_t>& out_t,
                         DBuf<int32_t>& out_s, int64_t& n_out, DBuf<int32_t>& targets, DBuf<int64_t>& offsets,
                         int64_t& n_targets) {
    S.flag.resize(n + 1);
    S.rank.resize(n + 1);
    k_state_flags<<<grid_for(n, 256), 256, 0, C.stream>>>(S.fst[cur].data(), n, state, S.flag.data());
    LAUNCH_CHECK();
    excl_scan(C, S.flag.data(), S.rank.data(), n);
    n_out = (int64_t)read_back(C, S.rank.data() + (n - 1)) + read_back(C, S.flag.data() + (n - 1));
    out_t.resize(std::max<int64_t>(1, n_out));
    out_s.resize(std::max<int64_t>(1, n_out));
    n_targets = 0;
    if (n_out == 0) { offsets.resize(1); targets.resize(1); HIP_CHECK(hipMemsetAsync(offsets.data(), 0, 8, C.stream)); return; }
    S.kt.resize(n_out); S.vs.resize(n_out); S.kt_sorted.resize(n_out); S.vs_sorted.resize(n_out);
    k_compact<<<grid_for(n, 256), 256, 0, C.stream>>>(S.ft[cur].data(), S.fs[cur].data(), S.flag.data(), S.rank.data(), n,
                                                     S.kt.data(), S.vs.data());
    LAUNCH_CHECK();
    size_t bytes = 0;   // stable LSD radix sort by target == sort_by_target (src/interaction_list.jl:183-213)
    HIP_CHECK(rocprim::radix_sort_pairs(nullptr, bytes, S.kt.data(), S.kt_sorted.data(), S.vs.data(), out_s.data(), (size_t)n_out,
                                        0, 32, C.stream));
    C.temp.resize(bytes + 16);
    HIP_CHECK(rocprim::radix_sort_pairs(C.temp.data(), bytes, S.kt.data(), S.kt_sorted.data(), S.vs.data(), out_s.data(),
                                        (size_t)n_out, 0, 32, C.stream));
    S.flag.resize(n_out + 1);
    S.rank.resize(n_out + 1);
    k_target_starts<<<grid_for(n_out, 256), 256, 0, C.stream>>>(S.kt_sorted.data(), n_out, S.flag.data());
    LAUNCH_CHECK();
    excl_scan(C, S.flag.data(), S.rank.data(), n_out);
    n_targets = (int64_t)read_back(C, S.rank.data() + (n_out - 1)) + read_back(C, S.flag.data() + (n_out - 1));
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
    ListScratch& S = lscratch();
    ListParams P{ls, mac, method, ff, nf, si};
    int cur = 0;
    int64_t n = 1;
    S.ft[0].resize(1); S.fs[0].resize(1); S.fst[0].resize(1);
    S.counter.resize(1);
    k_init_frontier<<<1, 1, 0, C.stream>>>(S.ft[0].data(), S.fs[0].data(), S.fst[0].data(), S.counter.data());
    LAUNCH_CHECK();
    for (int step = 0; step < 4096; ++step) {
        S.dec.resize(n); S.cnt.resize(n); S.pos.resize(n + 1);
        k_zero_counter<<<1, 1, 0, C.stream>>>(S.counter.data());
        k_list_decide<<<grid_for(n, 256), 256, 0, C.stream>>>(S.ft[cur].data(), S.fs[cur].data(), S.fst[cur].data(), n,
                                                              tt.branches.data(), st.branches.data(), P, S.dec.data(),
                                                              S.cnt.data(), S.counter.data());
        LAUNCH_CHECK();
        excl_scan(C, S.cnt.data(), S.pos.data(), n);
        const int64_t total = read_back(C, S.pos.data() + (n - 1)) + (int64_t)read_back(C, S.cnt.data() + (n - 1));
        const unsigned long long pending = read_back(C, S.counter.data());
        const int nxt = cur ^ 1;
        S.ft[nxt].resize(std::max<int64_t>(1, total)); S.fs[nxt].resize(std::max<int64_t>(1, total));
        S.fst[nxt].resize(std::max<int64_t>(1, total));
        k_list_write<<<grid_for(n, 256), 256, 0, C.stream>>>(S.ft[cur].data(), S.fs[cur].data(), S.fst[cur].data(), n,
                                                            tt.branches.data(), st.branches.data(), S.dec.data(), S.pos.data(),
                                                            S.ft[nxt].data(), S.fs[nxt].data(), S.fst[nxt].data());
        LAUNCH_CHECK();
        cur = nxt;
        n = total;
        if (pending == 0 || n == 0) break;
    }
    if (n == 0) return;
    extract_list(C, S, cur, n, ST_M2L, L.m2l_t, L.m2l_s, L.n_m2l, L.m2l_targets, L.m2l_offsets, L.n_m2l_targets);
    extract_list(C, S, cur, n, ST_DIRECT, L.dir_t, L.dir_s, L.n_direct, L.dir_targets, L.dir_offsets, L.n_direct_targets);
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
        k_zero_counter<<<1, 1, 0, C.stream>>>(S.counter.data());
        k_count_pairs<<<std::min(1024u, grid_for(L.n_direct, 256)), 256, 0, C.stream>>>(L.dir_t.data(), L.dir_s.data(), L.n_direct,
                                                                       tt.branches.data(), st.branches.data(), S.counter.data());
        LAUNCH_CHECK();
        L.n_pairs = (int64_t)read_back(C, S.counter.data());
    }
}

}  // namespace fmmhip
