// Octree construction on the GPU, bit-exact with FastMultipole's Tree()
// (src/tree.jl:3-122, 346-463, 581-592, 608-743, 845-1188).
//
// The reference partitions each node's bodies stably into octants, level by
// level, so the final body order is: lexicographic by the octant path down to
// the body's leaf, then by original index. Here every body gets its full octant
// path (up to 42 levels, two 63-bit words) in one pass, computed with exactly
// the reference's child-centre arithmetic (centre +/- box*0.5, box halved per
// level). A stable radix sort by path groups bodies by node at every depth; the
// node structure is then read off the sorted keys level by level with binary
// searches, using the reference's leaf test on the node counts. A second sort
// by (leaf, original index) restores the reference's in-leaf order.
//
// A forest (several independent systems, Tree.seg holding their body offsets) is
// built in the same launches: each system gets its own root cube from its own
// bodies, the keys are sorted within each system's range, and the roots are
// branches 0 .. n_systems - 1 of level 0, so every system's subtree is the tree
// the reference builds for that system alone.
//
// This file is compiled with -ffp-contract=off: every floating-point result
// here (centres, boxes, radii, the leaf test, the MAC) must match the
// reference's IEEE double arithmetic bit for bit.
#include "fmmhip_internal.hpp"
#include "scan3.hpp"
#include "radix_sort.hpp"
#include <algorithm>
#include <atomic>
#include <cmath>

namespace fmmhip {

static constexpr int LEVELS_PER_WORD = 21;
static constexpr int MAX_LEVELS = 2 * LEVELS_PER_WORD;

// Julia's min/max on Float64 (order -0.0 < 0.0; inputs are never NaN here)
__device__ __host__ inline double jmin(double a, double b) { return (b < a || (b == a && std::signbit(b))) ? b : a; }
__device__ __host__ inline double jmax(double a, double b) { return (b > a || (b == a && !std::signbit(b))) ? b : a; }

__device__ inline int digit_at(uint64_t hi, uint64_t lo, int level) {
    return level < LEVELS_PER_WORD ? (int)((hi >> (3 * (LEVELS_PER_WORD - 1 - level))) & 7u)
                                   : (int)((lo >> (3 * (MAX_LEVELS - 1 - level))) & 7u);
}

// exceeds (src/tree.jl:492-569), single system
__device__ __host__ inline bool exceeds(int64_t n, int64_t ls, int method) {
    if (method == FMMHIP_BARBA) return n > ls;
    double fraction = 0.0;
    fraction += (double)n / (double)(ls * ls);
    fraction *= (double)ls;
    if (method == FMMHIP_SELFTUNING) return fraction > 0.5 && n > 1;
    return fraction > 2.8284271247461903 && n > 1;
}

// exceeds for several systems in one tree (src/tree.jl:492-569): each system's count against its own leaf
// size; count(i) gives system i's bodies in the branch
template <class Count>
__device__ __host__ inline bool exceeds_systems(int n_sys, Count count, const int64_t* ls, int method) {
    if (method == FMMHIP_BARBA) {
        bool not_over = true;
        for (int i = 0; i < n_sys; ++i) not_over = not_over && count(i) <= ls[i];
        return !not_over;
    }
    double fraction = 0.0;
    int64_t n_bodies = 0, min_ls = ls[0];
    for (int i = 0; i < n_sys; ++i) {
        const int64_t n = count(i);
        n_bodies += n;
        fraction += (double)n / (double)(ls[i] * ls[i]);
        min_ls = min_ls < ls[i] ? min_ls : ls[i];
    }
    fraction *= (double)min_ls;
    if (method == FMMHIP_SELFTUNING) return fraction > 0.5 && n_bodies > 1;
    return fraction > 2.8284271247461903 && n_bodies > 1;
}

// flag[i] = 1 where body i of the key order (original index idx[i]) belongs to system s; flag[n] = 0
__global__ void k_sys_flag(const uint32_t* __restrict__ idx, int64_t n, const int64_t* __restrict__ off, int n_sys, int s,
                           int32_t* __restrict__ flag) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i > n) return;
    if (i == n) { flag[i] = 0; return; }
    const int64_t v = idx[i];
    flag[i] = (v >= off[s] && v < off[s + 1]) ? 1 : 0;
    (void)n_sys;
}

// ------------------------------------------------------------- bounding box
// block (x, y) reduces its share of system y's bodies (seg null: one system, bodies 0 .. n - 1)
__global__ void k_bbox_partial(const double* __restrict__ buf, int rows, int64_t n, const int64_t* __restrict__ seg,
                               double* __restrict__ part) {
    __shared__ double s[6][256];
    const int64_t f = seg ? seg[blockIdx.y] : 0, e = seg ? seg[blockIdx.y + 1] : n;
    double mn[3], mx[3];
    for (int d = 0; d < 3; ++d) { mn[d] = buf[f * rows + d]; mx[d] = buf[f * rows + d]; }
    for (int64_t i = f + blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < e; i += (int64_t)gridDim.x * blockDim.x)
        for (int d = 0; d < 3; ++d) {
            double v = buf[i * rows + d];
            mn[d] = jmin(mn[d], v);
            mx[d] = jmax(mx[d], v);
        }
    for (int d = 0; d < 3; ++d) { s[d][threadIdx.x] = mn[d]; s[3 + d][threadIdx.x] = mx[d]; }
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int d = 0; d < 3; ++d) {
                s[d][threadIdx.x] = jmin(s[d][threadIdx.x], s[d][threadIdx.x + w]);
                s[3 + d][threadIdx.x] = jmax(s[3 + d][threadIdx.x], s[3 + d][threadIdx.x + w]);
            }
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int d = 0; d < 6; ++d) part[((int64_t)blockIdx.y * gridDim.x + blockIdx.x) * 6 + d] = s[d][0];
}

// ---------------------------------------------------------------- path keys
// roots: (centre x, y, z, half side) of the root cube, one per system; seg: the systems' body offsets
// (n_seg + 1 entries), or null for one system
__global__ void k_path_keys(const double* __restrict__ buf, int rows, int64_t n, const double* __restrict__ roots,
                            const int64_t* __restrict__ seg, int n_seg, uint64_t* __restrict__ khi,
                            uint64_t* __restrict__ klo, uint32_t* __restrict__ idx) {
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    int r = 0;
    if (seg) {   // last system whose first body is at or before i
        int lo = 0, hi = n_seg - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (seg[mid] <= i) lo = mid; else hi = mid - 1;
        }
        r = lo;
    }
    const double x = buf[i * rows + 0], y = buf[i * rows + 1], z = buf[i * rows + 2];
    double cx = roots[4 * r], cy = roots[4 * r + 1], cz = roots[4 * r + 2], box = roots[4 * r + 3];
    uint64_t w[2] = {0, 0};
    for (int word = 0; word < 2; ++word)
        for (int l = 0; l < LEVELS_PER_WORD; ++l) {
            int o = (x > cx ? 1 : 0) | (y > cy ? 2 : 0) | (z > cz ? 4 : 0);   // get_octant - 1
            w[word] = (w[word] << 3) | (uint64_t)o;
            double delta = box * 0.5;                                       // get_child_center
            cx = cx + ((o & 1) ? delta : -delta);
            cy = cy + ((o & 2) ? delta : -delta);
            cz = cz + ((o & 4) ? delta : -delta);
            box = box * 0.5;                                                // child_box = box * 0.5
        }
    khi[i] = w[0];
    klo[i] = w[1];
    idx[i] = (uint32_t)i;
}

__global__ void k_gather_u64(const uint64_t* __restrict__ src, const uint32_t* __restrict__ idx, int64_t n,
                             uint64_t* __restrict__ dst) {
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[idx[i]];
}

// ------------------------------------------------------ structure per level
// one thread per (active node, octant): body sub-range with that digit
__global__ void k_split(const uint64_t* __restrict__ khi, const uint64_t* __restrict__ klo, int level, int64_t n_active,
                        const int64_t* __restrict__ act_first, const int64_t* __restrict__ act_count,
                        int64_t* __restrict__ sub_first, int64_t* __restrict__ sub_count, int32_t* __restrict__ nonempty,
                        int32_t* __restrict__ split_flag) {
    int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t <= n_active * 8) split_flag[t] = 0;   // entries past the level's children stay 0 for the scan
    if (t == n_active * 8) nonempty[t] = 0;     // so that the scan's last entry is the number of children
    if (t >= n_active * 8) return;
    int64_t a = t >> 3;
    int o = (int)(t & 7);
    int64_t f = act_first[a], e = f + act_count[a];
    auto lower = [&](int d) {   // first position in [f,e) whose digit >= d
        int64_t lo = f, hi = e;
        while (lo < hi) {
            int64_t mid = (lo + hi) >> 1;
            if (digit_at(khi[mid], klo[mid], level) < d) lo = mid + 1; else hi = mid;
        }
        return lo;
    };
    int64_t s0 = lower(o), s1 = (o == 7) ? e : lower(o + 1);
    sub_first[t] = s0;
    sub_count[t] = s1 - s0;
    nonempty[t] = (s1 > s0) ? 1 : 0;
}

// child branches (child_branches!, src/tree.jl:346-376 and branch!, :433-463)
__global__ void k_make_children(int64_t n_active, const int64_t* __restrict__ act_branch, const int64_t* __restrict__ sub_first,
                                const int64_t* __restrict__ sub_count, const int32_t* __restrict__ rank,
                                int64_t child_base, Branch* __restrict__ br, int64_t ls, int method,
                                int32_t* __restrict__ split_flag, const int32_t* __restrict__ sys_prefix, int n_sys,
                                int64_t n_bodies_all, const int64_t* __restrict__ sys_ls) {
    int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= n_active * 8) return;
    int64_t cnt = sub_count[t];
    if (cnt == 0) return;
    int64_t a = t >> 3;
    int o = (int)(t & 7);
    const Branch& P = br[act_branch[a]];
    int64_t ci = child_base + rank[t];
    Branch c;
    c.n_bodies = cnt;
    c.first = sub_first[t];
    c.n_branches = 0;
    c.child_first = -1;
    c.parent = act_branch[a];
    c.leaf = -1;
    const double child_radius = P.target_radius * 0.5;
    for (int d = 0; d < 3; ++d) {
        double delta = P.target_box[d] * 0.5;
        double cc = P.target_center[d] + (((o >> d) & 1) ? delta : -delta);
        c.source_center[d] = cc;
        c.target_center[d] = cc;
        c.source_box[d] = P.target_box[d] * 0.5;
        c.target_box[d] = P.target_box[d] * 0.5;
    }
    c.source_radius = child_radius;
    c.target_radius = child_radius;
    c.max_influence = 0.0;
    c.pad = 0.0;
    br[ci] = c;
    bool split;
    if (n_sys > 1) {   // several systems: their counts in the child's range from the per-system prefixes
        const int64_t b = c.first, e = c.first + cnt, stride = n_bodies_all + 1;
        split = exceeds_systems(
            n_sys, [&](int i) { return (int64_t)(sys_prefix[i * stride + e] - sys_prefix[i * stride + b]); }, sys_ls, method);
    } else split = exceeds(cnt, ls, method);
    split_flag[rank[t]] = split ? 1 : 0;
}

// parents' child ranges: first nonempty octant's child index, number of children
__global__ void k_link_parents(int64_t n_active, const int64_t* __restrict__ act_branch, const int32_t* __restrict__ nonempty,
                               const int32_t* __restrict__ rank, int64_t child_base, Branch* __restrict__ br) {
    int64_t a = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (a >= n_active) return;
    int nb = 0;
    int64_t first = -1;
    for (int o = 0; o < 8; ++o)
        if (nonempty[a * 8 + o]) {
            if (first < 0) first = child_base + rank[a * 8 + o];
            ++nb;
        }
    Branch& P = br[act_branch[a]];
    P.n_branches = nb;
    P.child_first = first;
}

// next active list from split flags
__global__ void k_next_active(int64_t n_children, const int32_t* __restrict__ split_flag, const int32_t* __restrict__ srank,
                              int64_t child_base, const Branch* __restrict__ br, int64_t* __restrict__ act_branch,
                              int64_t* __restrict__ act_first, int64_t* __restrict__ act_count) {
    int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (c >= n_children || !split_flag[c]) return;   // n_children: a bound (flags past the children are 0)
    int64_t k = srank[c];
    const Branch& B = br[child_base + c];
    act_branch[k] = child_base + c;
    act_first[k] = B.first;
    act_count[k] = B.n_bodies;
}

// ------------------------------------------------ leaves, final body order
__global__ void k_leaf_flags(const Branch* __restrict__ br, int64_t nb, int32_t* __restrict__ flag) {
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < nb) flag[i] = br[i].n_branches == 0 ? 1 : 0;
}

__global__ void k_leaf_number(Branch* __restrict__ br, int64_t nb, const int32_t* __restrict__ flag,
                              const int32_t* __restrict__ rank, int64_t* __restrict__ leaf_index) {
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= nb || !flag[i]) return;
    br[i].leaf = rank[i];
    leaf_index[rank[i]] = i;
}

// key (leaf first body << 32 | original index): sorting it keeps leaves in place and
// orders each leaf's bodies by original index
__global__ void k_leaf_keys(const Branch* __restrict__ br, const int64_t* __restrict__ leaf_index, int64_t n_leaves,
                            const uint32_t* __restrict__ idx_sorted, uint64_t* __restrict__ key2) {
    int64_t l = blockIdx.x;
    if (l >= n_leaves) return;
    const Branch& B = br[leaf_index[l]];
    for (int64_t i = B.first + threadIdx.x; i < B.first + B.n_bodies; i += blockDim.x)
        key2[i] = ((uint64_t)B.first << 32) | (uint64_t)idx_sorted[i];
}

// In-leaf order = original index order, one wavefront per leaf: a body's place in its leaf is the number of
// the leaf's bodies with a smaller original index (leaves of at most LEAF_ORDER_MAX bodies, staged in LDS)
constexpr int LEAF_ORDER_MAX = 1024;
__global__ __launch_bounds__(64) void k_leaf_order(const Branch* __restrict__ br, const int64_t* __restrict__ leaf_index,
                                                   int64_t n_leaves, const uint32_t* __restrict__ idx_sorted,
                                                   int64_t* __restrict__ sort_index, int64_t* __restrict__ inverse) {
    __shared__ uint32_t v[LEAF_ORDER_MAX];
    const int64_t l = blockIdx.x;
    if (l >= n_leaves) return;
    const Branch& B = br[leaf_index[l]];
    const int m = (int)(B.n_bodies < LEAF_ORDER_MAX ? B.n_bodies : LEAF_ORDER_MAX);   // the host keeps leaves within
    const int64_t f = B.first;
    for (int i = threadIdx.x; i < m; i += 64) v[i] = idx_sorted[f + i];
    __syncthreads();
    for (int i = threadIdx.x; i < m; i += 64) {
        const uint32_t x = v[i];
        int r = 0;
        for (int j = 0; j < m; ++j) r += v[j] < x ? 1 : 0;
        sort_index[f + r] = (int64_t)x;
        inverse[x] = f + r;
    }
}

__global__ void k_finish_sort(const uint64_t* __restrict__ key2, int64_t n, int64_t* __restrict__ sort_index,
                              int64_t* __restrict__ inverse) {
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    int64_t o = (int64_t)(key2[i] & 0xffffffffull);
    sort_index[i] = o;
    inverse[o] = i;
}

__global__ void k_gather_bodies(const double* __restrict__ buf, int rows, int64_t n, const int64_t* __restrict__ sort_index,
                                bool target, int sdims, double* __restrict__ x, double* __restrict__ y, double* __restrict__ z,
                                double* __restrict__ radius, double* __restrict__ strength) {
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double* b = buf + sort_index[i] * rows;
    x[i] = b[0];
    y[i] = b[1];
    z[i] = b[2];
    if (!target) {
        radius[i] = b[3];
        for (int d = 0; d < sdims; ++d) strength[d * n + i] = b[4 + d];
    }
}

// ---------------------------------------------------------- shrink_recenter!
// one branch from its bodies (a leaf) or its children (src/tree.jl:1141-1188)
__device__ inline void shrink_branch(Branch* br, int64_t ib, const double* __restrict__ x, const double* __restrict__ y,
                                     const double* __restrict__ z, const double* __restrict__ radius, bool use_radius) {
    Branch& b = br[ib];
    double tc[3], tb[3], sc[3], sbx[3], tr = 0.0, sr = 0.0;
    if (b.n_branches == 0) {
        int64_t f = b.first, e = b.first + b.n_bodies;
        double mn[3] = {x[f], y[f], z[f]}, mx[3] = {x[f], y[f], z[f]};
        double smn[3] = {x[f], y[f], z[f]}, smx[3] = {x[f], y[f], z[f]};
        for (int64_t i = f; i < e; ++i) {
            double r = use_radius ? radius[i] : 0.0;
            double p[3] = {x[i], y[i], z[i]};
            for (int d = 0; d < 3; ++d) {
                mn[d] = jmin(mn[d], p[d]);
                mx[d] = jmax(mx[d], p[d]);
                smn[d] = jmin(smn[d], p[d] - r);
                smx[d] = jmax(smx[d], p[d] + r);
            }
        }
        for (int d = 0; d < 3; ++d) {
            tc[d] = (mx[d] + mn[d]) * 0.5;
            tb[d] = mx[d] - tc[d];
            sc[d] = (smx[d] + smn[d]) * 0.5;
            sbx[d] = smx[d] - sc[d];
        }
        for (int64_t i = f; i < e; ++i) {
            double r = use_radius ? radius[i] : 0.0;
            double dx = x[i] - tc[0], dy = y[i] - tc[1], dz = z[i] - tc[2];
            tr = jmax(tr, sqrt(dx * dx + dy * dy + dz * dz));
            dx = x[i] - sc[0]; dy = y[i] - sc[1]; dz = z[i] - sc[2];
            sr = jmax(sr, sqrt(dx * dx + dy * dy + dz * dz) + r);
        }
    } else {
        const Branch& f = br[b.child_first];
        double mn[3], mx[3], smn[3], smx[3];
        for (int d = 0; d < 3; ++d) { mn[d] = mx[d] = f.target_center[d]; smn[d] = smx[d] = f.source_center[d]; }
        for (int64_t c = b.child_first; c < b.child_first + b.n_branches; ++c) {
            const Branch& ch = br[c];
            for (int d = 0; d < 3; ++d) {
                mn[d] = jmin(mn[d], ch.target_center[d] - ch.target_box[d]);
                mx[d] = jmax(mx[d], ch.target_center[d] + ch.target_box[d]);
                smn[d] = jmin(smn[d], ch.source_center[d] - ch.source_box[d]);
                smx[d] = jmax(smx[d], ch.source_center[d] + ch.source_box[d]);
            }
        }
        for (int d = 0; d < 3; ++d) {
            tc[d] = (mn[d] + mx[d]) * 0.5;
            tb[d] = mx[d] - tc[d];
            sc[d] = (smn[d] + smx[d]) * 0.5;
            sbx[d] = smx[d] - sc[d];
        }
        for (int64_t c = b.child_first; c < b.child_first + b.n_branches; ++c) {
            const Branch& ch = br[c];
            double dx = ch.target_center[0] - tc[0], dy = ch.target_center[1] - tc[1], dz = ch.target_center[2] - tc[2];
            tr = jmax(tr, sqrt(dx * dx + dy * dy + dz * dz) + ch.target_radius);
            dx = ch.source_center[0] - sc[0]; dy = ch.source_center[1] - sc[1]; dz = ch.source_center[2] - sc[2];
            sr = jmax(sr, sqrt(dx * dx + dy * dy + dz * dz) + ch.source_radius);
        }
    }
    for (int d = 0; d < 3; ++d) {
        b.target_center[d] = tc[d];
        b.target_box[d] = tb[d];
        b.source_center[d] = sc[d];
        b.source_box[d] = sbx[d];
    }
    b.target_radius = tr;
    b.source_radius = sr;
}

// shrink_recenter! of the whole tree in one launch: a thread per leaf shrinks its leaf, then climbs: the
// thread that finishes a branch's last child shrinks the branch (done[] counts finished children, zero at
// launch). Each branch is computed once, from its finished children, as the level-by-level pass computes it.
__global__ void k_shrink_up(Branch* br, const int64_t* __restrict__ leaf_index, int64_t n_leaves, const double* __restrict__ x,
                            const double* __restrict__ y, const double* __restrict__ z, const double* __restrict__ radius,
                            bool use_radius, int32_t* done) {
    const int64_t l = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (l >= n_leaves) return;
    int64_t ib = leaf_index[l];
    shrink_branch(br, ib, x, y, z, radius, use_radius);
    for (;;) {
        const int64_t pa = br[ib].parent;
        if (pa < 0) return;
        const int32_t last = (int32_t)br[pa].n_branches - 1;
        __threadfence();   // this branch's fields, visible on every XCD before the count
        if (atomicAdd(done + pa, 1) != last) return;
        __threadfence();   // the siblings' fields, written before their counts
        shrink_branch(br, pa, x, y, z, radius, use_radius);
        ib = pa;
    }
}

// the roots, one per system
__global__ void k_init_roots(const Branch* __restrict__ roots, int n_roots, Branch* __restrict__ br) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n_roots) br[i] = roots[i];
}

// ------------------------------------------------------------ host helpers
static inline unsigned grid_for(int64_t n, int block) { return (unsigned)std::max<int64_t>(1, (n + block - 1) / block); }

template <class T>
static void exclusive_scan_i32(Context& C, const T* in, T* out, int64_t n) {
    excl_scan3<T, T>(C, in, out, n);
}

// stable sort by a 63-bit key, within each system's range when seg (device, n_seg + 1 offsets) is given
// (the values are original body indices, so each lies in its system's range)
static void sort_pairs_u64(Context& C, const uint64_t* kin, uint64_t* kout, const uint32_t* vin, uint32_t* vout, int64_t n,
                           const int64_t* seg = nullptr, int n_seg = 1) {
    radix_sort<uint64_t>(C, kin, kout, vin, vout, n, 63, false, seg && n_seg > 1 ? seg : nullptr, n_seg);
}

static void sort_keys_u64(Context& C, const uint64_t* kin, uint64_t* kout, int64_t n) {
    radix_sort<uint64_t>(C, kin, kout, nullptr, nullptr, n, 64);
}

struct TreeScratch {
    DBuf<uint64_t> khi, klo, khi_s, klo_s, key2, key2_s;
    DBuf<uint32_t> idx, idx_s;
    DBuf<int64_t> act_branch, act_first, act_count, sub_first, sub_count;
    DBuf<int64_t> nact_branch, nact_first, nact_count;   // the next level's active branches
    DBuf<int32_t> nonempty, rank, split, srank, flag, frank, done;
    DBuf<int32_t> sys_prefix, sys_flag;   // several systems: per system, the prefix count over the key order
    DBuf<int64_t> sys_off, sys_ls;
    DBuf<double> part, roots;
    DBuf<int64_t> seg;
    DBuf<Branch> root_branches;
    // the build whose keys (khi, klo, idx unsorted and khi_s, klo_s, idx_s sorted) the arrays hold
    const Tree* last = nullptr;
    uint64_t last_id = 0;
    const double* buf = nullptr;
    int rows = 0;
    int64_t n = 0;
    std::vector<int64_t> seg_h;       // the systems' body offsets of that build (empty: one system)
    std::vector<double> roots_h;      // per system: root centre (3), half side, radius
};

static TreeScratch& scratch(Context& C) {   // per context: builds on different streams share nothing
    if (!C.tree_scratch) C.tree_scratch = new TreeScratch();
    return *C.tree_scratch;
}

void free_tree_scratch(TreeScratch* s) { delete s; }

const TreeScratch* tree_keys(const Context& C, const Tree& T) {
    const TreeScratch* S = C.tree_scratch;
    return S && S->last == &T && T.build_id != 0 && S->last_id == T.build_id ? S : nullptr;
}

// Level-by-level structure from sorted keys; returns false if deeper levels than
// max_level would be needed. roots_h: per system the root centre (3), half side and radius.
static bool build_structure(Context& C, Tree& T, TreeScratch& S, const std::vector<int64_t>& seg,
                            const std::vector<double>& roots_h, int64_t ls, int method, int max_level) {
    const int64_t n = T.n;
    const int n_roots = (int)seg.size() - 1;
    // several systems in one tree: the leaf test counts each system (per-system prefix counts over the key
    // order, which the branch ranges index)
    const bool multi = T.sys_off.size() > 2;
    const int n_sys = multi ? (int)T.sys_off.size() - 1 : 1;
    if (multi) {
        if (n_roots != 1 || (int)T.sys_ls.size() != n_sys || T.sys_off.front() != 0 || T.sys_off.back() != n)
            throw Error(FMMHIP_EINVAL, "several systems need one root, offsets covering the bodies and a leaf size each");
        S.sys_off.resize(n_sys + 1);
        S.sys_ls.resize(n_sys);
        HIP_CHECK(hipMemcpyAsync(S.sys_off.data(), T.sys_off.data(), (n_sys + 1) * sizeof(int64_t), hipMemcpyHostToDevice, C.stream));
        HIP_CHECK(hipMemcpyAsync(S.sys_ls.data(), T.sys_ls.data(), n_sys * sizeof(int64_t), hipMemcpyHostToDevice, C.stream));
        S.sys_flag.resize(n + 1);
        S.sys_prefix.resize((size_t)n_sys * (n + 1));
        for (int k = 0; k < n_sys; ++k) {
            k_sys_flag<<<grid_for(n + 1, 256), 256, 0, C.stream>>>(S.idx_s.data(), n, S.sys_off.data(), n_sys, k, S.sys_flag.data());
            LAUNCH_CHECK();
            exclusive_scan_i32(C, S.sys_flag.data(), S.sys_prefix.data() + (size_t)k * (n + 1), n + 1);
        }
    }
    std::vector<Branch> rb(n_roots);
    std::vector<int64_t> act(3 * (size_t)n_roots);   // active roots: branch, first body, count
    int64_t n_active = 0;
    for (int i = 0; i < n_roots; ++i) {
        Branch& r = rb[i];
        r.n_bodies = seg[i + 1] - seg[i]; r.first = seg[i]; r.n_branches = 0; r.child_first = -1; r.parent = -1; r.leaf = -1;
        for (int d = 0; d < 3; ++d) {
            r.source_center[d] = r.target_center[d] = roots_h[5 * i + d];
            r.source_box[d] = r.target_box[d] = roots_h[5 * i + 3];
        }
        r.source_radius = r.target_radius = roots_h[5 * i + 4];
        r.max_influence = 0.0; r.pad = 0.0;
        const bool split = multi ? exceeds_systems(
                                       n_sys, [&](int k) { return T.sys_off[k + 1] - T.sys_off[k]; }, T.sys_ls.data(), method)
                                 : exceeds(r.n_bodies, ls, method);
        if (split) {
            act[n_active] = i;
            act[n_roots + n_active] = r.first;
            act[2 * (size_t)n_roots + n_active] = r.n_bodies;
            ++n_active;
        }
    }
    // capacity: every split node adds at most 8 children; grow on demand
    size_t cap = std::max<size_t>(1024, (size_t)(2 * n / std::max<int64_t>(1, ls) + 64 + n_roots));
    T.branches.resize(cap);
    T.level_first.assign(1, 0);
    T.level_last.assign(1, n_roots - 1);
    int64_t n_branches = n_roots;
    S.root_branches.resize(n_roots);
    HIP_CHECK(hipMemcpyAsync(S.root_branches.data(), rb.data(), n_roots * sizeof(Branch), hipMemcpyHostToDevice, C.stream));
    k_init_roots<<<grid_for(n_roots, 256), 256, 0, C.stream>>>(S.root_branches.data(), n_roots, T.branches.data());
    LAUNCH_CHECK();
    S.act_branch.resize(std::max<int64_t>(1, n_active));
    S.act_first.resize(std::max<int64_t>(1, n_active));
    S.act_count.resize(std::max<int64_t>(1, n_active));
    if (n_active) {
        HIP_CHECK(hipMemcpyAsync(S.act_branch.data(), act.data(), n_active * sizeof(int64_t), hipMemcpyHostToDevice, C.stream));
        HIP_CHECK(hipMemcpyAsync(S.act_first.data(), act.data() + n_roots, n_active * sizeof(int64_t), hipMemcpyHostToDevice,
                                 C.stream));
        HIP_CHECK(hipMemcpyAsync(S.act_count.data(), act.data() + 2 * (size_t)n_roots, n_active * sizeof(int64_t),
                                 hipMemcpyHostToDevice, C.stream));
    }
    HIP_CHECK(hipStreamSynchronize(C.stream));   // rb and act are read by the copies
    int level = 0;
    // One host read-back per level: the children of a level are at most 8 per active branch, so every
    // buffer of the level is sized by that bound; the next level's active branches go to a second set of
    // arrays, swapped in afterwards.
    while (n_active > 0) {
        if (level >= max_level) return false;
        const int64_t m = n_active * 8;
        S.sub_first.resize(m); S.sub_count.resize(m); S.nonempty.resize(m + 1); S.rank.resize(m + 1);
        S.split.resize(m + 1); S.srank.resize(m + 1);
        k_split<<<grid_for(m + 1, 256), 256, 0, C.stream>>>(S.khi_s.data(), S.klo_s.data(), level, n_active, S.act_first.data(),
                                                            S.act_count.data(), S.sub_first.data(), S.sub_count.data(),
                                                            S.nonempty.data(), S.split.data());
        LAUNCH_CHECK();
        exclusive_scan_i32(C, S.nonempty.data(), S.rank.data(), m + 1);   // rank[m]: the level's children
        const int64_t child_base = n_branches;
        if ((size_t)(n_branches + m) > T.branches.cap) {
            DBuf<Branch> bigger;
            bigger.resize(2 * (n_branches + m));
            HIP_CHECK(hipMemcpyAsync(bigger.data(), T.branches.data(), n_branches * sizeof(Branch), hipMemcpyDeviceToDevice, C.stream));
            std::swap(bigger.p, T.branches.p);
            std::swap(bigger.cap, T.branches.cap);
        }
        k_make_children<<<grid_for(m, 256), 256, 0, C.stream>>>(n_active, S.act_branch.data(), S.sub_first.data(),
                                                                S.sub_count.data(), S.rank.data(), child_base, T.branches.data(),
                                                                ls, method, S.split.data(), multi ? S.sys_prefix.data() : nullptr,
                                                                n_sys, n, multi ? S.sys_ls.data() : nullptr);
        LAUNCH_CHECK();
        k_link_parents<<<grid_for(n_active, 256), 256, 0, C.stream>>>(n_active, S.act_branch.data(), S.nonempty.data(),
                                                                      S.rank.data(), child_base, T.branches.data());
        LAUNCH_CHECK();
        exclusive_scan_i32(C, S.split.data(), S.srank.data(), m + 1);   // srank[m]: the next level's active branches
        S.nact_branch.resize(m); S.nact_first.resize(m); S.nact_count.resize(m);
        k_next_active<<<grid_for(m, 256), 256, 0, C.stream>>>(m, S.split.data(), S.srank.data(), child_base, T.branches.data(),
                                                             S.nact_branch.data(), S.nact_first.data(), S.nact_count.data());
        LAUNCH_CHECK();
        int32_t* h = (int32_t*)pinned_scratch(C);
        HIP_CHECK(hipMemcpyAsync(h, S.rank.data() + m, sizeof(int32_t), hipMemcpyDeviceToHost, C.stream));
        HIP_CHECK(hipMemcpyAsync(h + 1, S.srank.data() + m, sizeof(int32_t), hipMemcpyDeviceToHost, C.stream));
        HIP_CHECK(hipStreamSynchronize(C.stream));
        const int64_t n_children = h[0], next_active = h[1];
        std::swap(S.act_branch.p, S.nact_branch.p); std::swap(S.act_branch.cap, S.nact_branch.cap);
        std::swap(S.act_first.p, S.nact_first.p); std::swap(S.act_first.cap, S.nact_first.cap);
        std::swap(S.act_count.p, S.nact_count.p); std::swap(S.act_count.cap, S.nact_count.cap);
        n_branches += n_children;
        T.level_first.push_back(child_base);
        T.level_last.push_back(child_base + n_children - 1);
        n_active = next_active;
        ++level;
    }
    T.n_branches = n_branches;
    T.branches.n = n_branches;
    T.depth = level;
    return true;
}

void build_tree(Context& C, Tree& T, int64_t n, const double* d_buf, int rows, bool target, int64_t ls, int method,
                bool shrink, int p, int sdims, const TreeScratch* keys) {
    T.n = n;
    T.rows = rows;
    T.p = p;
    T.strength_dims = target ? 0 : sdims;
    T.n_branches = 0;
    T.level_first.clear();
    T.level_last.clear();
    T.leaf_index_h.clear();
    if (n == 0) return;
    if (n >= (int64_t)1 << 32) throw Error(FMMHIP_EINVAL, "more than 2^32 bodies in one system");
    if (ls < 1) throw Error(FMMHIP_EINVAL, "leaf size must be positive");
    // the systems' body offsets: T.seg for a forest, else the one range [0, n)
    std::vector<int64_t> seg = T.seg.size() > 2 ? T.seg : std::vector<int64_t>{0, n};
    const int n_seg = (int)seg.size() - 1;
    if (seg.front() != 0 || seg.back() != n) throw Error(FMMHIP_EINVAL, "system offsets do not cover the bodies");
    for (int i = 0; i < n_seg; ++i)
        if (seg[i + 1] <= seg[i]) throw Error(FMMHIP_EINVAL, "every system of a forest needs at least one body");
    TreeScratch& S = scratch(C);
    S.khi.resize(n); S.klo.resize(n); S.idx.resize(n); S.khi_s.resize(n); S.klo_s.resize(n); S.idx_s.resize(n);
    S.seg.resize(n_seg + 1);
    HIP_CHECK(hipMemcpyAsync(S.seg.data(), seg.data(), (n_seg + 1) * sizeof(int64_t), hipMemcpyHostToDevice, C.stream));
    const int64_t* dseg = n_seg > 1 ? S.seg.data() : nullptr;
    std::vector<double> roots_h(5 * n_seg);
    if (keys && keys != &S && keys->n == n && keys->buf == d_buf && keys->rows == rows &&
        (keys->seg_h.empty() ? n_seg == 1 : keys->seg_h == seg)) {
        // the root cubes, the octant paths and their order depend only on the positions: take them
        // from the other context's build of the same buffer (its stream has been waited for)
        roots_h = keys->roots_h;
        auto copy = [&](auto& dst, const auto& src) {
            HIP_CHECK(hipMemcpyAsync(dst.data(), src.data(), (size_t)n * sizeof(*src.data()), hipMemcpyDeviceToDevice, C.stream));
        };
        copy(S.khi, keys->khi); copy(S.klo, keys->klo); copy(S.idx, keys->idx);
        copy(S.khi_s, keys->khi_s); copy(S.klo_s, keys->klo_s); copy(S.idx_s, keys->idx_s);
    } else {
        // root cube of each system (center_box, src/tree.jl:894-907; :20-26)
        const int64_t longest = [&] { int64_t m = 0; for (int i = 0; i < n_seg; ++i) m = std::max(m, seg[i + 1] - seg[i]); return m; }();
        const int nblk = (int)std::min<int64_t>(n_seg > 1 ? 64 : 1024, (longest + 255) / 256);
        S.part.resize((size_t)6 * nblk * n_seg);
        k_bbox_partial<<<dim3(nblk, n_seg), 256, 0, C.stream>>>(d_buf, rows, n, dseg, S.part.data());
        LAUNCH_CHECK();
        std::vector<double> part((size_t)6 * nblk * n_seg);
        HIP_CHECK(hipMemcpyAsync(part.data(), S.part.data(), part.size() * sizeof(double), hipMemcpyDeviceToHost, C.stream));
        HIP_CHECK(hipStreamSynchronize(C.stream));
        std::vector<double> cube(4 * n_seg);
        for (int i = 0; i < n_seg; ++i) {
            const double* q = part.data() + (size_t)6 * nblk * i;
            double mn[3] = {q[0], q[1], q[2]}, mx[3] = {q[3], q[4], q[5]};
            for (int b = 1; b < nblk; ++b)
                for (int d = 0; d < 3; ++d) { mn[d] = jmin(mn[d], q[6 * b + d]); mx[d] = jmax(mx[d], q[6 * b + 3 + d]); }
            double box[3];
            for (int d = 0; d < 3; ++d) {
                roots_h[5 * i + d] = (mx[d] + mn[d]) * 0.5;
                box[d] = mx[d] - roots_h[5 * i + d];
            }
            const double bmax = std::max(std::max(box[0], box[1]), box[2]);
            roots_h[5 * i + 3] = bmax;
            roots_h[5 * i + 4] = std::sqrt(bmax * bmax * 3.0);
            for (int d = 0; d < 4; ++d) cube[4 * i + d] = roots_h[5 * i + d];
        }
        S.roots.resize(4 * n_seg);
        HIP_CHECK(hipMemcpyAsync(S.roots.data(), cube.data(), cube.size() * sizeof(double), hipMemcpyHostToDevice, C.stream));
        // full octant paths, then a stable sort (within each system) by the first 21 levels
        k_path_keys<<<grid_for(n, 256), 256, 0, C.stream>>>(d_buf, rows, n, S.roots.data(), dseg, n_seg, S.khi.data(),
                                                           S.klo.data(), S.idx.data());
        LAUNCH_CHECK();
        sort_pairs_u64(C, S.khi.data(), S.khi_s.data(), S.idx.data(), S.idx_s.data(), n, dseg, n_seg);
        k_gather_u64<<<grid_for(n, 256), 256, 0, C.stream>>>(S.klo.data(), S.idx_s.data(), n, S.klo_s.data());
        LAUNCH_CHECK();
        HIP_CHECK(hipStreamSynchronize(C.stream));   // cube is read by the copy
    }
    static std::atomic<uint64_t> builds{0};
    S.last = &T;
    S.last_id = T.build_id = ++builds;
    S.buf = d_buf;
    S.rows = rows;
    S.n = n;
    S.seg_h = n_seg > 1 ? seg : std::vector<int64_t>();
    S.roots_h = roots_h;
    if (!build_structure(C, T, S, seg, roots_h, ls, method, LEVELS_PER_WORD)) {
        // deeper than 21 levels: order by the second word first, then stably by the first
        DBuf<uint64_t> tmpk;
        DBuf<uint32_t> tmpi;
        tmpk.resize(n); tmpi.resize(n);
        sort_pairs_u64(C, S.klo.data(), tmpk.data(), S.idx.data(), tmpi.data(), n, dseg, n_seg);
        k_gather_u64<<<grid_for(n, 256), 256, 0, C.stream>>>(S.khi.data(), tmpi.data(), n, S.khi_s.data());
        LAUNCH_CHECK();
        // khi_s now holds hi keys in lo-sorted order; stable sort by hi carrying the index
        DBuf<uint64_t> hi2;
        hi2.resize(n);
        sort_pairs_u64(C, S.khi_s.data(), hi2.data(), tmpi.data(), S.idx_s.data(), n, dseg, n_seg);
        HIP_CHECK(hipMemcpyAsync(S.khi_s.data(), hi2.data(), n * sizeof(uint64_t), hipMemcpyDeviceToDevice, C.stream));
        k_gather_u64<<<grid_for(n, 256), 256, 0, C.stream>>>(S.klo.data(), S.idx_s.data(), n, S.klo_s.data());
        LAUNCH_CHECK();
        if (!build_structure(C, T, S, seg, roots_h, ls, method, 2 * LEVELS_PER_WORD))
            throw Error(FMMHIP_EDEPTH, "octree deeper than 42 levels (coincident bodies?)");
    }
    const int64_t nb = T.n_branches;
    // leaf numbering in branch order (src/tree.jl:93-100)
    S.flag.resize(nb + 1); S.frank.resize(nb + 1);
    k_leaf_flags<<<grid_for(nb, 256), 256, 0, C.stream>>>(T.branches.data(), nb, S.flag.data());
    LAUNCH_CHECK();
    exclusive_scan_i32(C, S.flag.data(), S.frank.data(), nb);
    T.n_leaves = read_back_sum(C, S.frank.data() + (nb - 1), S.flag.data() + (nb - 1));
    T.leaf_index.resize(T.n_leaves);
    k_leaf_number<<<grid_for(nb, 256), 256, 0, C.stream>>>(T.branches.data(), nb, S.flag.data(), S.frank.data(), T.leaf_index.data());
    LAUNCH_CHECK();
    // in-leaf order = original index order. A leaf holds at most the largest count the leaf test keeps
    // whole (a branch that exceeds it is split, down to the depth limit, past which the build fails):
    // up to LEAF_ORDER_MAX bodies, each leaf ranks its bodies in LDS; beyond, a sort by (leaf, index)
    T.sort_index.resize(n); T.inverse_sort.resize(n);
    int64_t leaf_max = 0;
    if (T.sys_off.size() > 2) {   // a leaf's per-system terms each stay within the threshold: bound each count
        const double min_ls = (double)*std::min_element(T.sys_ls.begin(), T.sys_ls.end());
        const double thr = method == FMMHIP_SELFTUNING ? 0.5 : 2.8284271247461903;
        for (int64_t l : T.sys_ls)
            leaf_max += method == FMMHIP_BARBA ? l : (int64_t)(thr * (double)l * (double)l / min_ls) + 2;
    } else {
        leaf_max = 1;
        while (leaf_max <= LEAF_ORDER_MAX && !exceeds(leaf_max + 1, ls, method)) ++leaf_max;
    }
    if (leaf_max <= LEAF_ORDER_MAX) {
        k_leaf_order<<<(unsigned)T.n_leaves, 64, 0, C.stream>>>(T.branches.data(), T.leaf_index.data(), T.n_leaves,
                                                                S.idx_s.data(), T.sort_index.data(), T.inverse_sort.data());
        LAUNCH_CHECK();
    } else {
        S.key2.resize(n); S.key2_s.resize(n);
        k_leaf_keys<<<(unsigned)T.n_leaves, 64, 0, C.stream>>>(T.branches.data(), T.leaf_index.data(), T.n_leaves, S.idx_s.data(),
                                                               S.key2.data());
        LAUNCH_CHECK();
        sort_keys_u64(C, S.key2.data(), S.key2_s.data(), n);
        k_finish_sort<<<grid_for(n, 256), 256, 0, C.stream>>>(S.key2_s.data(), n, T.sort_index.data(), T.inverse_sort.data());
        LAUNCH_CHECK();
    }
    // sorted body data (target_to_buffer! / system_to_buffer!, src/compatibility.jl:340-399)
    T.x.resize(n); T.y.resize(n); T.z.resize(n);
    if (!target) { T.radius.resize(n); T.strength.resize((size_t)sdims * n); }
    k_gather_bodies<<<grid_for(n, 256), 256, 0, C.stream>>>(d_buf, rows, n, T.sort_index.data(), target, sdims, T.x.data(),
                                                           T.y.data(), T.z.data(), target ? nullptr : T.radius.data(),
                                                           target ? nullptr : T.strength.data());
    LAUNCH_CHECK();
    if (shrink) {   // bottom-up in one launch
        S.done.resize(nb);
        HIP_CHECK(hipMemsetAsync(S.done.data(), 0, sizeof(int32_t) * nb, C.stream));
        k_shrink_up<<<grid_for(T.n_leaves, 64), 64, 0, C.stream>>>(T.branches.data(), T.leaf_index.data(), T.n_leaves, T.x.data(),
                                                                   T.y.data(), T.z.data(), target ? nullptr : T.radius.data(),
                                                                   !target, S.done.data());
        LAUNCH_CHECK();
    }
    T.expansions.resize((size_t)nb * 4 * n_harmonics(p));
}

// a target tree's source fields are its target fields: shrink_recenter! adds no body radius to
// either in a target tree, so both come out the same
__global__ void k_source_from_target(Branch* __restrict__ br, int64_t nb) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= nb) return;
    Branch& b = br[i];
    for (int d = 0; d < 3; ++d) {
        b.source_center[d] = b.target_center[d];
        b.source_box[d] = b.target_box[d];
    }
    b.source_radius = b.target_radius;
}

// Tree structure, body order, positions and the branches' target centres, boxes and radii
// depend only on the positions, leaf size and options, not on the tree's role, so a target tree
// of the same bodies is the source tree's copy with its source fields replaced by its target
// fields: the tree a target build of those bodies gives, bit for bit.
void clone_target_tree(Context& C, const Tree& s, Tree& t) {
    t.n = s.n;
    t.seg = s.seg;
    t.rows = s.rows;
    t.p = s.p;
    t.strength_dims = 0;
    t.n_branches = s.n_branches;
    t.depth = s.depth;
    t.level_first = s.level_first;
    t.level_last = s.level_last;
    t.leaf_index_h.clear();
    t.n_leaves = s.n_leaves;
    if (s.n == 0) return;
    auto copy = [&](auto& dst, const auto& src, size_t count) {
        dst.resize(count);
        if (count)
            HIP_CHECK(hipMemcpyAsync(dst.data(), src.data(), count * sizeof(*src.data()), hipMemcpyDeviceToDevice, C.stream));
    };
    copy(t.branches, s.branches, (size_t)s.n_branches);
    copy(t.sort_index, s.sort_index, (size_t)s.n);
    copy(t.inverse_sort, s.inverse_sort, (size_t)s.n);
    copy(t.leaf_index, s.leaf_index, (size_t)s.n_leaves);
    copy(t.x, s.x, (size_t)s.n);
    copy(t.y, s.y, (size_t)s.n);
    copy(t.z, s.z, (size_t)s.n);
    t.expansions.resize(s.expansions.size());
    k_source_from_target<<<grid_for(s.n_branches, 256), 256, 0, C.stream>>>(t.branches.data(), s.n_branches);
    LAUNCH_CHECK();
}

}  // namespace fmmhip
