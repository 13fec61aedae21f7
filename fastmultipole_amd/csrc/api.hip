// C ABI of libfmmhip (include/fmmhip.h): contexts, evaluations and the fmm!
// pipeline (src/fmm.jl:862-1176) on one MI355X.
#include "fmmhip_internal.hpp"
#include "scan3.hpp"
#include <memory>
#include <cstring>
#include <algorithm>
#include <cmath>

using namespace fmmhip;

struct fmmhip_ctx { Context c; };
struct fmmhip_eval { Eval e; };

namespace fmmhip {

static inline unsigned grid_for(int64_t n, int block) { return (unsigned)std::max<int64_t>(1, (n + block - 1) / block); }

void ctx_disable_graphs(fmmhip_ctx* ctx) {
    if (ctx) ctx->c.graphs = false;
}

__global__ void k_gather_outputs(const double* __restrict__ x, const double* __restrict__ y, const double* __restrict__ z,
                                 const int64_t* __restrict__ sort_index, const double* __restrict__ tout, int64_t nt,
                                 double* __restrict__ aos) {
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= nt) return;
    double* o = aos + sort_index[i] * 16;
    o[0] = x[i]; o[1] = y[i]; o[2] = z[i];
    for (int k = 0; k < 13; ++k) o[3 + k] = tout[(int64_t)k * nt + i];
}

void gather_outputs(Context& C, const Tree& tt, const double* tout, double* aos, int64_t nt) {
    if (nt == 0) return;
    k_gather_outputs<<<grid_for(nt, 256), 256, 0, C.stream>>>(tt.x.data(), tt.y.data(), tt.z.data(), tt.sort_index.data(), tout,
                                                              nt, aos);
    LAUNCH_CHECK();
}

// influence of each main-tree target from an estimate evaluation (src/fmm.jl:1206-1227)
__global__ void k_map_influence(const int64_t* __restrict__ est_sort, const int64_t* __restrict__ main_inverse,
                                const double* __restrict__ est_tout, int64_t nt, int use_potential, double* __restrict__ infl) {
    int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (j >= nt) return;
    int64_t jm = main_inverse[est_sort[j]];
    if (use_potential) infl[jm] = est_tout[j];
    else {
        double vx = est_tout[nt + j], vy = est_tout[2 * nt + j], vz = est_tout[3 * nt + j];
        infl[jm] = sqrt(vx * vx + vy * vy + vz * vz);
    }
}

// max_influence of every branch over its bodies (src/fmm.jl:1231-1262): one wavefront per leaf takes
// the max over the leaf's bodies and raises the leaf and its ancestors to it. Values are >= +0.0 (the
// reference's max starts from 0.0, and Julia's max(0.0, -0.0) is 0.0), and non-negative doubles order
// like their bit patterns, so the raise is an integer atomic max; the max of a set is exact in any order.
// A walk stops at an ancestor that already holds as much: whichever leaf raised it raises the rest.
__global__ __launch_bounds__(64) void k_branch_max(Branch* __restrict__ br, const int64_t* __restrict__ leaf_index,
                                                   int64_t n_leaves, const double* __restrict__ infl) {
    const int64_t l = blockIdx.x;
    if (l >= n_leaves) return;
    const int64_t ib = leaf_index[l];
    const int64_t f = br[ib].first, e = f + br[ib].n_bodies;
    double m = 0.0;
    for (int64_t i = f + threadIdx.x; i < e; i += 64) m = fmax(m, infl[i]);
    for (int w = 32; w > 0; w >>= 1) m = fmax(m, __shfl_xor(m, w, 64));
    if (threadIdx.x != 0) return;
    const unsigned long long v = (unsigned long long)__double_as_longlong(m > 0.0 ? m : 0.0);
    for (int64_t b = ib; b >= 0; b = br[b].parent) {
        unsigned long long* slot = (unsigned long long*)&br[b].max_influence;
        if (__hip_atomic_load(slot, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= v) break;
        atomicMax(slot, v);
    }
}

// per system of a batch: the largest expansion order its M2L interactions used, and whether each of its
// target boxes met the tolerance; a box belongs to the system holding its first body
__global__ void k_init_sys(int32_t* __restrict__ sys, int n_seg) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n_seg) { sys[2 * i] = 0; sys[2 * i + 1] = 1; }
}

__global__ void k_system_status(const int32_t* __restrict__ m2l_targets, const int64_t* __restrict__ m2l_offsets,
                                int64_t n_boxes, const int32_t* __restrict__ orders, const int32_t* __restrict__ box_ok,
                                const Branch* __restrict__ tbr, const int64_t* __restrict__ seg, int n_seg,
                                int32_t* __restrict__ sys) {
    const int64_t b = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (b >= n_boxes) return;
    const int64_t first = tbr[m2l_targets[b]].first;
    int lo = 0, hi = n_seg - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (seg[mid] <= first) lo = mid; else hi = mid - 1;
    }
    int mx = 0;
    for (int64_t li = m2l_offsets[b]; li < m2l_offsets[b + 1]; ++li) mx = max(mx, orders[li]);
    atomicMax(sys + 2 * lo, mx);
    if (!box_ok[b]) atomicAnd(sys + 2 * lo + 1, 0);
}

// Several systems in one tree: within a leaf the bodies are in original order, so each system's bodies
// (original indices off[s] .. off[s + 1] - 1) are one sub-range of the leaf. Per system, a copy of the
// branches whose leaves hold that sub-range (other branches keep their full range).
__global__ void k_sys_branches(const Branch* __restrict__ br, int64_t nb, const int64_t* __restrict__ sort_index,
                               const int64_t* __restrict__ off, int n_sys, Branch* __restrict__ out) {
    const int64_t b = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (b >= nb) return;
    const Branch B = br[b];
    auto lower = [&](int64_t v) {   // first position of the leaf whose original index is >= v
        int64_t lo = B.first, hi = B.first + B.n_bodies;
        while (lo < hi) {
            const int64_t mid = (lo + hi) >> 1;
            if (sort_index[mid] < v) lo = mid + 1; else hi = mid;
        }
        return lo;
    };
    for (int s = 0; s < n_sys; ++s) {
        Branch Bs = B;
        if (B.n_branches == 0) {
            const int64_t f = lower(off[s]);
            Bs.first = f;
            Bs.n_bodies = lower(off[s + 1]) - f;
        }
        out[(int64_t)s * nb + b] = Bs;
    }
}

// one system's body count of every direct-list source leaf; entry n is 0 (the scan then ends with the total)
__global__ void k_sys_dir_counts(const int32_t* __restrict__ dir_s, int64_t n, const Branch* __restrict__ sbr,
                                 int64_t* __restrict__ cnt) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < n) cnt[i] = sbr[dir_s[i]].n_bodies;
    else if (i == n) cnt[i] = 0;
}

static bool multi(const Eval& E) { return !E.sys_kind.empty(); }
// rows per body of the source buffer, and strength rows of the source tree
static int eval_rows(const Eval& E) { return multi(E) ? E.sys_rows : data_per_body(E.kernel); }
static int eval_sdims(const Eval& E) { return multi(E) ? E.sys_sdims : strength_dims(E.kernel); }

static void sys_branches(Eval& E) {
    if (!multi(E) || E.st.n_branches == 0) return;
    Context& C = *E.ctx;
    const int S = (int)E.sys_kind.size();
    const int64_t nb = E.st.n_branches;
    E.sys_branches.resize((size_t)S * nb);
    k_sys_branches<<<grid_for(nb, 256), 256, 0, C.stream>>>(E.st.branches.data(), nb, E.st.sort_index.data(),
                                                           E.sys_off_dev.data(), S, E.sys_branches.data());
    LAUNCH_CHECK();
}

static void sys_dir_cum(Eval& E) {
    if (!multi(E)) return;
    Context& C = *E.ctx;
    const int S = (int)E.sys_kind.size();
    const int64_t nd = E.lists.n_direct, nb = E.st.n_branches;
    E.sys_dir_cum.resize((size_t)S * (nd + 1));
    if (nd == 0) return;
    DBuf<int64_t> cnt;
    cnt.resize(nd + 1);
    for (int s = 0; s < S; ++s) {
        k_sys_dir_counts<<<grid_for(nd + 1, 256), 256, 0, C.stream>>>(E.lists.dir_s.data(), nd,
                                                                     E.sys_branches.data() + (size_t)s * nb, cnt.data());
        LAUNCH_CHECK();
        excl_scan3<int64_t, int64_t>(C, cnt.data(), E.sys_dir_cum.data() + (size_t)s * (nd + 1), nd + 1);
    }
    HIP_CHECK(hipStreamSynchronize(C.stream));   // cnt is freed on return
}

// P2M of the source tree: once per system of a tuple, with that system's kernel and leaf ranges
static void eval_p2m(Context& F, Eval& E, int p) {
    if (!multi(E)) { run_p2m(F, E.kernel, E.st, p); return; }
    for (size_t s = 0; s < E.sys_kind.size(); ++s)
        run_p2m(F, E.sys_kind[s], E.st, p, E.sys_branches.data() + s * E.st.n_branches);
}

// nearfield: once per source system of a tuple (the reference's order, src/fmm.jl:3-40), each with its kernel
static void eval_p2p(Context& C, Eval& E, bool PS, bool GS, bool HS, int precision, int slots) {
    if (!multi(E)) {
        run_p2p(C, E.kernel, E.tt, E.st, E.lists, E.tout.data(), E.nt, PS, GS, HS, precision, 0, -1, slots);
        return;
    }
    const int64_t nb = E.st.n_branches, nd = E.lists.n_direct;
    for (size_t s = 0; s < E.sys_kind.size(); ++s) {
        SourceView sv;
        sv.branches = E.sys_branches.data() + s * nb;
        sv.dir_cum = E.sys_dir_cum.data() + s * (nd + 1);
        run_p2p(C, E.sys_kind[s], E.tt, E.st, E.lists, E.tout.data(), E.nt, PS, GS, HS, precision, 0, -1, slots, sv);
    }
}

static void check_options(const fmmhip_options* o) {
    if (!o) throw Error(FMMHIP_EINVAL, "options is NULL");
    if (o->expansion_order < 1 || o->expansion_order > 20) throw Error(FMMHIP_EINVAL, "expansion_order must be in 1..20");
    if (o->leaf_size_source < 1 || o->leaf_size_target < 1) throw Error(FMMHIP_EINVAL, "leaf sizes must be positive");
    if (!(o->multipole_acceptance > 0.0)) throw Error(FMMHIP_EINVAL, "multipole_acceptance must be positive");
    if (o->list_method < 0 || o->list_method > 2) throw Error(FMMHIP_EINVAL, "unknown interaction list method");
    if (o->error_kind < 0 || o->error_kind > 3) throw Error(FMMHIP_EINVAL, "unknown error method");
    if (o->precision != FMMHIP_FP64 && o->precision != FMMHIP_MIXED) throw Error(FMMHIP_EINVAL, "unknown precision");
}

struct Timer {
    Context& C;
    hipEvent_t a, b;
    explicit Timer(Context& c) : C(c) {
        HIP_CHECK(hipEventCreate(&a));
        HIP_CHECK(hipEventCreate(&b));
        HIP_CHECK(hipEventRecord(a, C.stream));
    }
    bool marked = false;
    // records the end without waiting for it (read later by stop)
    void mark() {
        HIP_CHECK(hipEventRecord(b, C.stream));
        marked = true;
    }
    double stop() {
        if (!marked) HIP_CHECK(hipEventRecord(b, C.stream));
        HIP_CHECK(hipEventSynchronize(b));
        float ms = 0;
        HIP_CHECK(hipEventElapsedTime(&ms, a, b));
        HIP_CHECK(hipEventDestroy(a));
        HIP_CHECK(hipEventDestroy(b));
        return ms;
    }
};

// several systems in one tree: the leaf test counts each system against its leaf size
static void tree_systems(const Eval& E, Tree& T, int64_t leaf_size) {
    if (!multi(E) || E.sys_kind.size() < 2) {
        T.sys_off.clear();
        T.sys_ls.clear();
        return;
    }
    T.sys_off = E.sys_off;
    T.sys_ls = E.sys_leaf.empty() ? std::vector<int64_t>(E.sys_kind.size(), leaf_size) : E.sys_leaf;
}

static void eval_build_trees(Eval& E, const fmmhip_options& o) {
    Context& C = *E.ctx;
    const int dpb = eval_rows(E);
    const int p = o.expansion_order;
    E.tt.seg = E.seg;
    E.st.seg = E.seg;
    tree_systems(E, E.st, o.leaf_size_source);
    tree_systems(E, E.tt, o.leaf_size_target);
    if (!E.t_xyz && o.leaf_size_target == o.leaf_size_source) {
        // the targets are the sources with one leaf size: the target tree is the source tree's copy
        build_tree(C, E.st, E.ns, E.s_buf, dpb, false, o.leaf_size_source, o.list_method, o.shrink_recenter != 0, p,
                   eval_sdims(E));
        clone_target_tree(C, E.st, E.tt);
    } else {
        if (E.t_xyz) build_tree(C, E.tt, E.nt, E.t_xyz, 3, true, o.leaf_size_target, o.list_method, o.shrink_recenter != 0, p, 0);
        else build_tree(C, E.tt, E.nt, E.s_buf, dpb, true, o.leaf_size_target, o.list_method, o.shrink_recenter != 0, p, 0);
        build_tree(C, E.st, E.ns, E.s_buf, dpb, false, o.leaf_size_source, o.list_method, o.shrink_recenter != 0, p,
                   eval_sdims(E));
    }
    sys_branches(E);
    E.trees_built = true;
    E.lists_built = false;
    E.p_built = p;
    E.keys_fresh = true;
}

static void eval_build_lists(Eval& E, const fmmhip_options& o, bool count_pairs = true) {
    if (!E.trees_built) throw Error(FMMHIP_EINVAL, "trees are not built");
    if (multi(E) && o.list_method == FMMHIP_SELFTUNING)
        throw Error(FMMHIP_EINVAL, "several systems in one tree take SelfTuningTreeStop or Barba lists");
    build_lists(*E.ctx, E.tt, E.st, E.lists, o.leaf_size_source, o.multipole_acceptance, o.list_method, o.farfield != 0,
                o.nearfield != 0, o.self_induced != 0, count_pairs);
    sys_dir_cum(E);
    E.lists_built = true;
}

static void eval_core(Eval& E, const fmmhip_options& o, fmmhip_stats* stats, bool is_estimate);

// releases everything a context holds besides its handle
static void release_context(Context& c) {
    if (c.stream) (void)hipStreamSynchronize(c.stream);
    free_tables(c.tables);
    free_tree_scratch(c.tree_scratch);
    free_list_scratch(c.list_scratch);
    c.tables = nullptr;
    c.tree_scratch = nullptr;
    c.list_scratch = nullptr;
    if (c.pinned) (void)hipHostFree(c.pinned);
    c.pinned = nullptr;
    if (c.stream) (void)hipStreamDestroy(c.stream);
    c.stream = nullptr;
}

Eval::~Eval() {
    delete estimate;
    free_fgs(fgs);
    if (side) {
        release_context(*side);
        delete side;
    }
}

// the evaluation's second stream, with its own scratch and tables
static Context& side_context(Eval& E) {
    Context*& slot = E.side;
    if (!slot) {
        Context* c = new Context();
        c->device = E.ctx->device;
        if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
            delete c;
            throw Error(FMMHIP_EDEVICE, "hipStreamCreateWithFlags failed");
        }
        slot = c;
    }
    return *slot;
}

// makes `waiter` wait for the work enqueued on `signaller` so far
static void stream_wait(Context& waiter, Context& signaller) {
    hipEvent_t e;
    HIP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    const hipError_t a = hipEventRecord(e, signaller.stream);
    const hipError_t b = a == hipSuccess ? hipStreamWaitEvent(waiter.stream, e, 0) : a;
    (void)hipEventDestroy(e);
    HIP_CHECK(b);
}

// estimate_influence! for relative tolerances (src/fmm.jl:1193-1276): a full
// low-order evaluation (p = 3, leaf 5, MAC 0.6, potential and gradient) on its
// own trees; max |grad| (or potential) per main-tree branch. In two parts: the
// trees and lists (many small, host-synchronised launches), then the evaluation.
static fmmhip_options estimate_options(const fmmhip_options& o) {
    fmmhip_options eo = o;
    eo.scalar_potential = 1; eo.gradient = 1; eo.hessian = 0;
    eo.leaf_size_source = 5; eo.leaf_size_target = 5;
    eo.expansion_order = 3; eo.multipole_acceptance = 0.6;
    eo.error_kind = FMMHIP_ERR_NONE; eo.error_relative = 0;
    eo.shrink_recenter = 1;
    eo.list_method = FMMHIP_SELFTUNING_TREESTOP;
    eo.farfield = eo.nearfield = eo.self_induced = 1;
    return eo;
}

static void estimate_build(Eval& E, const fmmhip_options& o, Context& C) {
    if (!E.estimate) E.estimate = new Eval();
    Eval& X = *E.estimate;
    X.ctx = &C;
    X.kernel = E.kernel;
    X.nt = E.nt;
    X.ns = E.ns;
    X.t_xyz = E.t_xyz;
    X.s_buf = E.s_buf;
    X.seg = E.seg;
    X.st.seg = E.seg;
    X.sys_kind = E.sys_kind;
    X.sys_off = E.sys_off;
    X.sys_rows = E.sys_rows;
    X.sys_sdims = E.sys_sdims;
    X.sys_leaf.clear();   // estimate_influence: leaf size 5 for every system
    if (multi(E)) {
        X.sys_off_dev.resize(E.sys_off.size());
        HIP_CHECK(hipMemcpyAsync(X.sys_off_dev.data(), E.sys_off_dev.data(), E.sys_off.size() * sizeof(int64_t),
                                 hipMemcpyDeviceToDevice, C.stream));
    }
    const fmmhip_options eo = estimate_options(o);
    if (X.t_xyz) eval_build_trees(X, eo);
    else {   // the targets are the sources: one build gives both trees; the Morton keys and their
             // order are the main source tree's (same buffer), built on the evaluation's stream
        tree_systems(X, X.st, eo.leaf_size_source);
        build_tree(C, X.st, X.ns, X.s_buf, eval_rows(X), false, eo.leaf_size_source, eo.list_method, true,
                   eo.expansion_order, eval_sdims(X), E.keys_fresh ? tree_keys(*E.ctx, E.st) : nullptr);
        clone_target_tree(C, X.st, X.tt);
        sys_branches(X);
        X.trees_built = true;
        X.lists_built = false;
        X.p_built = eo.expansion_order;
    }
    eval_build_lists(X, eo, false);
}

static void estimate_eval(Eval& E, const fmmhip_options& o, Context& C) {
    Eval& X = *E.estimate;
    eval_core(X, estimate_options(o), nullptr, true);
    DBuf<double>& infl = E.influence;
    infl.resize(std::max<int64_t>(1, E.nt));
    k_map_influence<<<grid_for(E.nt, 256), 256, 0, C.stream>>>(X.tt.sort_index.data(), E.tt.inverse_sort.data(), X.tout.data(),
                                                              E.nt, o.error_kind == FMMHIP_ERR_POWER_POTENTIAL ? 1 : 0,
                                                              infl.data());
    LAUNCH_CHECK();
    k_branch_max<<<(unsigned)std::max<int64_t>(1, E.tt.n_leaves), 64, 0, C.stream>>>(E.tt.branches.data(), E.tt.leaf_index.data(),
                                                                                   E.tt.n_leaves, infl.data());
    LAUNCH_CHECK();
}

static void eval_core(Eval& E, const fmmhip_options& o, fmmhip_stats* stats, bool is_estimate) {
    Context& C = *E.ctx;
    if (!E.lists_built) throw Error(FMMHIP_EINVAL, "interaction lists are not built");
    const int p = o.expansion_order;
    if (p != E.p_built) throw Error(FMMHIP_EINVAL, "expansion_order differs from the one the trees were built with");
    const bool LH = has_vector_potential(E.kernel);
    const bool PS = o.scalar_potential != 0, GS = o.gradient != 0, HS = o.hessian != 0;
    Timer total(C);
    double t_est = 0, t_gather = 0;
    E.tout.resize((size_t)13 * std::max<int64_t>(1, E.nt));
    HIP_CHECK(hipMemsetAsync(E.tout.data(), 0, E.tout.size() * sizeof(double), C.stream));
    if (E.tt.n_branches) HIP_CHECK(hipMemsetAsync(E.tt.expansions.data(), 0, E.tt.expansions.size() * sizeof(double), C.stream));
    if (E.st.n_branches) HIP_CHECK(hipMemsetAsync(E.st.expansions.data(), 0, E.st.expansions.size() * sizeof(double), C.stream));
    E.pmax = 0;
    E.error_success = 1;
    double t_p2p = 0, t_up = 0, t_m2l = 0, t_down = 0;
    if (E.nt > 0 && E.ns > 0) {
        // The nearfield (P2P) runs on the evaluation's stream while the far field
        // (estimate_influence, upward pass, M2L, L2L) runs on a second stream; L2P adds
        // into the nearfield's results once both are done. The nearfield's launch fills the
        // device, so far-field phase times include waiting for it (the estimate's small
        // launches mostly run after it); the overlap takes the host synchronisations of the
        // estimate's tree and list builds off the critical path. The estimate's own
        // evaluation runs on one stream.
        Context& F = is_estimate ? C : side_context(E);
        ensure_tables(C, p);
        if (&F != &C) {
            ensure_tables(F, p);
            stream_wait(F, C);
        }
        const bool estimate = !is_estimate && o.error_kind != FMMHIP_ERR_NONE && o.error_relative;
        Timer tp(C);
        // Beside the far-field stream the nearfield leaves one workgroup slot per CU free: its
        // rate is about the same at 3 and 4 workgroups per CU, and the far field's kernels (the
        // estimate's tree, list and sort launches) are placed at once instead of waiting for
        // the nearfield's last workgroups.
        const int slots = &F != &C ? -C.p2p_free_slots : 0;
        eval_p2p(C, E, PS, GS, HS, o.precision, slots);
        tp.mark();
        // With estimate_influence on the far-field stream, the upward pass follows the nearfield on its
        // stream: it needs nothing of the estimate, so it runs beside the estimate's evaluation once the
        // nearfield is done (103.0 against 104.9 ms per evaluation of the 1M ring with it on the far-field
        // stream after the estimate; the estimate's own nearfield and far field on two streams gained
        // nothing; 102.3 against 100.9 ms with it before the nearfield, beside the estimate's trees).
        // Without the estimate it stays on the far-field stream, beside the nearfield.
        const bool up_on_c = estimate && &F != &C;
        std::unique_ptr<Timer> tu;
        if (up_on_c) {
            tu.reset(new Timer(C));
            eval_p2m(C, E, p);
            run_m2m(C, E.st, p, LH);
            tu->mark();
        }
        if (estimate) {
            // estimate_influence (trees, lists, evaluation) on the far-field stream beside the nearfield.
            // Measured alternatives: its trees and lists first, alone on the device (110.1 against 107.5 ms
            // per evaluation of the 1M ring: its evaluation then waits for the nearfield), and the upward
            // pass on a third stream beside it (109.0 ms: both then share the one free slot per CU).
            Timer t(F);
            estimate_build(E, o, F);
            estimate_eval(E, o, F);
            t_est = t.stop();
        }
        if (!up_on_c) {
            Timer t(F);
            eval_p2m(F, E, p);
            run_m2m(F, E.st, p, LH);
            t_up = t.stop();
        } else {
            stream_wait(F, C);
        }
        {
            Timer t(F);
            int32_t pmax = 0, ok = 1;
            const bool batch = !is_estimate && !E.seg.empty();
            E.lists.track_box_ok = batch;
            if (batch) {
                const int n_seg = (int)E.seg.size() - 1;
                E.sys_status.resize(2 * n_seg);
                k_init_sys<<<grid_for(n_seg, 256), 256, 0, F.stream>>>(E.sys_status.data(), n_seg);
                LAUNCH_CHECK();
            }
            run_m2l(F, E.tt, E.st, E.tt, E.lists, p, LH, o.error_kind, o.error_relative != 0, o.error_tolerance,
                    o.error_bonus != 0, o.precision, batch ? nullptr : &pmax, &ok);
            if (batch) {
                if (E.lists.n_m2l_targets > 0) {
                    k_system_status<<<grid_for(E.lists.n_m2l_targets, 256), 256, 0, F.stream>>>(
                        E.lists.m2l_targets.data(), E.lists.m2l_offsets.data(), E.lists.n_m2l_targets, E.lists.m2l_order.data(),
                        E.lists.m2l_box_ok.data(), E.tt.branches.data(), E.seg_dev.data(), (int)E.seg.size() - 1,
                        E.sys_status.data());
                    LAUNCH_CHECK();
                }
                const int n_seg = (int)E.seg.size() - 1;
                std::vector<int32_t> h(2 * n_seg);
                HIP_CHECK(hipMemcpyAsync(h.data(), E.sys_status.data(), h.size() * sizeof(int32_t), hipMemcpyDeviceToHost,
                                         F.stream));
                HIP_CHECK(hipStreamSynchronize(F.stream));
                E.sys_pmax.assign(n_seg, 0);
                E.sys_ok.assign(n_seg, 1);
                for (int i = 0; i < n_seg; ++i) {
                    E.sys_ok[i] = h[2 * i + 1];
                    E.sys_pmax[i] = h[2 * i] + (h[2 * i + 1] ? 0 : 1);   // as the evaluation's pmax below
                    pmax = std::max(pmax, h[2 * i]);
                    ok = ok && h[2 * i + 1];
                }
            }
            E.pmax = pmax;
            E.error_success = ok;
            if (!ok) E.pmax += 1;
            t_m2l = t.stop();
        }
        {
            Timer t(F);
            run_l2l(F, E.tt, p, LH);
            t_down = t.stop();
        }
        if (tu) t_up = tu->stop();
        if (&F != &C) stream_wait(C, F);
        {
            Timer t(C);
            run_l2p(C, E.tt, E.tout.data(), E.nt, p, LH, PS, GS, HS);
            t_down += t.stop();
        }
        t_p2p = tp.stop();
    }
    if (!is_estimate) {
        Timer t(C);
        E.tout_aos.resize((size_t)16 * std::max<int64_t>(1, E.nt));
        gather_outputs(C, E.tt, E.tout.data(), E.tout_aos.data(), E.nt);
        t_gather = t.stop();
    }
    const double t_tot = total.stop();
    E.ran = true;
    if (!is_estimate) E.keys_fresh = false;
    if (stats) {
        stats->pmax = E.pmax;
        stats->error_success = E.error_success;
        stats->n_target_branches = E.tt.n_branches;
        stats->n_source_branches = E.st.n_branches;
        stats->n_m2l = E.lists.n_m2l;
        stats->n_direct = E.lists.n_direct;
        stats->n_p2p_pairs = E.lists.n_pairs;
        stats->ms_p2p = t_p2p;
        stats->ms_upward = t_up;
        stats->ms_m2l = t_m2l;
        stats->ms_downward = t_down;
        stats->ms_total = t_tot;
        stats->ms_estimate = t_est;
        stats->ms_gather = t_gather;
    }
}

}  // namespace fmmhip

// ====================================================================== ABI
#define ABI_TRY(ctxptr, ...)                                             \
    try {                                                                \
        __VA_ARGS__;                                                     \
        return FMMHIP_OK;                                                \
    } catch (const Error& err) {                                         \
        if (ctxptr) (ctxptr)->last_error = err.what();                   \
        return err.status;                                               \
    } catch (const std::exception& err) {                                \
        if (ctxptr) (ctxptr)->last_error = err.what();                   \
        return FMMHIP_EDEVICE;                                           \
    }

extern "C" {

void fmmhip_options_default(fmmhip_options* o) {
    if (!o) return;
    std::memset(o, 0, sizeof(*o));
    o->expansion_order = 5;
    o->leaf_size_source = 20;
    o->leaf_size_target = 20;
    o->multipole_acceptance = 0.4;
    o->list_method = FMMHIP_SELFTUNING_TREESTOP;
    o->shrink_recenter = 1;
    o->error_kind = FMMHIP_ERR_NONE;
    o->error_relative = 0;
    o->error_tolerance = 0.0;
    o->error_bonus = 1;
    o->scalar_potential = 0;
    o->gradient = 1;
    o->hessian = 0;
    o->farfield = o->nearfield = o->self_induced = 1;
    o->precision = FMMHIP_FP64;
}

fmmhip_status fmmhip_ctx_create(int device, fmmhip_ctx** out) {
    if (!out) return FMMHIP_EINVAL;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return FMMHIP_EDEVICE;
    fmmhip_ctx* h = new fmmhip_ctx();
    try {
        if (device >= 0) HIP_CHECK(hipSetDevice(device));
        HIP_CHECK(hipGetDevice(&h->c.device));
        HIP_CHECK(hipStreamCreateWithFlags(&h->c.stream, hipStreamNonBlocking));
    } catch (const Error& e) {
        delete h;
        return e.status;
    }
    *out = h;
    return FMMHIP_OK;
}

void fmmhip_ctx_destroy(fmmhip_ctx* h) {
    if (!h) return;
    release_context(h->c);
    delete h;
}

const char* fmmhip_last_error(const fmmhip_ctx* h) { return h ? h->c.last_error.c_str() : "null context"; }

fmmhip_status fmmhip_sync(fmmhip_ctx* h) {
    if (!h) return FMMHIP_EINVAL;
    Context* C = &h->c;
    ABI_TRY(C, HIP_CHECK(hipStreamSynchronize(C->stream)));
}

fmmhip_status fmmhip_ctx_set_p2p_free_slots(fmmhip_ctx* h, int32_t slots) {
    if (!h) return FMMHIP_EINVAL;
    Context* C = &h->c;
    ABI_TRY(C, {
        if (slots < 0 || slots > 3) throw Error(FMMHIP_EINVAL, "free slots per compute unit must be 0..3");
        C->p2p_free_slots = slots;
    });
}

fmmhip_status fmmhip_eval_create(fmmhip_ctx* h, fmmhip_eval** out) {
    if (!h || !out) return FMMHIP_EINVAL;
    fmmhip_eval* e = new fmmhip_eval();
    e->e.ctx = &h->c;
    *out = e;
    return FMMHIP_OK;
}

void fmmhip_eval_destroy(fmmhip_eval* e) {
    if (!e) return;
    (void)hipStreamSynchronize(e->e.ctx->stream);
    delete e;
}

fmmhip_status fmmhip_eval_set_systems(fmmhip_eval* ev, int32_t kernel, int64_t nt, const double* t_xyz, int64_t ns,
                                      const double* s_buf, int32_t on_device) {
    if (!ev) return FMMHIP_EINVAL;
    Eval& E = ev->e;
    Context* C = E.ctx;
    ABI_TRY(C, {
        if (!valid_kernel(kernel)) throw Error(FMMHIP_EINVAL, "unknown kernel");
        if (nt < 0 || ns < 0) throw Error(FMMHIP_EINVAL, "negative system size");
        if ((ns > 0 && !s_buf) || (nt > 0 && !t_xyz && nt != ns)) throw Error(FMMHIP_EINVAL, "missing system buffer");
        E.kernel = kernel;
        E.nt = nt;
        E.ns = ns;
        E.seg.clear();
        E.sys_pmax.clear();
        E.sys_ok.clear();
        E.sys_kind.clear();
        E.sys_off.clear();
        const int dpb = data_per_body(kernel);
        if (on_device) {
            E.t_xyz = t_xyz;
            E.s_buf = s_buf;
        } else {
            E.s_buf_in.resize(std::max<int64_t>(1, (int64_t)dpb * ns));
            if (ns) HIP_CHECK(hipMemcpyAsync(E.s_buf_in.data(), s_buf, sizeof(double) * dpb * ns, hipMemcpyHostToDevice, C->stream));
            E.s_buf = E.s_buf_in.data();
            if (t_xyz) {
                E.t_xyz_in.resize(std::max<int64_t>(1, 3 * nt));
                if (nt) HIP_CHECK(hipMemcpyAsync(E.t_xyz_in.data(), t_xyz, sizeof(double) * 3 * nt, hipMemcpyHostToDevice, C->stream));
                E.t_xyz = E.t_xyz_in.data();
            } else E.t_xyz = nullptr;
            HIP_CHECK(hipStreamSynchronize(C->stream));
        }
        E.trees_built = E.lists_built = E.ran = false;
    });
}

fmmhip_status fmmhip_eval_set_batch(fmmhip_eval* ev, int32_t kernel, int64_t n_systems, const int64_t* counts,
                                     const double* s_buf, int32_t on_device) {
    if (!ev) return FMMHIP_EINVAL;
    Eval& E = ev->e;
    Context* C = E.ctx;
    ABI_TRY(C, {
        if (!valid_kernel(kernel)) throw Error(FMMHIP_EINVAL, "unknown kernel");
        if (n_systems < 1 || !counts) throw Error(FMMHIP_EINVAL, "a batch needs at least one system");
        std::vector<int64_t> seg(n_systems + 1, 0);
        for (int64_t i = 0; i < n_systems; ++i) {
            if (counts[i] < 1) throw Error(FMMHIP_EINVAL, "every system of a batch needs at least one body");
            seg[i + 1] = seg[i] + counts[i];
        }
        const int64_t n = seg[n_systems];
        if (!s_buf) throw Error(FMMHIP_EINVAL, "missing system buffer");
        if (n >= (int64_t)1 << 31) throw Error(FMMHIP_EINVAL, "more than 2^31 bodies in one batch");
        E.kernel = kernel;
        E.nt = E.ns = n;
        E.t_xyz = nullptr;
        E.sys_kind.clear();
        E.sys_off.clear();
        const int dpb = data_per_body(kernel);
        if (on_device) E.s_buf = s_buf;
        else {
            E.s_buf_in.resize((int64_t)dpb * n);
            HIP_CHECK(hipMemcpyAsync(E.s_buf_in.data(), s_buf, sizeof(double) * dpb * n, hipMemcpyHostToDevice, C->stream));
            HIP_CHECK(hipStreamSynchronize(C->stream));
            E.s_buf = E.s_buf_in.data();
        }
        E.seg = n_systems > 1 ? seg : std::vector<int64_t>();
        if (n_systems > 1) {
            E.seg_dev.resize(n_systems + 1);
            HIP_CHECK(hipMemcpyAsync(E.seg_dev.data(), seg.data(), seg.size() * sizeof(int64_t), hipMemcpyHostToDevice,
                                     C->stream));
            HIP_CHECK(hipStreamSynchronize(C->stream));
        }
        E.sys_pmax.clear();
        E.sys_ok.clear();
        E.trees_built = E.lists_built = E.ran = false;
    });
}

fmmhip_status fmmhip_eval_set_tuple(fmmhip_eval* ev, int64_t n_systems, const int32_t* kernels, const int64_t* counts,
                                     const double* buffer, int32_t on_device) {
    if (!ev) return FMMHIP_EINVAL;
    Eval& E = ev->e;
    Context* C = E.ctx;
    ABI_TRY(C, {
        if (n_systems < 1 || !kernels || !counts) throw Error(FMMHIP_EINVAL, "a tuple needs at least one system");
        std::vector<int64_t> off(n_systems + 1, 0);
        bool lh = false;
        int rows = 7;
        for (int64_t i = 0; i < n_systems; ++i) {
            const int k = kernels[i];
            if (!valid_kernel(k)) throw Error(FMMHIP_EINVAL, "unknown kernel in a tuple");
            if (counts[i] < 0) throw Error(FMMHIP_EINVAL, "negative system size");
            lh = lh || has_vector_potential(k);
            rows = std::max(rows, data_per_body(k));
            off[i + 1] = off[i] + counts[i];
        }
        const int64_t n = off[n_systems];
        if (n > 0 && !buffer) throw Error(FMMHIP_EINVAL, "missing system buffer");
        if (n >= (int64_t)1 << 31) throw Error(FMMHIP_EINVAL, "more than 2^31 bodies in one tuple");
        // the tree's bodies carry the largest row count of the systems, [x, y, z, radius, strength rows...]
        // (7 for point systems, 14 with triangles, 17 with quadrilaterals); the kernel tag switches the
        // Lamb-Helmholtz terms on if any system is a vortex system
        E.kernel = lh ? FMMHIP_POINT_VORTEX : FMMHIP_POINT_DIPOLE;
        E.sys_rows = rows;
        E.sys_sdims = rows - 4;
        E.nt = E.ns = n;
        E.t_xyz = nullptr;
        E.seg.clear();
        E.sys_pmax.clear();
        E.sys_ok.clear();
        if (on_device) E.s_buf = buffer;
        else {
            E.s_buf_in.resize(std::max<int64_t>(1, rows * n));
            if (n) HIP_CHECK(hipMemcpyAsync(E.s_buf_in.data(), buffer, sizeof(double) * rows * n, hipMemcpyHostToDevice, C->stream));
            E.s_buf = E.s_buf_in.data();
        }
        E.sys_kind.assign(kernels, kernels + n_systems);
        E.sys_off = off;
        E.sys_leaf.clear();
        E.sys_off_dev.resize(off.size());
        HIP_CHECK(hipMemcpyAsync(E.sys_off_dev.data(), off.data(), off.size() * sizeof(int64_t), hipMemcpyHostToDevice, C->stream));
        HIP_CHECK(hipStreamSynchronize(C->stream));
        E.trees_built = E.lists_built = E.ran = false;
    });
}

fmmhip_status fmmhip_eval_set_tuple_leaf_sizes(fmmhip_eval* ev, int64_t n_systems, const int64_t* leaf_sizes) {
    if (!ev) return FMMHIP_EINVAL;
    Eval& E = ev->e;
    Context* C = E.ctx;
    ABI_TRY(C, {
        if (!multi(E)) throw Error(FMMHIP_EINVAL, "leaf sizes per system need a tuple (fmmhip_eval_set_tuple)");
        if (n_systems != (int64_t)E.sys_kind.size() || !leaf_sizes) throw Error(FMMHIP_EINVAL, "one leaf size per system");
        for (int64_t i = 0; i < n_systems; ++i)
            if (leaf_sizes[i] < 1) throw Error(FMMHIP_EINVAL, "leaf sizes must be positive");
        E.sys_leaf.assign(leaf_sizes, leaf_sizes + n_systems);
        E.trees_built = E.lists_built = E.ran = false;
    });
}

fmmhip_status fmmhip_eval_get_batch_stats(fmmhip_eval* ev, int32_t* pmax, int32_t* error_success) {
    if (!ev) return FMMHIP_EINVAL;
    Eval& E = ev->e;
    Context* C = E.ctx;
    ABI_TRY(C, {
        if (!E.ran) throw Error(FMMHIP_EINVAL, "the evaluation has not run");
        if (E.seg.empty()) {
            if (pmax) pmax[0] = E.pmax;
            if (error_success) error_success[0] = E.error_success;
        } else {
            if (E.sys_pmax.size() + 1 != E.seg.size()) throw Error(FMMHIP_EINVAL, "no per-system results");
            if (pmax) std::copy(E.sys_pmax.begin(), E.sys_pmax.end(), pmax);
            if (error_success) std::copy(E.sys_ok.begin(), E.sys_ok.end(), error_success);
        }
    });
}

fmmhip_status fmmhip_eval_build_trees(fmmhip_eval* ev, const fmmhip_options* o) {
    if (!ev) return FMMHIP_EINVAL;
    Context* C = ev->e.ctx;
    ABI_TRY(C, {
        check_options(o);
        eval_build_trees(ev->e, *o);
        HIP_CHECK(hipStreamSynchronize(C->stream));
    });
}

fmmhip_status fmmhip_eval_build_lists(fmmhip_eval* ev, const fmmhip_options* o) {
    if (!ev) return FMMHIP_EINVAL;
    Context* C = ev->e.ctx;
    ABI_TRY(C, {
        check_options(o);
        eval_build_lists(ev->e, *o);
        HIP_CHECK(hipStreamSynchronize(C->stream));
    });
}

fmmhip_status fmmhip_eval_run(fmmhip_eval* ev, const fmmhip_options* o, fmmhip_stats* stats) {
    if (!ev) return FMMHIP_EINVAL;
    Context* C = ev->e.ctx;
    ABI_TRY(C, {
        check_options(o);
        eval_core(ev->e, *o, stats, false);
        HIP_CHECK(hipStreamSynchronize(C->stream));
    });
}

fmmhip_status fmmhip_fmm(fmmhip_eval* ev, const fmmhip_options* o, fmmhip_stats* stats) {
    if (!ev) return FMMHIP_EINVAL;
    Context* C = ev->e.ctx;
    ABI_TRY(C, {
        check_options(o);
        Timer t(*C);
        double t_tree, t_lists;
        {
            Timer a(*C);
            eval_build_trees(ev->e, *o);
            t_tree = a.stop();
        }
        {
            Timer a(*C);
            eval_build_lists(ev->e, *o);
            t_lists = a.stop();
        }
        eval_core(ev->e, *o, stats, false);
        double tt = t.stop();
        if (stats) {
            stats->ms_tree = t_tree;
            stats->ms_lists = t_lists;
            stats->ms_total = tt;
        }
    });
}

fmmhip_status fmmhip_eval_get_targets(fmmhip_eval* ev, double* out, int32_t on_device) {
    if (!ev || !out) return FMMHIP_EINVAL;
    Eval& E = ev->e;
    Context* C = E.ctx;
    ABI_TRY(C, {
        if (!E.ran) throw Error(FMMHIP_EINVAL, "no evaluation has run");
        if (E.nt)
            HIP_CHECK(hipMemcpyAsync(out, E.tout_aos.data(), sizeof(double) * 16 * E.nt,
                                     on_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, C->stream));
        HIP_CHECK(hipStreamSynchronize(C->stream));
    });
}

static Tree& which_tree(Eval& E, int32_t which) { return which == 0 ? E.tt : E.st; }

int64_t fmmhip_eval_n_branches(fmmhip_eval* ev, int32_t which) { return ev ? which_tree(ev->e, which).n_branches : -1; }
int64_t fmmhip_eval_n_levels(fmmhip_eval* ev, int32_t which) { return ev ? (int64_t)which_tree(ev->e, which).level_first.size() : -1; }
int64_t fmmhip_eval_n_leaves(fmmhip_eval* ev, int32_t which) { return ev ? which_tree(ev->e, which).n_leaves : -1; }

fmmhip_status fmmhip_eval_get_branches(fmmhip_eval* ev, int32_t which, int64_t* ints, double* reals) {
    if (!ev || !ints || !reals) return FMMHIP_EINVAL;
    Context* C = ev->e.ctx;
    ABI_TRY(C, {
        Tree& T = which_tree(ev->e, which);
        std::vector<Branch> h(T.n_branches);
        if (T.n_branches)
            HIP_CHECK(hipMemcpy(h.data(), T.branches.data(), sizeof(Branch) * T.n_branches, hipMemcpyDeviceToHost));
        for (int64_t i = 0; i < T.n_branches; ++i) {
            const Branch& b = h[i];
            int64_t* I = ints + 6 * i;
            I[0] = b.n_bodies; I[1] = b.first; I[2] = b.n_branches; I[3] = b.child_first; I[4] = b.parent; I[5] = b.leaf;
            double* D = reals + 17 * i;
            for (int d = 0; d < 3; ++d) {
                D[d] = b.source_center[d]; D[3 + d] = b.target_center[d];
                D[8 + d] = b.source_box[d]; D[11 + d] = b.target_box[d];
            }
            D[6] = b.source_radius; D[7] = b.target_radius; D[14] = b.max_influence; D[15] = 0; D[16] = 0;
        }
    });
}

fmmhip_status fmmhip_eval_get_levels(fmmhip_eval* ev, int32_t which, int64_t* out) {
    if (!ev || !out) return FMMHIP_EINVAL;
    Tree& T = which_tree(ev->e, which);
    for (size_t i = 0; i < T.level_first.size(); ++i) { out[2 * i] = T.level_first[i]; out[2 * i + 1] = T.level_last[i]; }
    return FMMHIP_OK;
}

fmmhip_status fmmhip_eval_get_leaf_index(fmmhip_eval* ev, int32_t which, int64_t* out) {
    if (!ev || !out) return FMMHIP_EINVAL;
    Context* C = ev->e.ctx;
    ABI_TRY(C, {
        Tree& T = which_tree(ev->e, which);
        if (T.n_leaves) HIP_CHECK(hipMemcpy(out, T.leaf_index.data(), sizeof(int64_t) * T.n_leaves, hipMemcpyDeviceToHost));
    });
}

fmmhip_status fmmhip_eval_get_sort_index(fmmhip_eval* ev, int32_t which, int64_t* out) {
    if (!ev || !out) return FMMHIP_EINVAL;
    Context* C = ev->e.ctx;
    ABI_TRY(C, {
        Tree& T = which_tree(ev->e, which);
        if (T.n) HIP_CHECK(hipMemcpy(out, T.sort_index.data(), sizeof(int64_t) * T.n, hipMemcpyDeviceToHost));
    });
}

int64_t fmmhip_eval_n_list(fmmhip_eval* ev, int32_t which) {
    if (!ev) return -1;
    return which == 0 ? ev->e.lists.n_m2l : ev->e.lists.n_direct;
}

fmmhip_status fmmhip_eval_get_list(fmmhip_eval* ev, int32_t which, int32_t* out) {
    if (!ev || !out) return FMMHIP_EINVAL;
    Context* C = ev->e.ctx;
    ABI_TRY(C, {
        Lists& L = ev->e.lists;
        int64_t n = which == 0 ? L.n_m2l : L.n_direct;
        if (n) {
            std::vector<int32_t> t(n), s(n);
            HIP_CHECK(hipMemcpy(t.data(), which == 0 ? L.m2l_t.data() : L.dir_t.data(), 4 * n, hipMemcpyDeviceToHost));
            HIP_CHECK(hipMemcpy(s.data(), which == 0 ? L.m2l_s.data() : L.dir_s.data(), 4 * n, hipMemcpyDeviceToHost));
            for (int64_t i = 0; i < n; ++i) { out[2 * i] = t[i]; out[2 * i + 1] = s[i]; }
        }
    });
}

fmmhip_status fmmhip_eval_get_m2l_orders(fmmhip_eval* ev, int32_t* out) {
    if (!ev || !out) return FMMHIP_EINVAL;
    Context* C = ev->e.ctx;
    ABI_TRY(C, {
        Lists& L = ev->e.lists;
        if (L.n_m2l) HIP_CHECK(hipMemcpy(out, L.m2l_order.data(), 4 * L.n_m2l, hipMemcpyDeviceToHost));
    });
}

fmmhip_status fmmhip_eval_get_expansions(fmmhip_eval* ev, int32_t which, double* out) {
    if (!ev || !out) return FMMHIP_EINVAL;
    Context* C = ev->e.ctx;
    ABI_TRY(C, {
        Tree& T = which_tree(ev->e, which);
        size_t n = (size_t)T.n_branches * 4 * n_harmonics(T.p);
        if (n) HIP_CHECK(hipMemcpy(out, T.expansions.data(), sizeof(double) * n, hipMemcpyDeviceToHost));
    });
}

fmmhip_status fmmhip_direct(fmmhip_ctx* h, int32_t kernel, int64_t nt, const double* t_xyz, int64_t ns, const double* s_buf,
                            int32_t PS, int32_t GS, int32_t HS, int32_t precision, double* out, int32_t on_device) {
    if (!h) return FMMHIP_EINVAL;
    Context* C = &h->c;
    ABI_TRY(C, {
        if (!valid_kernel(kernel)) throw Error(FMMHIP_EINVAL, "unknown kernel");
        if (nt < 0 || ns < 0 || (nt && (!t_xyz || !out)) || (ns && !s_buf)) throw Error(FMMHIP_EINVAL, "bad direct arguments");
        const int dpb = data_per_body(kernel);
        if (on_device) {
            run_direct_all(*C, kernel, nt, t_xyz, ns, s_buf, out, PS, GS, HS, precision);
        } else {
            DBuf<double> dt, ds, dout;
            dt.resize(std::max<int64_t>(1, 3 * nt));
            ds.resize(std::max<int64_t>(1, (int64_t)dpb * ns));
            dout.resize(std::max<int64_t>(1, 16 * nt));
            if (nt) HIP_CHECK(hipMemcpyAsync(dt.data(), t_xyz, sizeof(double) * 3 * nt, hipMemcpyHostToDevice, C->stream));
            if (ns) HIP_CHECK(hipMemcpyAsync(ds.data(), s_buf, sizeof(double) * dpb * ns, hipMemcpyHostToDevice, C->stream));
            if (nt) HIP_CHECK(hipMemcpyAsync(dout.data(), out, sizeof(double) * 16 * nt, hipMemcpyHostToDevice, C->stream));
            run_direct_all(*C, kernel, nt, dt.data(), ns, ds.data(), dout.data(), PS, GS, HS, precision);
            if (nt) HIP_CHECK(hipMemcpyAsync(out, dout.data(), sizeof(double) * 16 * nt, hipMemcpyDeviceToHost, C->stream));
        }
        HIP_CHECK(hipStreamSynchronize(C->stream));
    });
}


// The allocation-free form of fmmhip_direct: device pointers only, and the chunk sums go to the
// library's static per-device workspace instead of memory allocated here. No hipMalloc / hipFree.
// Calls that overlap on one device share that workspace and race.
fmmhip_status fmmhip_direct_noalloc(fmmhip_ctx* h, int32_t kernel, int64_t nt, const double* t_xyz, int64_t ns,
                                    const double* s_buf, int32_t PS, int32_t GS, int32_t HS, int32_t precision,
                                    double* out) {
    if (!h) return FMMHIP_EINVAL;
    Context* C = &h->c;
    ABI_TRY(C, {
        if (!valid_kernel(kernel)) throw Error(FMMHIP_EINVAL, "unknown kernel");
        if (nt < 0 || ns < 0 || (nt && (!t_xyz || !out)) || (ns && !s_buf)) throw Error(FMMHIP_EINVAL, "bad direct arguments");
        if (precision != FMMHIP_FP64 && precision != FMMHIP_MIXED) throw Error(FMMHIP_EINVAL, "unknown precision");
        const size_t need = direct_part_bytes(kernel, nt, ns);
        if (need > direct_scratch_bytes())
            throw Error(FMMHIP_EINVAL, "direct_noalloc needs " + std::to_string(need) + " bytes of workspace; the static "
                                       "workspace holds " + std::to_string(direct_scratch_bytes()));
        double* part = need ? direct_scratch(C->device) : nullptr;
        run_direct_all(*C, kernel, nt, t_xyz, ns, s_buf, out, PS, GS, HS, precision, part);
        HIP_CHECK(hipStreamSynchronize(C->stream));
    });
}

// --------------------------------------------------------- FastGaussSeidel
fmmhip_status fmmhip_fgs_setup(fmmhip_eval* ev, const fmmhip_options* opt, int32_t influence) {
    if (!ev) return FMMHIP_EINVAL;
    Eval& E = ev->e;
    Context* C = E.ctx;
    ABI_TRY(C, {
        check_options(opt);
        if (E.kernel != FMMHIP_POINT_SOURCE && E.kernel != FMMHIP_PANEL_SOURCE_TRI && E.kernel != FMMHIP_PANEL_SOURCE_QUAD)
            throw Error(FMMHIP_EINVAL, "FastGaussSeidel supports point sources and source panels");
        if (influence != FMMHIP_INFLUENCE_POTENTIAL && influence != FMMHIP_INFLUENCE_NORMAL_GRADIENT)
            throw Error(FMMHIP_EINVAL, "unknown influence");
        if (influence == FMMHIP_INFLUENCE_NORMAL_GRADIENT && !is_panel(E.kernel))
            throw Error(FMMHIP_EINVAL, "the normal-gradient influence needs panels");
        if (E.t_xyz || E.nt != E.ns || !E.seg.empty() || !E.sys_kind.empty())
            throw Error(FMMHIP_EINVAL, "FastGaussSeidel needs one system as target and source");
        // FastGaussSeidel's fixed choices (src/solve.jl:538-571): Barba lists, farfield and
        // nearfield without self pairs, one leaf size, potential and gradient
        fmmhip_options o = *opt;
        o.list_method = FMMHIP_BARBA;
        o.leaf_size_target = o.leaf_size_source;
        o.farfield = 1; o.nearfield = 1; o.self_induced = 0;
        o.error_kind = FMMHIP_ERR_NONE;
        o.scalar_potential = 1; o.gradient = 1; o.hessian = 0;
        o.precision = FMMHIP_FP64;
        eval_build_trees(E, o);
        E.lists.keep_discovery = true;
        try {
            eval_build_lists(E, o, false);
        } catch (...) {
            E.lists.keep_discovery = false;
            throw;
        }
        E.lists.keep_discovery = false;
        fgs_setup(E, o, influence);
    });
}

fmmhip_status fmmhip_fgs_solve(fmmhip_eval* ev, const double* external, int32_t on_device, int32_t max_iterations, double tolerance,
                               double* strengths_out, int32_t* sweeps, double* mse_history) {
    if (!ev) return FMMHIP_EINVAL;
    Eval& E = ev->e;
    Context* C = E.ctx;
    ABI_TRY(C, {
        if (!E.fgs) throw Error(FMMHIP_EINVAL, "fmmhip_fgs_setup has not run");
        if (E.ns && !external) throw Error(FMMHIP_EINVAL, "external influence is NULL");
        if (max_iterations < 0) throw Error(FMMHIP_EINVAL, "max_iterations must be >= 0");
        const double* ext = external;
        DBuf<double> tmp;
        if (!on_device) {
            tmp.resize(std::max<int64_t>(1, E.ns));
            if (E.ns) HIP_CHECK(hipMemcpyAsync(tmp.data(), external, sizeof(double) * E.ns, hipMemcpyHostToDevice, C->stream));
            ext = tmp.data();
        }
        fgs_solve(E, ext, max_iterations, tolerance, strengths_out, sweeps, mse_history);
    });
}

int64_t fmmhip_fgs_n_leaves(fmmhip_eval* ev) { return ev ? fgs_n_leaves(ev->e) : 0; }

fmmhip_status fmmhip_fgs_leaf_info(fmmhip_eval* ev, int64_t* out) {
    if (!ev) return FMMHIP_EINVAL;
    Context* C = ev->e.ctx;
    ABI_TRY(C, {
        if (!ev->e.fgs) throw Error(FMMHIP_EINVAL, "fmmhip_fgs_setup has not run");
        fgs_leaf_info(ev->e, out);
    });
}

fmmhip_status fmmhip_fgs_get_matrix(fmmhip_eval* ev, int32_t which, int64_t leaf, double* out) {
    if (!ev) return FMMHIP_EINVAL;
    Context* C = ev->e.ctx;
    ABI_TRY(C, {
        if (!ev->e.fgs) throw Error(FMMHIP_EINVAL, "fmmhip_fgs_setup has not run");
        if (which != 0 && which != 1) throw Error(FMMHIP_EINVAL, "which must be 0 (self) or 1 (non-self)");
        fgs_matrix(ev->e, which, leaf, out);
    });
}

fmmhip_status fmmhip_fgs_nonself_rows(fmmhip_eval* ev, int64_t leaf, int64_t* out) {
    if (!ev) return FMMHIP_EINVAL;
    Context* C = ev->e.ctx;
    ABI_TRY(C, {
        if (!ev->e.fgs) throw Error(FMMHIP_EINVAL, "fmmhip_fgs_setup has not run");
        fgs_nonself_rows(ev->e, leaf, out);
    });
}

}  // extern "C"
