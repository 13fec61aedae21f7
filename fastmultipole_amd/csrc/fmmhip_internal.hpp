// Internal declarations of libfmmhip (MI355X / gfx950). Host orchestration and
// device kernels share these structures; nothing here crosses the C ABI.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>
#include <stdexcept>
#include "../../include/fmmhip.h"

namespace fmmhip {

// ------------------------------------------------------------------ errors
struct Error : std::runtime_error {
    fmmhip_status status;
    Error(fmmhip_status s, const std::string& m) : std::runtime_error(m), status(s) {}
};

#define HIP_CHECK(expr)                                                                          \
    do {                                                                                         \
        hipError_t _e = (expr);                                                                  \
        if (_e != hipSuccess)                                                                    \
            throw ::fmmhip::Error(_e == hipErrorOutOfMemory ? FMMHIP_ENOMEM : FMMHIP_EDEVICE,    \
                                  std::string(#expr) + ": " + hipGetErrorString(_e));            \
    } while (0)

#define LAUNCH_CHECK() HIP_CHECK(hipGetLastError())

// -------------------------------------------------------- device buffers
// Grow-only device allocation owned by an evaluation; contents survive calls.
template <class T>
struct DBuf {
    T* p = nullptr;
    size_t cap = 0;
    size_t n = 0;
    DBuf() = default;
    DBuf(const DBuf&) = delete;
    DBuf& operator=(const DBuf&) = delete;
    ~DBuf() { if (p) (void)hipFree(p); }
    void resize(size_t count) {
        n = count;
        if (count <= cap) return;
        if (p) HIP_CHECK(hipFree(p));
        p = nullptr;
        size_t c = count + count / 8 + 64;
        HIP_CHECK(hipMalloc(&p, c * sizeof(T)));
        cap = c;
    }
    T* data() const { return p; }
    size_t size() const { return n; }
};

// ---------------------------------------------------------------- branch
// Device record of a tree branch (Branch, src/containers.jl:150-164), 0-based.
struct Branch {
    int64_t n_bodies, first, n_branches, child_first, parent, leaf;
    double source_center[3], target_center[3];
    double source_radius, target_radius;
    double source_box[3], target_box[3];
    double max_influence;
    double pad;
};

struct Tree {
    int64_t n = 0;              // bodies
    int rows = 0;               // data rows per body (16 target / data_per_body source)
    int64_t n_branches = 0;
    int depth = 0;
    std::vector<int64_t> level_first, level_last;   // host copy of levels_index
    std::vector<int64_t> leaf_index_h;               // host copy of leaf indices (lazy)
    int64_t n_leaves = 0;
    DBuf<Branch> branches;
    DBuf<int64_t> sort_index;        // sorted -> original
    DBuf<int64_t> inverse_sort;      // original -> sorted
    DBuf<int64_t> leaf_index;        // leaf k -> branch
    DBuf<double> x, y, z;            // sorted positions (SoA)
    DBuf<double> radius;             // sorted body radius (source trees)
    DBuf<double> strength;           // sorted strengths, SoA [dim][n]
    int strength_dims = 0;
    DBuf<double> expansions;         // [branch][H][2][2] (reference layout per branch)
    int p = 0;
    uint64_t build_id = 0;           // identifies the build whose keys a context's scratch may still hold
    std::vector<int64_t> seg;        // forest of independent systems: their body offsets (n_systems + 1); else empty
    // several systems in one tree (one root): their original index offsets (n_systems + 1) and leaf sizes, for
    // the leaf test that counts each system; else empty
    std::vector<int64_t> sys_off, sys_ls;
};

struct Lists {
    int64_t n_m2l = 0, n_direct = 0, n_pairs = 0;
    DBuf<int32_t> m2l_t, m2l_s, dir_t, dir_s;        // sorted by target, DFS order within target
    DBuf<int32_t> m2l_order;                         // expansion order used per M2L
    // per M2L target box: 1 while every interaction of the box met its tolerance (written by the
    // error-controlled M2L when track_box_ok is set; batches derive each system's error_success from it)
    bool track_box_ok = false;
    DBuf<int32_t> m2l_box_ok;
    DBuf<int32_t> m2l_perm, m2l_iota;                // processing order of the M2L list within each target box
    DBuf<float> m2l_key, m2l_key_sorted;
    // CSR over target branches for the direct list: unique targets and offsets
    int64_t n_direct_targets = 0;
    DBuf<int32_t> dir_targets;                       // unique target branches in the direct list
    DBuf<int64_t> dir_offsets;                       // n_direct_targets + 1
    DBuf<int64_t> dir_cum;                           // n_direct + 1: exclusive prefix of source leaf sizes
    // direct targets can nest (SelfTuning pairs branches that are not leaves): the positions in
    // dir_targets where each level of the target tree starts, plus n_direct_targets; else empty
    std::vector<int64_t> dir_level_split;
    int64_t n_m2l_targets = 0;
    DBuf<int32_t> m2l_targets;
    DBuf<int64_t> m2l_offsets;
    bool keep_discovery = false;                     // also keep the direct list in discovery order
    DBuf<uint32_t> dfs_t;                            // (build_interaction_lists' order, before sort_by_target)
    DBuf<int32_t> dfs_s;
};

struct Tables;        // device constant tables (rotation normalisations), expansions.hip
struct TreeScratch;   // tree build work buffers, tree.hip
struct ListScratch;   // interaction-list work buffers, lists.hip

// One stream and everything a build or an evaluation on it needs besides the
// evaluation's own state: contexts on different threads share nothing.
struct Context {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string last_error;
    Tables* tables = nullptr;
    TreeScratch* tree_scratch = nullptr;
    ListScratch* list_scratch = nullptr;
    DBuf<int32_t> m2l_flags;    // pmax and success of the last M2L pass
    DBuf<double> l2p_coef;      // per-leaf local-to-body coefficients of the last L2P pass
    DBuf<double> m2m_slot;      // per-child translated multipoles of one level of the upward pass
    DBuf<unsigned long long> p2p_queue;   // work counter of the persistent nearfield launch
    DBuf<double> direct_part;   // chunk sums of the split direct! form
    int n_cu = 0;               // compute units of the device (0 until first needed)
    DBuf<unsigned char> temp;   // scratch for sorts/scans
    DBuf<unsigned char> sort_work;   // radix sort: digit counts, offsets, keys and values between passes
    void* pinned = nullptr;     // 64 bytes of pinned host memory for scalar read-backs
    // Launch sequences without host synchronisation may be captured in graphs and replayed. Off for
    // contexts that share the device with other threads (the request server's workers): a
    // device-synchronising call there (hipFree) while a stream is being captured invalidates the capture.
    bool graphs = true;
    int p2p_free_slots = 1;     // workgroup slots per CU the nearfield leaves to the far-field stream
};

// Device scalars read back through the context's pinned buffer, one stream synchronisation per call
inline void* pinned_scratch(Context& C) {
    if (!C.pinned) HIP_CHECK(hipHostMalloc(&C.pinned, 64, hipHostMallocDefault));
    return C.pinned;
}
template <class T>
inline T read_back(Context& C, const T* d) {
    T* h = (T*)pinned_scratch(C);
    HIP_CHECK(hipMemcpyAsync(h, d, sizeof(T), hipMemcpyDeviceToHost, C.stream));
    HIP_CHECK(hipStreamSynchronize(C.stream));
    return *h;
}
// a + b of two device scalars (an exclusive scan's last entry plus its last flag)
template <class T>
inline int64_t read_back_sum(Context& C, const T* a, const T* b) {
    T* h = (T*)pinned_scratch(C);
    HIP_CHECK(hipMemcpyAsync(h, a, sizeof(T), hipMemcpyDeviceToHost, C.stream));
    HIP_CHECK(hipMemcpyAsync(h + 1, b, sizeof(T), hipMemcpyDeviceToHost, C.stream));
    HIP_CHECK(hipStreamSynchronize(C.stream));
    return (int64_t)h[0] + (int64_t)h[1];
}
void free_tree_scratch(TreeScratch* s);
// the request server's worker contexts capture no graphs (api.hip)
void ctx_disable_graphs(fmmhip_ctx* ctx);
void free_list_scratch(ListScratch* s);

struct Eval {
    Context* ctx = nullptr;
    int kernel = 0;
    int64_t nt = 0, ns = 0;
    DBuf<double> t_xyz_in, s_buf_in;          // host-copied inputs (when on_device == 0)
    const double* t_xyz = nullptr;            // device pointer to 3 x nt
    const double* s_buf = nullptr;            // device pointer to dpb x ns
    std::vector<int64_t> seg;                 // several independent systems (each its own target and source):
                                              // their body offsets; empty for one target and one source system
    DBuf<int64_t> seg_dev;                    // seg on the device
    // several systems in one tree, each a target and a source of all (fmm!((s1, s2, ...))): their kernels
    // and body offsets (empty: one system of `kernel`), and the rows and strength rows of their shared
    // buffer (the largest of the systems')
    std::vector<int32_t> sys_kind;
    std::vector<int64_t> sys_off;
    int sys_rows = 0, sys_sdims = 0;
    std::vector<int64_t> sys_leaf;            // leaf size per system (empty: the options' leaf size for each)
    DBuf<int64_t> sys_off_dev;
    DBuf<Branch> sys_branches;                // per system: the source tree's branches with its leaf ranges
    DBuf<int64_t> sys_dir_cum;                // per system: prefix of its counts over the direct list
    DBuf<int32_t> sys_status;                 // per system of a batch: largest order used, tolerance met
    std::vector<int32_t> sys_pmax, sys_ok;    // host copies after the evaluation (pmax + 1 where not met)
    Tree tt, st;
    Lists lists;
    DBuf<double> tout;                        // sorted target outputs, SoA [13][nt]: phi, v(3), h(9)
    DBuf<double> tout_aos;                    // 16 x nt original order
    bool trees_built = false, lists_built = false, ran = false;
    // the context's tree scratch holds the sorted keys of the positions st was built from, and no
    // evaluation has run since: estimate_influence may reuse them (positions handed over in place
    // can change between an evaluation and the next one)
    bool keys_fresh = false;
    int32_t pmax = 0, error_success = 1;
    int p_built = -1;
    Eval* estimate = nullptr;                 // persistent low-order evaluation of estimate_influence
    Context* side = nullptr;                  // second stream: the far field runs beside the nearfield
    DBuf<double> influence;                   // per-target max |grad| (or |phi|) of the estimate
    struct FgsState* fgs = nullptr;           // FastGaussSeidel solver state (fgs.hip)
    ~Eval();
};

__host__ __device__ constexpr bool panel_source_dipole(int kernel) {
    return kernel == FMMHIP_PANEL_SOURCE_DIPOLE_TRI || kernel == FMMHIP_PANEL_SOURCE_DIPOLE_QUAD;
}
__host__ __device__ constexpr bool is_panel(int kernel) {
    return (kernel >= FMMHIP_PANEL_SOURCE_TRI && kernel <= FMMHIP_PANEL_DIPOLE_QUAD) || panel_source_dipole(kernel);
}
__host__ __device__ constexpr int panel_vertices(int kernel) {
    return kernel == FMMHIP_PANEL_SOURCE_QUAD || kernel == FMMHIP_PANEL_DIPOLE_QUAD || kernel == FMMHIP_PANEL_SOURCE_DIPOLE_QUAD ? 4 : 3;
}
__host__ __device__ constexpr bool panel_dipole(int kernel) {
    return kernel == FMMHIP_PANEL_DIPOLE_TRI || kernel == FMMHIP_PANEL_DIPOLE_QUAD;
}
// strength rows of a panel ahead of its vertices
__host__ __device__ constexpr int panel_strengths(int kernel) { return panel_source_dipole(kernel) ? 2 : 1; }
__host__ __device__ inline int data_per_body(int kernel) {
    return kernel == FMMHIP_POINT_SOURCE ? 5
         : is_panel(kernel)                    ? 4 + panel_strengths(kernel) + 3 * panel_vertices(kernel)
         : kernel == FMMHIP_POINT_SOURCE_VORTEX ? 8
                                                : 7;
}
// rows after [x, y, z, radius] kept with every sorted body: strengths (and panel vertices)
inline int strength_dims(int kernel) { return kernel == FMMHIP_POINT_SOURCE ? 1 : data_per_body(kernel) - 4; }
inline bool valid_kernel(int kernel) { return kernel >= FMMHIP_POINT_SOURCE && kernel <= FMMHIP_PANEL_SOURCE_DIPOLE_QUAD; }
// has_vector_potential (test/vortex.jl:145-147): Lamb-Helmholtz chi expansions
inline bool has_vector_potential(int kernel) {
    return kernel == FMMHIP_POINT_VORTEX || kernel == FMMHIP_VORTEX_GAUSSIAN || kernel == FMMHIP_POINT_SOURCE_VORTEX;
}
inline int n_harmonics(int p) { return ((p + 1) * (p + 2)) / 2; }

// ---------------------------------------------------------- module entry
// tree.hip
// keys: the scratch of another context whose last build sorted the same body buffer (its Morton
// keys and order are reused instead of computed again), or null
void build_tree(Context& C, Tree& T, int64_t n, const double* d_buf, int rows, bool target, int64_t leaf_size,
                int method, bool shrink, int p, int strength_dims, const TreeScratch* keys = nullptr);
// the scratch of context C if it still holds the sorted keys of T's current build, else null
const TreeScratch* tree_keys(const Context& C, const Tree& T);
// the target tree of the bodies of source tree s (same leaf size and options), by device copies
void clone_target_tree(Context& C, const Tree& s, Tree& t);
// lists.hip
void build_lists(Context& C, const Tree& tt, const Tree& st, Lists& L, int64_t leaf_size_source, double mac,
                 int method, bool ff, bool nf, bool si, bool count_pairs);
// p2p.hip
// nearfield of the target leaves b0 .. b1 - 1 of the direct list (all by default). The point
// kernels run as workgroups taking leaves from the context's queue: at most `slots` per CU
// (0: as many as fit, < 0: that many fewer than fit).
// One system of several sharing the source tree: its bodies' range in each leaf (a copy of the tree's
// branches with the leaves' first / n_bodies replaced) and the direct list's prefix of its counts.
struct SourceView {
    const Branch* branches = nullptr;
    const int64_t* dir_cum = nullptr;
};
void run_p2p(Context& C, int kernel, const Tree& tt, const Tree& st, const Lists& L, double* tout, int64_t nt,
             bool PS, bool GS, bool HS, int precision,
             int64_t b0 = 0, int64_t b1 = -1, int slots = 0, SourceView sv = SourceView());
// part: workspace of direct_part_bytes() for the chunk sums of the split form (null: the context's own)
void run_direct_all(Context& C, int kernel, int64_t nt, const double* t_xyz, int64_t ns, const double* s_buf,
                    double* tout_aos, bool PS, bool GS, bool HS, int precision, double* part = nullptr);
// source chunks of direct! over nt targets and ns sources, and the bytes of their partial sums (0: no workspace)
int direct_split(int64_t nt, int64_t ns);
size_t direct_part_bytes(int kernel, int64_t nt, int64_t ns);
// device address of the library's static direct! workspace on the current device (looked up once per
// device) and its size in bytes
double* direct_scratch(int device);
size_t direct_scratch_bytes();
// expansions.hip
void ensure_tables(Context& C, int p);
void free_tables(Tables* t);
// sbr: the source tree's branches with one system's leaf ranges (several systems in one tree), or null
void run_p2m(Context& C, int kernel, Tree& st, int p, const Branch* sbr = nullptr);
void run_m2m(Context& C, Tree& st, int p, bool LH);
// pmax / success null: no read-back of the launch's largest order and tolerance flag (no host synchronisation)
void run_m2l(Context& C, const Tree& tt, Tree& st_src, Tree& tt_dst, Lists& L, int p, bool LH, int err_kind,
             bool err_rel, double eps, bool BE, int precision, int32_t* pmax, int32_t* success);
void run_l2l(Context& C, Tree& tt, int p, bool LH);
// zeroes n doubles with a kernel (graph-capturable, unlike hipMemsetAsync)
void fill_zero(Context& C, double* p, size_t n);
void run_l2p(Context& C, Tree& tt, double* tout, int64_t nt, int p, bool LH, bool PS, bool GS, bool HS);
// api.hip helpers
void gather_outputs(Context& C, const Tree& tt, const double* tout, double* aos, int64_t nt);
// fgs.hip: FastGaussSeidel (src/solve.jl)
void fgs_setup(Eval& E, const fmmhip_options& o, int influence);
void fgs_solve(Eval& E, const double* ext_dev, int max_iterations, double tolerance, double* strengths_host, int32_t* sweeps,
               double* mse_host);
int64_t fgs_n_leaves(const Eval& E);
void fgs_leaf_info(const Eval& E, int64_t* out_host);
void fgs_matrix(const Eval& E, int which, int64_t leaf, double* out_host);
void fgs_nonself_rows(const Eval& E, int64_t leaf, int64_t* out_host);
void free_fgs(struct FgsState* f);

}  // namespace fmmhip
