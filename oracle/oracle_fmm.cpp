// Oracle (test infrastructure only): CPU FP64 restatement of FastMultipole's
// octree (src/tree.jl), interaction lists (src/interaction_list.jl), pass
// drivers (src/fmm.jl) and the user kernels of test/gravitational.jl and
// test/vortex.jl. Only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg may load it, and only as the checker.
//
// Index convention: branches, bodies and leaves are 0-based here; the
// reference's 1-based values are these plus one (root parent 0 <-> -1).
#include "oracle_math.hpp"
#include <omp.h>
#include <cstring>
#include <memory>
#include <array>

namespace orc {

struct OBranch {
    int64_t n_bodies, first;           // bodies_index = first : first+n_bodies-1
    int64_t n_branches, child_first;   // branch_index = child_first : child_first+n_branches-1
    int64_t i_parent, i_leaf;          // -1 when absent
    double source_center[3], target_center[3];
    double source_radius, target_radius;
    double source_box[3], target_box[3];
    double max_influence;
};

enum { METHOD_BARBA = 0, METHOD_SELFTUNING = 1, METHOD_SELFTUNING_TREESTOP = 2 };
// kernel kinds: point source, singular point vortex (the reference's test kernels),
// point dipole and Gaussian-core vortex (kernels of this library with no
// reference direct! counterpart; their formulas are stated at the functions)
enum { KIND_SOURCE = 0, KIND_VORTEX = 1, KIND_DIPOLE = 2, KIND_VORTEX_GAUSS = 3,
       // constant-strength planar panels (test/panels.jl), triangles and quadrilaterals;
       // buffer rows [centroid(3), radius, strength, v1(3), v2(3), v3(3), v4(3) for quads]
       KIND_PANEL_SOURCE_TRI = 4, KIND_PANEL_DIPOLE_TRI = 5, KIND_PANEL_SOURCE_QUAD = 6, KIND_PANEL_DIPOLE_QUAD = 7,
       // Point{SourceVortex} (src/bodytomultipole.jl:491-496): rows [x, y, z, sigma, q, Gx, Gy, Gz]
       KIND_SOURCE_VORTEX = 8,
       // Panel{SourceDipole} (test/panels.jl:323-429; src/bodytomultipole.jl:717-720): rows [centroid(3), radius,
       // source strength, dipole strength, v1(3), v2(3), v3(3), v4(3) for quads]
       KIND_PANEL_SOURCE_DIPOLE_TRI = 9, KIND_PANEL_SOURCE_DIPOLE_QUAD = 10 };
static bool panel_source_dipole(int kind) { return kind == KIND_PANEL_SOURCE_DIPOLE_TRI || kind == KIND_PANEL_SOURCE_DIPOLE_QUAD; }
static bool is_panel(int kind) {
    return (kind >= KIND_PANEL_SOURCE_TRI && kind <= KIND_PANEL_DIPOLE_QUAD) || panel_source_dipole(kind);
}
static int panel_vertices(int kind) {
    return kind == KIND_PANEL_SOURCE_QUAD || kind == KIND_PANEL_DIPOLE_QUAD || kind == KIND_PANEL_SOURCE_DIPOLE_QUAD ? 4 : 3;
}
static bool panel_dipole(int kind) { return kind == KIND_PANEL_DIPOLE_TRI || kind == KIND_PANEL_DIPOLE_QUAD; }
static int panel_strengths(int kind) { return panel_source_dipole(kind) ? 2 : 1; }   // rows ahead of the vertices

struct OTree {
    std::vector<OBranch> branches;
    std::vector<std::array<int64_t, 2>> levels;   // [first, last] 0-based inclusive
    std::vector<int64_t> leaf_index, sort_index, inverse_sort_index;
    std::vector<double> buffer;   // rows x N, column-major, sorted order
    int rows = 0;
    std::vector<double> expansions;
    int p = 0;
};

// -------------------------------------------------------------- octree
static bool exceeds(int64_t n, int64_t ls, int method) {   // src/tree.jl:492-569 (single system)
    if (method == METHOD_BARBA) return n > ls;
    double fraction = 0.0;
    fraction += (double)n / (double)(ls * ls);
    fraction *= (double)ls;
    if (method == METHOD_SELFTUNING) return fraction > 0.5 && n > 1;
    return fraction > 2.8284271247461903 && n > 1;
}

// exceeds for several systems in one tree (src/tree.jl:492-569): the count of each system against its
// own leaf size; count(i) gives system i's bodies in the branch
template <class Count>
static bool exceeds_systems(int n_sys, Count count, const int64_t* ls, int method) {
    if (method == METHOD_BARBA) {
        bool not_over = true;
        for (int i = 0; i < n_sys; ++i) not_over = not_over && count(i) <= ls[i];
        return !not_over;
    }
    double fraction = 0.0;
    int64_t n_bodies = 0;
    int64_t min_ls = ls[0];
    for (int i = 0; i < n_sys; ++i) {
        const int64_t n = count(i);
        n_bodies += n;
        fraction += (double)n / (double)(ls[i] * ls[i]);
        min_ls = std::min(min_ls, ls[i]);
    }
    fraction *= (double)min_ls;
    if (method == METHOD_SELFTUNING) return fraction > 0.5 && n_bodies > 1;
    return fraction > 2.8284271247461903 && n_bodies > 1;
}

static inline int get_octant(const double* pos, const double* c) {   // src/tree.jl:474-480
    int o = 1;
    if (pos[0] > c[0]) o += 1;
    if (pos[1] > c[1]) o += 2;
    if (pos[2] > c[2]) o += 4;
    return o;
}

static void get_child_center(const double* pc, const double* pbox, int oct, double* cc) {   // src/tree.jl:581-592
    double delta = pbox[0] * 0.5;
    oct -= 1;
    double dx = (oct % 2 == 0) ? -delta : delta;
    delta = pbox[1] * 0.5;
    oct >>= 1;
    double dy = (oct % 2 == 0) ? -delta : delta;
    delta = pbox[2] * 0.5;
    oct >>= 1;
    double dz = (oct % 2 == 0) ? -delta : delta;
    cc[0] = pc[0] + dx; cc[1] = pc[1] + dy; cc[2] = pc[2] + dz;
}

struct TreeBuilder {
    OTree& T;
    int64_t N;
    int64_t ls;
    int method;
    std::vector<double> small;        // 3 x N
    std::vector<int64_t> sort_buf;
    TreeBuilder(OTree& t, int64_t n, int64_t l, int m) : T(t), N(n), ls(l), method(m), small(3 * n), sort_buf(n) {}
    double* pos(int64_t i) { return &T.buffer[(size_t)i * T.rows]; }

    // several systems: original index ranges of the systems and their leaf sizes (null: one system)
    const std::vector<int64_t>* sys_off = nullptr;
    const std::vector<int64_t>* sys_ls = nullptr;

    // the leaf test of the branch holding bodies first .. first + count - 1
    bool splits(int64_t first, int64_t count) {
        if (!sys_off) return exceeds(count, ls, method);
        const int n_sys = (int)sys_off->size() - 1;
        std::vector<int64_t> c(n_sys, 0);
        for (int64_t i = first; i < first + count; ++i) {
            const int64_t o = T.sort_index[i];
            c[std::upper_bound(sys_off->begin(), sys_off->end(), o) - sys_off->begin() - 1] += 1;
        }
        return exceeds_systems(n_sys, [&](int i) { return c[i]; }, sys_ls->data(), method);
    }

    void census(int64_t first, int64_t count, const double* c, int64_t* cum) {   // :608-625
        for (int k = 0; k < 8; ++k) cum[k] = 0;
        for (int64_t i = first; i < first + count; ++i) cum[get_octant(pos(i), c) - 1] += 1;
        for (int k = 1; k < 8; ++k) cum[k] += cum[k - 1];
    }

    // branch! (src/tree.jl:433-463)
    OBranch branch(int64_t i_first_branch, int64_t first, int64_t count, const double* sc, const double* tc,
                   double sr, double tr, const double* sbox, const double* tbox, int64_t i_parent, int64_t& i_leaf,
                   int64_t& n_children) {
        int64_t cum[8];
        census(first, count, tc, cum);
        n_children = 0;
        if (splits(first, count))
            for (int o = 0; o < 8; ++o) if ((o == 0 ? cum[0] : cum[o] - cum[o - 1]) > 0) n_children += 1;
        if (n_children > 0) {
            int64_t oi[8];   // octant_beginning_index! (0-based)
            for (int o = 7; o >= 1; --o) oi[o] = cum[o - 1] + first;
            oi[0] = first;
            for (int64_t i = first; i < first + count; ++i) {   // sort_bodies! :712-737
                int o = get_octant(pos(i), tc) - 1;
                int64_t ti = oi[o];
                for (int d = 0; d < 3; ++d) small[3 * ti + d] = pos(i)[d];
                sort_buf[ti] = T.sort_index[i];
                oi[o] += 1;
            }
            for (int64_t i = first; i < first + count; ++i) {
                for (int d = 0; d < 3; ++d) pos(i)[d] = small[3 * i + d];
                T.sort_index[i] = sort_buf[i];
            }
        }
        OBranch b;
        b.n_bodies = count;
        b.first = first;
        b.n_branches = n_children;
        b.child_first = i_first_branch;
        b.i_parent = i_parent;
        if (n_children == 0) { b.i_leaf = i_leaf; i_leaf += 1; } else b.i_leaf = -1;
        for (int d = 0; d < 3; ++d) {
            b.source_center[d] = sc[d]; b.target_center[d] = tc[d];
            b.source_box[d] = sbox[d]; b.target_box[d] = tbox[d];
        }
        b.source_radius = sr;
        b.target_radius = tr;
        b.max_influence = 0.0;
        return b;
    }

    // child_branches! (src/tree.jl:346-376)
    void child_branches(int64_t& pfirst, int64_t& plast, int64_t& n_children, int64_t& i_leaf) {
        int64_t i_first_branch = plast + n_children + 1;
        for (int64_t ip = pfirst; ip <= plast; ++ip) {
            OBranch parent = T.branches[ip];
            if (parent.n_branches > 0) {
                double child_radius = parent.target_radius * 0.5;
                double child_box[3] = {parent.target_box[0] * 0.5, parent.target_box[1] * 0.5, parent.target_box[2] * 0.5};
                int64_t cum[8];
                census(parent.first, parent.n_bodies, parent.target_center, cum);
                if (splits(parent.first, parent.n_bodies)) {
                    for (int o = 1; o <= 8; ++o) {
                        int64_t pop = (o == 1) ? cum[0] : cum[o - 1] - cum[o - 2];
                        if (pop > 0) {
                            int64_t off = (o == 1) ? 0 : cum[o - 2];
                            double cc[3];
                            get_child_center(parent.target_center, parent.target_box, o, cc);
                            int64_t ngc;
                            OBranch child = branch(i_first_branch, parent.first + off, pop, cc, cc, child_radius, child_radius,
                                                   child_box, child_box, ip, i_leaf, ngc);
                            i_first_branch += ngc;
                            T.branches.push_back(child);
                        }
                    }
                }
            }
        }
        n_children = i_first_branch - (int64_t)T.branches.size();
        pfirst = plast + 1;
        plast = (int64_t)T.branches.size() - 1;
    }
};

// center_box / source_center_box / shrink_radius (src/tree.jl:845-1139)
static void shrink_recenter(OTree& T, bool use_radius) {   // src/tree.jl:1176-1188
    const int rows = T.rows;
    auto P = [&](int64_t i) { return &T.buffer[(size_t)i * rows]; };
    auto R = [&](int64_t i) { return use_radius ? T.buffer[(size_t)i * rows + 3] : 0.0; };
    for (int64_t lv = (int64_t)T.levels.size() - 1; lv >= 0; --lv) {
        for (int64_t ib = T.levels[lv][0]; ib <= T.levels[lv][1]; ++ib) {
            OBranch& b = T.branches[ib];
            double tc[3], tb[3], sc[3], sbx[3], tr = 0.0, sr = 0.0;
            if (b.n_branches == 0) {   // shrink_leaf!
                double mn[3], mx[3], smn[3], smx[3];
                for (int d = 0; d < 3; ++d) { mn[d] = mx[d] = smn[d] = smx[d] = P(b.first)[d]; }
                for (int64_t i = b.first; i < b.first + b.n_bodies; ++i) {
                    double r = R(i);
                    for (int d = 0; d < 3; ++d) {
                        double x = P(i)[d];
                        mn[d] = std::min(mn[d], x); mx[d] = std::max(mx[d], x);
                        smn[d] = std::min(smn[d], x - r); smx[d] = std::max(smx[d], x + r);
                    }
                }
                for (int d = 0; d < 3; ++d) {
                    tc[d] = (mx[d] + mn[d]) * 0.5; tb[d] = mx[d] - tc[d];
                    sc[d] = (smx[d] + smn[d]) * 0.5; sbx[d] = smx[d] - sc[d];
                }
                for (int64_t i = b.first; i < b.first + b.n_bodies; ++i) {
                    double dx = P(i)[0] - tc[0], dy = P(i)[1] - tc[1], dz = P(i)[2] - tc[2];
                    double dist = std::sqrt(dx * dx + dy * dy + dz * dz);
                    tr = std::max(tr, dist);
                    dx = P(i)[0] - sc[0]; dy = P(i)[1] - sc[1]; dz = P(i)[2] - sc[2];
                    dist = std::sqrt(dx * dx + dy * dy + dz * dz);
                    sr = std::max(sr, dist + R(i));
                }
            } else {   // shrink_branch!
                const OBranch& f = T.branches[b.child_first];
                double mn[3], mx[3], smn[3], smx[3];
                for (int d = 0; d < 3; ++d) { mn[d] = mx[d] = f.target_center[d]; smn[d] = smx[d] = f.source_center[d]; }
                for (int64_t c = b.child_first; c < b.child_first + b.n_branches; ++c) {
                    const OBranch& ch = T.branches[c];
                    for (int d = 0; d < 3; ++d) {
                        mn[d] = std::min(mn[d], ch.target_center[d] - ch.target_box[d]);
                        mx[d] = std::max(mx[d], ch.target_center[d] + ch.target_box[d]);
                        smn[d] = std::min(smn[d], ch.source_center[d] - ch.source_box[d]);
                        smx[d] = std::max(smx[d], ch.source_center[d] + ch.source_box[d]);
                    }
                }
                for (int d = 0; d < 3; ++d) {
                    tc[d] = (mn[d] + mx[d]) * 0.5; tb[d] = mx[d] - tc[d];
                    sc[d] = (smn[d] + smx[d]) * 0.5; sbx[d] = smx[d] - sc[d];
                }
                for (int64_t c = b.child_first; c < b.child_first + b.n_branches; ++c) {
                    const OBranch& ch = T.branches[c];
                    double dx = ch.target_center[0] - tc[0], dy = ch.target_center[1] - tc[1], dz = ch.target_center[2] - tc[2];
                    double dist = std::sqrt(dx * dx + dy * dy + dz * dz);
                    tr = std::max(tr, dist + ch.target_radius);
                    dx = ch.source_center[0] - sc[0]; dy = ch.source_center[1] - sc[1]; dz = ch.source_center[2] - sc[2];
                    dist = std::sqrt(dx * dx + dy * dy + dz * dz);
                    sr = std::max(sr, dist + ch.source_radius);
                }
            }
            for (int d = 0; d < 3; ++d) {
                b.target_center[d] = tc[d]; b.target_box[d] = tb[d];
                b.source_center[d] = sc[d]; b.source_box[d] = sbx[d];
            }
            b.target_radius = tr;
            b.source_radius = sr;
        }
    }
}

// Tree(systems, target; ...) for one system (src/tree.jl:3-122).
// `full` is the system in its original order: rows x N (target: 16 rows with
// positions only; source: data_per_body rows).
static void build_tree(OTree& T, int64_t N, int rows, const double* full, bool target, int64_t ls, int method,
                       bool shrink, int p, const std::vector<int64_t>* sys_off = nullptr,
                       const std::vector<int64_t>* sys_ls = nullptr, int n_divisions = 20) {
    T = OTree();
    T.rows = rows;
    T.p = p;
    if (N == 0) return;
    T.buffer.assign((size_t)rows * N, 0.0);
    for (int64_t i = 0; i < N; ++i)
        for (int d = 0; d < 3; ++d) T.buffer[(size_t)i * rows + d] = full[(size_t)i * rows + d];   // target_to_buffer!
    T.sort_index.resize(N);
    for (int64_t i = 0; i < N; ++i) T.sort_index[i] = i;
    // center_box (:894-907)
    double mn[3], mx[3];
    for (int d = 0; d < 3; ++d) mn[d] = mx[d] = T.buffer[d];
    for (int64_t i = 0; i < N; ++i)
        for (int d = 0; d < 3; ++d) {
            double x = T.buffer[(size_t)i * rows + d];
            mn[d] = std::min(mn[d], x);
            mx[d] = std::max(mx[d], x);
        }
    double center[3], box[3];
    for (int d = 0; d < 3; ++d) { center[d] = (mx[d] + mn[d]) * 0.5; box[d] = mx[d] - center[d]; }
    double bmax = std::max(std::max(box[0], box[1]), box[2]);
    double radius = std::sqrt(bmax * bmax * 3.0);
    double cbox[3] = {bmax, bmax, bmax};
    TreeBuilder B(T, N, ls, method);
    if (sys_off && sys_off->size() > 2) { B.sys_off = sys_off; B.sys_ls = sys_ls; }
    int64_t i_leaf = 0, n_children = 0;
    OBranch root = B.branch(1, 0, N, center, center, radius, radius, cbox, cbox, -1, i_leaf, n_children);
    T.branches.push_back(root);
    int64_t pf = 0, pl = 0;
    T.levels.push_back({0, 0});
    for (int k = 0; k < n_divisions; ++k)
        if (n_children > 0) { B.child_branches(pf, pl, n_children, i_leaf); T.levels.push_back({pf, pl}); }
    while (n_children > 0) { B.child_branches(pf, pl, n_children, i_leaf); T.levels.push_back({pf, pl}); }
    if (!target)   // system_to_buffer!(buffers, systems, sort_index)
        for (int64_t i = 0; i < N; ++i)
            for (int r = 0; r < rows; ++r) T.buffer[(size_t)i * rows + r] = full[(size_t)T.sort_index[i] * rows + r];
    T.inverse_sort_index.resize(N);
    for (int64_t i = 0; i < N; ++i) T.inverse_sort_index[T.sort_index[i]] = i;
    if (shrink) shrink_recenter(T, !target);
    for (size_t ib = 0; ib < T.branches.size(); ++ib)
        if (T.branches[ib].n_branches == 0) T.leaf_index.push_back((int64_t)ib);
    T.expansions.assign((size_t)4 * n_harmonics(p) * T.branches.size(), 0.0);
}

// ----------------------------------------------------- interaction lists
struct ListBuilder {   // src/interaction_list.jl:3-166
    const OTree& Tt;
    const OTree& Ts;
    int64_t ls;
    double mac;
    bool ff, nf, si;
    int method;
    std::vector<std::array<int32_t, 2>> m2l, direct;
    void run(int32_t it, int32_t js) {
        const OBranch& s = Ts.branches[js];
        const OBranch& t = Tt.branches[it];
        if (method == METHOD_SELFTUNING_TREESTOP || method == METHOD_BARBA) {
            bool both_leaves = (s.n_branches == 0 && t.n_branches == 0);
            if (method == METHOD_SELFTUNING_TREESTOP && both_leaves) {
                if ((nf || (it == js && si)) && (it != js || si)) direct.push_back({it, js});
                return;
            }
            double dx = t.target_center[0] - s.source_center[0];
            double dy = t.target_center[1] - s.source_center[1];
            double dz = t.target_center[2] - s.source_center[2];
            double sep2 = dx * dx + dy * dy + dz * dz;
            double sr = s.source_radius + t.target_radius;
            if (sep2 * mac * mac > sr * sr) {
                if (ff) m2l.push_back({it, js});
                return;
            }
            if (method == METHOD_BARBA && both_leaves) {
                if ((nf || (it == js && si)) && (it != js || si)) direct.push_back({it, js});
                return;
            }
            bool split_target;
            if (method == METHOD_BARBA)
                split_target = s.n_branches == 0 || (t.target_radius >= s.source_radius && t.n_branches != 0);
            else
                split_target = s.n_branches == 0 || (t.n_bodies >= s.n_bodies && t.n_branches != 0);
            if (split_target)
                for (int64_t c = t.child_first; c < t.child_first + t.n_branches; ++c) run((int32_t)c, js);
            else
                for (int64_t c = s.child_first; c < s.child_first + s.n_branches; ++c) run(it, (int32_t)c);
            return;
        }
        // SelfTuning (:66-120)
        double fraction = 0.0;
        fraction += (double)s.n_bodies / (double)(ls * ls);
        fraction *= (double)t.n_bodies;
        if (fraction < 1.0 || (s.n_branches == 0 && t.n_branches == 0)) {
            if (nf && (it != js || si)) direct.push_back({it, js});
            return;
        }
        double dx = t.target_center[0] - s.source_center[0];
        double dy = t.target_center[1] - s.source_center[1];
        double dz = t.target_center[2] - s.source_center[2];
        double sep2 = dx * dx + dy * dy + dz * dz;
        double sr = s.source_radius + t.target_radius;
        if (sep2 * mac * mac > sr * sr) {
            if (ff) m2l.push_back({it, js});
            return;
        }
        if (s.n_branches == 0 || (t.n_bodies >= s.n_bodies && t.n_branches != 0))
            for (int64_t c = t.child_first; c < t.child_first + t.n_branches; ++c) run((int32_t)c, js);
        else
            for (int64_t c = s.child_first; c < s.child_first + s.n_branches; ++c) run(it, (int32_t)c);
    }
};

static std::vector<std::array<int32_t, 2>> sort_by_target(const std::vector<std::array<int32_t, 2>>& list, size_t nt) {  // :183-213
    std::vector<int64_t> cnt(nt + 1, 0);
    for (auto& e : list) cnt[e[0] + 1] += 1;
    for (size_t i = 1; i <= nt; ++i) cnt[i] += cnt[i - 1];
    std::vector<std::array<int32_t, 2>> out(list.size());
    for (auto& e : list) out[cnt[e[0]]++] = e;
    return out;
}

// ----------------------------------------------------------- user kernels
// test/gravitational.jl:80-104 (ignores the derivative switches, as the reference kernel does)
static void direct_gravitational(double* tbuf, int64_t t0, int64_t nt, const double* sbuf, int sr, int64_t s0, int64_t ns) {
    for (int64_t js = s0; js < s0 + ns; ++js) {
        const double* s = sbuf + (size_t)js * sr;
        double sx = s[0], sy = s[1], sz = s[2], q = s[4];
        for (int64_t it = t0; it < t0 + nt; ++it) {
            double* t = tbuf + (size_t)it * 16;
            double dx = t[0] - sx, dy = t[1] - sy, dz = t[2] - sz;
            double r2 = dx * dx + dy * dy + dz * dz;
            if (r2 > 0) {
                double r = std::sqrt(r2);
                double dphi = q / r * ONE_OVER_4PI;
                t[3] += dphi;
                double r3 = r2 * r;
                t[4] += dx * q / r3 * ONE_OVER_4PI;
                t[5] += dy * q / r3 * ONE_OVER_4PI;
                t[6] += dz * q / r3 * ONE_OVER_4PI;
            }
        }
    }
}

// test/vortex.jl:83-128
static void direct_vortex(double* tbuf, int64_t t0, int64_t nt, const double* sbuf, int sr, int64_t s0, int64_t ns, bool V, bool VG) {
    for (int64_t js = s0; js < s0 + ns; ++js) {
        const double* s = sbuf + (size_t)js * sr;
        double sx = s[0], sy = s[1], sz = s[2], Gx = s[4], Gy = s[5], Gz = s[6];
        for (int64_t it = t0; it < t0 + nt; ++it) {
            double* t = tbuf + (size_t)it * 16;
            double dx = t[0] - sx, dy = t[1] - sy, dz = t[2] - sz;
            double r2 = dx * dx + dy * dy + dz * dz;
            if (r2 > 0) {
                double r = std::sqrt(r2);
                double rinv = 1 / r;
                double rinv2 = rinv * rinv;
                double denom = ONE_OVER_4PI * rinv * rinv2;
                if (V) {
                    t[4] += (dz * Gy - dy * Gz) * denom;
                    t[5] += (dx * Gz - dz * Gx) * denom;
                    t[6] += (dy * Gx - dx * Gy) * denom;
                }
                if (VG) {
                    denom *= rinv2;
                    double h[9];
                    h[0] = -3 * dx * (Gy * dz - Gz * dy) * denom;
                    h[1] = (-3 * dx * (Gz * dx - Gx * dz) + Gz * r2) * denom;
                    h[2] = (-3 * dx * (Gx * dy - Gy * dx) - Gy * r2) * denom;
                    h[3] = (-3 * dy * (Gy * dz - Gz * dy) - Gz * r2) * denom;
                    h[4] = -3 * dy * (Gz * dx - Gx * dz) * denom;
                    h[5] = (-3 * dy * (Gx * dy - Gy * dx) + Gx * r2) * denom;
                    h[6] = (-3 * dz * (Gy * dz - Gz * dy) + Gy * r2) * denom;
                    h[7] = (-3 * dz * (Gz * dx - Gx * dz) - Gx * r2) * denom;
                    h[8] = -3 * dz * (Gx * dy - Gy * dx) * denom;
                    for (int k = 0; k < 9; ++k) t[7 + k] += h[k];
                }
            }
        }
    }
}

// Point{SourceVortex}: the reference expands it as a point source plus a point vortex
// (src/bodytomultipole.jl:491-496) and holds no direct! for it; its nearfield here is, per source body, the
// point source's terms (test/gravitational.jl:80-104: potential and gradient, plus the gradient's
// derivative q (delta_ij / r^3 - 3 d_i d_j / r^5) / (4 pi), which the far field carries) then the singular
// vortex's (test/vortex.jl:83-128: velocity and its gradient), each written as the switches ask.
static void direct_source_vortex(double* tbuf, int64_t t0, int64_t nt, const double* sbuf, int sr, int64_t s0, int64_t ns,
                                 bool PS, bool GS, bool HS) {
    for (int64_t js = s0; js < s0 + ns; ++js) {
        const double* s = sbuf + (size_t)js * sr;
        const double sx = s[0], sy = s[1], sz = s[2], q = s[4], Gx = s[5], Gy = s[6], Gz = s[7];
        for (int64_t it = t0; it < t0 + nt; ++it) {
            double* t = tbuf + (size_t)it * 16;
            const double dx = t[0] - sx, dy = t[1] - sy, dz = t[2] - sz;
            const double r2 = dx * dx + dy * dy + dz * dz;
            if (!(r2 > 0)) continue;
            const double r = std::sqrt(r2);
            if (PS) t[3] += q / r * ONE_OVER_4PI;
            const double r3 = r2 * r;
            if (GS) {
                t[4] += dx * q / r3 * ONE_OVER_4PI;
                t[5] += dy * q / r3 * ONE_OVER_4PI;
                t[6] += dz * q / r3 * ONE_OVER_4PI;
            }
            const double rinv = 1 / r, rinv2 = rinv * rinv;
            double denom = ONE_OVER_4PI * rinv * rinv2;
            if (GS) {
                t[4] += (dz * Gy - dy * Gz) * denom;
                t[5] += (dx * Gz - dz * Gx) * denom;
                t[6] += (dy * Gx - dx * Gy) * denom;
            }
            if (HS) {
                const double d[3] = {dx, dy, dz};
                const double qs = q * ONE_OVER_4PI / (r3 * r2);
                for (int i = 0; i < 3; ++i)
                    for (int j = 0; j < 3; ++j) t[7 + 3 * i + j] += qs * ((i == j ? r2 : 0.0) - 3 * d[i] * d[j]);
                denom *= rinv2;
                double h[9];
                h[0] = -3 * dx * (Gy * dz - Gz * dy) * denom;
                h[1] = (-3 * dx * (Gz * dx - Gx * dz) + Gz * r2) * denom;
                h[2] = (-3 * dx * (Gx * dy - Gy * dx) - Gy * r2) * denom;
                h[3] = (-3 * dy * (Gy * dz - Gz * dy) - Gz * r2) * denom;
                h[4] = -3 * dy * (Gz * dx - Gx * dz) * denom;
                h[5] = (-3 * dy * (Gx * dy - Gy * dx) + Gx * r2) * denom;
                h[6] = (-3 * dz * (Gy * dz - Gz * dy) + Gy * r2) * denom;
                h[7] = (-3 * dz * (Gz * dx - Gx * dz) - Gx * r2) * denom;
                h[8] = -3 * dz * (Gx * dy - Gy * dx) * denom;
                for (int k = 0; k < 9; ++k) t[7 + k] += h[k];
            }
        }
    }
}

// Point dipole of moment q at x_s. The potential is the one the reference's
// dipole multipole expansion represents (body_to_multipole_point!(Point{Dipole}),
// src/bodytomultipole.jl:431-463; checked analytically at
// test/bodytomultipole_test.jl:508-514): phi = q.d / (4 pi r^3), d = x_t - x_s.
// Outputs follow the expansion evaluation's convention (src/evaluate_expansions.jl):
// slot "gradient" = -grad phi, slot "hessian" = -grad grad phi.
static void direct_dipole(double* tbuf, int64_t t0, int64_t nt, const double* sbuf, int sr, int64_t s0, int64_t ns, bool PS,
                          bool GS, bool HS) {
    for (int64_t js = s0; js < s0 + ns; ++js) {
        const double* s = sbuf + (size_t)js * sr;
        const double qx = s[4], qy = s[5], qz = s[6];
        for (int64_t it = t0; it < t0 + nt; ++it) {
            double* t = tbuf + (size_t)it * 16;
            const double d[3] = {t[0] - s[0], t[1] - s[1], t[2] - s[2]};
            const double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
            if (!(r2 > 0)) continue;
            const double r = std::sqrt(r2), ri = 1 / r, ri2 = ri * ri, ri3 = ri * ri2, ri5 = ri3 * ri2, ri7 = ri5 * ri2;
            const double qd = qx * d[0] + qy * d[1] + qz * d[2];
            const double q[3] = {qx, qy, qz};
            if (PS) t[3] += qd * ri3 * ONE_OVER_4PI;
            if (GS)
                for (int a = 0; a < 3; ++a) t[4 + a] += (3 * qd * d[a] * ri5 - q[a] * ri3) * ONE_OVER_4PI;
            if (HS)
                for (int a = 0; a < 3; ++a)
                    for (int b = 0; b < 3; ++b)
                        t[7 + 3 * a + b] += (3 * (q[b] * d[a] + q[a] * d[b] + (a == b ? qd : 0.0)) * ri5 -
                                             15 * qd * d[a] * d[b] * ri7) * ONE_OVER_4PI;
        }
    }
}

// Gaussian-core vortex (vortex particle with a Gaussian vorticity blob of
// width sigma = buffer row 4): the singular vortex kernel of test/vortex.jl
// scaled by the fraction of the blob inside radius r,
//   g(rho) = erf(rho/sqrt 2) - sqrt(2/pi) rho exp(-rho^2/2),  rho = r/sigma
// (the chi distribution's CDF with 3 degrees of freedom), and
//   grad v = g grad v_singular + g'(rho) (d / (sigma r)) (x) v_singular,
//   g'(rho) = sqrt(2/pi) rho^2 exp(-rho^2/2).
// sigma = 0 gives the singular kernel.
static void direct_vortex_gauss(double* tbuf, int64_t t0, int64_t nt, const double* sbuf, int sr, int64_t s0, int64_t ns,
                                bool V, bool VG) {
    const double c2pi = std::sqrt(2.0 / 3.141592653589793);
    for (int64_t js = s0; js < s0 + ns; ++js) {
        const double* s = sbuf + (size_t)js * sr;
        const double sig = s[3], Gx = s[4], Gy = s[5], Gz = s[6];
        for (int64_t it = t0; it < t0 + nt; ++it) {
            double* t = tbuf + (size_t)it * 16;
            const double dx = t[0] - s[0], dy = t[1] - s[1], dz = t[2] - s[2];
            const double r2 = dx * dx + dy * dy + dz * dz;
            if (!(r2 > 0)) continue;
            const double r = std::sqrt(r2), rinv = 1 / r, rinv2 = rinv * rinv;
            double g = 1.0, gp = 0.0;
            if (sig > 0) {
                const double rho = r / sig, e = std::exp(-0.5 * rho * rho);
                g = std::erf(rho / std::sqrt(2.0)) - c2pi * rho * e;
                gp = c2pi * rho * rho * e;
            }
            const double d3 = ONE_OVER_4PI * rinv * rinv2;
            const double c[3] = {dz * Gy - dy * Gz, dx * Gz - dz * Gx, dy * Gx - dx * Gy};
            if (V)
                for (int a = 0; a < 3; ++a) t[4 + a] += g * c[a] * d3;
            if (VG) {
                const double d5 = d3 * rinv2, dd[3] = {dx, dy, dz};
                double h[9];
                h[0] = -3 * dx * c[0] * d5;
                h[1] = (-3 * dx * c[1] + Gz * r2) * d5;
                h[2] = (-3 * dx * c[2] - Gy * r2) * d5;
                h[3] = (-3 * dy * c[0] - Gz * r2) * d5;
                h[4] = -3 * dy * c[1] * d5;
                h[5] = (-3 * dy * c[2] + Gx * r2) * d5;
                h[6] = (-3 * dz * c[0] + Gy * r2) * d5;
                h[7] = (-3 * dz * c[1] - Gx * r2) * d5;
                h[8] = -3 * dz * c[2] * d5;
                const double f = sig > 0 ? gp / (sig * r) : 0.0;
                for (int a = 0; a < 3; ++a)
                    for (int b = 0; b < 3; ++b) t[7 + 3 * a + b] += g * h[3 * a + b] + f * dd[a] * c[b] * d3;
            }
        }
    }
}


// ---------------------------------------------------------- panels
// Constant-strength planar source / normal-doublet panel with NS vertices:
// induced() of test/panels.jl (:1-11 entry, :78-93 rotate_to_panel, :95-133
// source_dipole_preliminaries, :139-216 source, :224-315 dipole). Returns the
// potential, the velocity slot and the Hessian slot.
static void panel_induced(const double* tgt, const double (*v)[3], int NS, const double* normal, double strength,
                          const double* centroid, bool dipole, bool PS, bool GS, bool HS, double& phi, double vel[3],
                          double H[9]) {
    const double EPS = 2.220446049250313e-16;   // Julia eps()
    // rotate_to_panel: R = [new_x new_y new_z], new_z = normal, new_x = (v3 - v1)/|v3 - v1|, new_y = new_z x new_x
    double nx[3] = {v[2][0] - v[0][0], v[2][1] - v[0][1], v[2][2] - v[0][2]};
    const double nn = std::sqrt(nx[0] * nx[0] + nx[1] * nx[1] + nx[2] * nx[2]);
    for (int a = 0; a < 3; ++a) nx[a] /= nn;
    const double* nz = normal;
    const double ny[3] = {nz[1] * nx[2] - nz[2] * nx[1], nz[2] * nx[0] - nz[0] * nx[2], nz[0] * nx[1] - nz[1] * nx[0]};
    const double* Rc[3] = {nx, ny, nz};   // columns of R = rows of R'
    auto dot = [](const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; };
    const double dt[3] = {tgt[0] - centroid[0], tgt[1] - centroid[1], tgt[2] - centroid[2]};
    const double tR[3] = {dot(nx, dt), dot(ny, dt), dot(nz, dt)};
    double vx[5], vy[5];
    for (int i = 0; i < NS; ++i) {
        const double d[3] = {v[i][0] - centroid[0], v[i][1] - centroid[1], v[i][2] - centroid[2]};
        vx[i] = dot(nx, d);
        vy[i] = dot(ny, d);
    }
    vx[NS] = vx[0];
    vy[NS] = vy[0];
    double dxs[4], dys[4], ds[4], ms[4];
    for (int i = 0; i < NS; ++i) {
        dxs[i] = vx[i + 1] - vx[i];
        dys[i] = vy[i + 1] - vy[i];
        ds[i] = std::sqrt(dxs[i] * dxs[i] + dys[i] * dys[i]);
        ms[i] = dys[i] / dxs[i];
    }
    double dz = tR[2];
    if (dz == 0.0) dz = EPS;
    const double dz2 = dz * dz;
    double rxs[5], rys[5], es[5], hs[5], rs[5], rxr[5], ryr[5];
    for (int i = 0; i <= NS; ++i) {
        rxs[i] = tR[0] - vx[i];
        rys[i] = tR[1] - vy[i];
        es[i] = rxs[i] * rxs[i] + dz2;
        hs[i] = rxs[i] * rys[i];
        const double dy = vy[i] - tR[1];
        rs[i] = std::sqrt(es[i] + dy * dy);
        rxr[i] = rxs[i] / rs[i];
        ryr[i] = rys[i] / rs[i];
    }
    double pot = 0.0, u[3] = {0, 0, 0}, h[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int i = 0; i < NS; ++i) {
        double num = rs[i] + rs[i + 1] - ds[i];
        if (num == 0.0) num += EPS;
        const double log_term = std::log(num / (rs[i] + rs[i + 1] + ds[i]));
        double ri = rs[i];
        if (ri == 0.0) ri += EPS;
        double rip1 = rs[i + 1];
        if (rip1 == 0.0) rip1 += EPS;
        const double tan_term = std::atan((ms[i] * es[i] - hs[i]) / dz / ri) - std::atan((ms[i] * es[i + 1] - hs[i + 1]) / dz / rip1);
        const double dx = dxs[i], dy = dys[i];
        const double r_plus = rs[i] + rs[i + 1], r_times = rs[i] * rs[i + 1];
        const double rho = r_times + rxs[i] * rxs[i + 1] + rys[i] * rys[i + 1] + dz2;
        const double lambda = rxs[i] * rys[i + 1] - rxs[i + 1] * rys[i];
        if (!dipole) {
            if (PS) {
                pot += (rxs[i] * dy - rys[i] * dx) / ds[i] * log_term;
                pot += dz * tan_term;
            }
            if (GS) {
                u[0] += dy / ds[i] * log_term;
                u[1] += -dx / ds[i] * log_term;
                u[2] += tan_term;
            }
            if (HS) {
                const double d2 = ds[i] * ds[i];
                const double val1 = r_plus * r_plus - d2;
                const double val2 = rxr[i] + rxr[i + 1], val3 = ryr[i] + ryr[i + 1];
                const double val4 = r_plus / (r_times * rho);
                const double xx = 2 * dy / val1 * val2, xy = -2 * dx / val1 * val2, xz = dz * dy * val4;
                const double yy = -2 * dx / val1 * val3, yz = -dz * dx * val4, zz = lambda * val4;
                const double m[9] = {xx, xy, xz, xy, yy, yz, xz, yz, zz};
                for (int k = 0; k < 9; ++k) h[k] += m[k];
            }
        } else {
            if (PS) pot -= tan_term;
            if (GS) {
                const double val4 = r_plus / (r_times * rho);
                u[0] += dz * dys[i] * val4;
                u[1] += -dz * dxs[i] * val4;
                u[2] += lambda * val4;
            }
            if (HS) {
                const double rp2 = r_plus * r_plus;
                double val1 = r_times * rp2 + rho * rs[i + 1] * rs[i + 1];
                val1 /= rho * rs[i] * r_plus;
                double val2 = r_times * rp2 + rho * rs[i] * rs[i];
                val2 /= rho * rs[i + 1] * r_plus;
                const double val3 = r_plus / (rho * r_times * r_times);
                const double xx = dz * dys[i] * val3 * (rxs[i] * val1 + rxs[i + 1] * val2);
                const double xy = dz * dys[i] * val3 * (rys[i] * val1 + rys[i + 1] * val2);
                const double yy = -dz * dxs[i] * val3 * (rys[i] * val1 + rys[i + 1] * val2);
                const double val4 = rp2 / rho;
                const double val5 = (rs[i] * rs[i] - r_times + rs[i + 1] * rs[i + 1]) / r_times;
                const double val6 = dz * (val4 + val5);
                const double zz = lambda * val3 * val6;
                const double val7 = r_times - dz * val6;
                const double val8 = val3 * val7;
                const double xz = -dys[i] * val8, yz = dxs[i] * val8;
                const double m[9] = {xx, xy, xz, xy, yy, yz, xz, yz, zz};
                for (int k = 0; k < 9; ++k) h[k] += m[k];
            }
        }
    }
    const double s4 = strength * ONE_OVER_4PI;
    phi = PS ? pot * s4 : 0.0;
    const double sv = dipole ? s4 : -s4, sh = -s4;
    // R u: sum over columns of R
    for (int a = 0; a < 3; ++a) vel[a] = GS ? sv * (Rc[0][a] * u[0] + Rc[1][a] * u[1] + Rc[2][a] * u[2]) : 0.0;
    // R h R': (R h R')[a][b] = sum_jk R[a][j] h[j][k] R[b][k]; h is symmetric
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) {
            double acc = 0.0;
            if (HS)
                for (int j = 0; j < 3; ++j)
                    for (int k = 0; k < 3; ++k) acc += Rc[j][a] * h[3 * j + k] * Rc[k][b];
            H[3 * a + b] = sh * acc;
        }
}

// unit normal of a panel from its first three vertices (get_normal, src/compatibility.jl:65-72)
static void panel_normal(const double (*v)[3], double* n) {
    const double a[3] = {v[1][0] - v[0][0], v[1][1] - v[0][1], v[1][2] - v[0][2]};
    const double b[3] = {v[2][0] - v[0][0], v[2][1] - v[0][1], v[2][2] - v[0][2]};
    n[0] = a[1] * b[2] - a[2] * b[1];
    n[1] = a[2] * b[0] - a[0] * b[2];
    n[2] = a[0] * b[1] - a[1] * b[0];
    const double nn = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    for (int k = 0; k < 3; ++k) n[k] /= nn;
}

static void direct_panels(int kind, double* tbuf, int64_t t0, int64_t nt, const double* sbuf, int sr, int64_t s0, int64_t ns,
                          bool PS, bool GS, bool HS) {
    const int NV = panel_vertices(kind), vo = 4 + panel_strengths(kind);
    for (int64_t js = s0; js < s0 + ns; ++js) {
        const double* s = sbuf + (size_t)js * sr;
        double v[4][3];
        for (int k = 0; k < NV; ++k)
            for (int a = 0; a < 3; ++a) v[k][a] = s[vo + 3 * k + a];
        double n[3];
        panel_normal(v, n);
        for (int64_t it = t0; it < t0 + nt; ++it) {
            double* t = tbuf + (size_t)it * 16;
            double phi, vel[3], H[9];
            // a source-dipole panel is the source's closed form with strength 1 plus the dipole's with strength 2,
            // as test/panels.jl:323-429 sums them edge by edge
            const int n_parts = panel_source_dipole(kind) ? 2 : 1;
            for (int part = 0; part < n_parts; ++part) {
                const bool dip = panel_source_dipole(kind) ? part == 1 : panel_dipole(kind);
                panel_induced(t, v, NV, n, s[4 + part], s, dip, PS, GS, HS, phi, vel, H);
                if (PS) t[3] += phi;
                if (GS) for (int a = 0; a < 3; ++a) t[4 + a] += vel[a];
                if (HS) for (int k = 0; k < 9; ++k) t[7 + k] += H[k];
            }
        }
    }
}

// body_to_multipole! for panels (src/bodytomultipole.jl:643-715): triangles, and quadrilaterals as
// the triangles (v1, v2, v3) and (v1, v3, v4). The reference's quadrilateral path leaves the second
// triangle's normal unnormalised (:705 discards the division); here it is normalised, so that the
// dipole expansion of a quadrilateral matches its induced() field.
static void p2m_panel(int kind, Exp M, Exp harm, const double* s, const double* center, int p) {
    const int NV = panel_vertices(kind), vo = 4 + panel_strengths(kind);
    double v[4][3];
    for (int k = 0; k < NV; ++k)
        for (int a = 0; a < 3; ++a) v[k][a] = s[vo + 3 * k + a];
    for (int tri = 0; tri < NV - 2; ++tri) {
        const double* a = v[0];
        const double* b = v[tri + 1];
        const double* c = v[tri + 2];
        const double x0[3] = {a[0] - center[0], a[1] - center[1], a[2] - center[2]};
        const double xu[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
        const double xv[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
        if (!panel_dipole(kind)) p2m_panel_source(M, harm, x0, xu, xv, s[4], p);
        if (panel_dipole(kind) || panel_source_dipole(kind)) {   // Panel{SourceDipole}: the source's, then the dipole's
            const double vv[3][3] = {{a[0], a[1], a[2]}, {b[0], b[1], b[2]}, {c[0], c[1], c[2]}};
            double n[3];
            panel_normal(vv, n);
            p2m_panel_dipole(M, harm, x0, xu, xv, n, panel_source_dipole(kind) ? s[5] : s[4], p);
        }
    }
}

static void user_direct(int kind, double* tbuf, int64_t t0, int64_t nt, const double* sbuf, int sr, int64_t s0, int64_t ns,
                        bool PS, bool GS, bool HS) {
    if (is_panel(kind)) { direct_panels(kind, tbuf, t0, nt, sbuf, sr, s0, ns, PS, GS, HS); return; }
    if (kind == KIND_SOURCE) direct_gravitational(tbuf, t0, nt, sbuf, sr, s0, ns);
    else if (kind == KIND_SOURCE_VORTEX) direct_source_vortex(tbuf, t0, nt, sbuf, sr, s0, ns, PS, GS, HS);
    else if (kind == KIND_VORTEX) direct_vortex(tbuf, t0, nt, sbuf, sr, s0, ns, GS, HS);
    else if (kind == KIND_DIPOLE) direct_dipole(tbuf, t0, nt, sbuf, sr, s0, ns, PS, GS, HS);
    else direct_vortex_gauss(tbuf, t0, nt, sbuf, sr, s0, ns, GS, HS);
}

static int data_per_body(int kind) {
    return kind == KIND_SOURCE ? 5 : is_panel(kind) ? 4 + panel_strengths(kind) + 3 * panel_vertices(kind)
                                   : kind == KIND_SOURCE_VORTEX ? 8 : 7;
}
static bool has_vector_potential(int kind) { return kind == KIND_VORTEX || kind == KIND_VORTEX_GAUSS || kind == KIND_SOURCE_VORTEX; }

// ------------------------------------------------------------ fmm driver
struct Options {
    int p = 5, ls_source = 20, ls_target = 20, method = METHOD_SELFTUNING_TREESTOP;
    bool shrink = true;
    ErrorMethod em;
    bool PS = false, GS = true, HS = false;
    bool ff = true, nf = true, si = true;
    double theta = 0.4;
};

struct Run {
    int kind;
    // several systems in one tree (fmm! on a tuple, src/fmm.jl:3-40 and src/fmm.jl:331-345): the kernel of
    // each system and its original index range off[s] .. off[s + 1] - 1 of the concatenated 7-row buffer;
    // empty for one system
    std::vector<int> kinds;
    std::vector<int64_t> off;
    int rows = 0;   // rows per body of a tuple's buffer: the largest data_per_body of its kinds
    std::vector<int64_t> sys_ls;   // a tuple's leaf size per system (empty: the options' leaf size for each)
    int64_t nt, ns;
    std::vector<double> tfull, sfull;   // original order: 16 x nt, dpb x ns
    Options opt;
    OTree tt, st;
    std::vector<std::array<int32_t, 2>> m2l, direct;
    int Pmax = 0;
    bool error_success = true;
    std::vector<int32_t> m2l_P;   // expansion order chosen per M2L interaction
    std::vector<double> m2l_margin;   // last_m2l_margin of each M2L interaction
};

static int run_rows(const Run& R) { return R.kinds.empty() ? data_per_body(R.kind) : R.rows; }

// both trees of a run; a tuple's trees take the leaf test of several systems
static void build_trees(Run& R) {
    const Options& o = R.opt;
    if (R.kinds.empty()) {
        build_tree(R.tt, R.nt, 16, R.tfull.data(), true, o.ls_target, o.method, o.shrink, o.p);
        build_tree(R.st, R.ns, run_rows(R), R.sfull.data(), false, o.ls_source, o.method, o.shrink, o.p);
        return;
    }
    std::vector<int64_t> ls = R.sys_ls;
    if (ls.empty()) ls.assign(R.kinds.size(), o.ls_source);
    build_tree(R.tt, R.nt, 16, R.tfull.data(), true, o.ls_target, o.method, o.shrink, o.p, &R.off, &ls);
    build_tree(R.st, R.ns, run_rows(R), R.sfull.data(), false, o.ls_source, o.method, o.shrink, o.p, &R.off, &ls);
}

static BranchGeom geom(const OBranch& b) {
    BranchGeom g;
    for (int d = 0; d < 3; ++d) {
        g.source_center[d] = b.source_center[d]; g.target_center[d] = b.target_center[d];
        g.source_box[d] = b.source_box[d]; g.target_box[d] = b.target_box[d];
    }
    g.source_radius = b.source_radius; g.target_radius = b.target_radius; g.max_influence = b.max_influence;
    return g;
}

static void fmm_run(Run& R, bool estimate_pass);

// estimate_influence! (src/fmm.jl:1193-1276)
static void estimate_influence(Run& R) {
    Run E;
    E.kind = R.kind; E.nt = R.nt; E.ns = R.ns; E.tfull = R.tfull; E.sfull = R.sfull;
    E.kinds = R.kinds; E.off = R.off; E.rows = R.rows;   // leaf size 5 for every system
    E.opt = R.opt;
    E.opt.PS = true; E.opt.GS = true; E.opt.HS = false;
    E.opt.ls_source = 5; E.opt.ls_target = 5;
    E.opt.p = 3; E.opt.theta = 0.6;
    E.opt.em = ErrorMethod();
    E.opt.shrink = true;
    E.opt.method = METHOD_SELFTUNING_TREESTOP;
    E.opt.ff = true; E.opt.nf = true; E.opt.si = true;
    fmm_run(E, true);
    // map estimate buffer (sorted by the estimate tree) onto the main target tree buffer
    std::vector<double> infl(R.nt, 0.0);
    for (int64_t je = 0; je < E.nt; ++je) {
        int64_t jsys = E.tt.sort_index[je];
        int64_t jb = R.tt.inverse_sort_index[jsys];
        const double* e = &E.tt.buffer[(size_t)je * 16];
        double pot = e[3], vx = e[4], vy = e[5], vz = e[6];
        (void)pot;
        if (R.opt.em.kind == 2) infl[jb] = pot;   // PowerRelativePotential reads the scalar potential
        else infl[jb] = std::sqrt(vx * vx + vy * vy + vz * vz);
    }
    for (auto& b : R.tt.branches) {
        double mi = b.max_influence;
        for (int64_t j = b.first; j < b.first + b.n_bodies; ++j) mi = std::max(mi, infl[j]);
        b.max_influence = mi;
    }
}

// bodies of system s in branch b of a source tree: the partition is stable, so within a leaf the bodies
// keep their original order and each system's bodies are one sub-range (bodies_index[i_system])
static void system_range(const Run& R, const OBranch& b, size_t s, int64_t& first, int64_t& n) {
    const int64_t* si = R.st.sort_index.data();
    const int64_t* f = std::lower_bound(si + b.first, si + b.first + b.n_bodies, R.off[s]);
    const int64_t* l = std::lower_bound(si + b.first, si + b.first + b.n_bodies, R.off[s + 1]);
    first = f - si;
    n = l - f;
}

// P2M of one body (src/bodytomultipole.jl:19-26)
static void p2m_body(int kind, Exp M, Exp harm, const double* x, const double* center, int p) {
    double dx[3] = {x[0] - center[0], x[1] - center[1], x[2] - center[2]};
    if (is_panel(kind)) p2m_panel(kind, M, harm, x, center, p);
    else if (kind == KIND_SOURCE) p2m_point_source(M, harm, dx, x[4], p);
    else if (kind == KIND_SOURCE_VORTEX) {   // src/bodytomultipole.jl:491-496: the source, then the vortex
        p2m_point_source(M, harm, dx, x[4], p);
        p2m_point_vortex(M, harm, dx, x + 5, p);
    }
    else if (kind == KIND_DIPOLE) p2m_point_dipole(M, harm, dx, x + 4, p);
    else p2m_point_vortex(M, harm, dx, x + 4, p);
}

static void fmm_run(Run& R, bool estimate_pass) {
    const Options& o = R.opt;
    const bool LH = has_vector_potential(R.kind);
    const int dpb = run_rows(R);
    build_trees(R);
    R.m2l.clear(); R.direct.clear();
    if (!R.tt.branches.empty() && !R.st.branches.empty()) {
        ListBuilder L{R.tt, R.st, o.ls_source, o.theta, o.ff, o.nf, o.si, o.method, {}, {}};
        L.run(0, 0);
        R.m2l = sort_by_target(L.m2l, R.tt.branches.size());
        R.direct = sort_by_target(L.direct, R.tt.branches.size());
    }
    if (!estimate_pass && o.em.kind != 0 && o.em.relative) estimate_influence(R);
    R.Pmax = 0;
    R.error_success = true;
    if (R.nt == 0 || R.ns == 0) return;
    tables().ensure(o.p);
    int p = o.p;
    int Hn = n_harmonics(p);
    // The passes are split over OpenMP threads the way the reference's multithreaded drivers split
    // them (src/fmm.jl:57-93, 348-438, 492-598, 642-757): by target leaf (nearfield, L2P), by source
    // leaf (P2M), by parent branch within a level (M2M), by target branch (M2L) and by child branch
    // within a level (L2L). Every expansion and target value is written by one thread and accumulates
    // its terms in the serial order, so results are bit-identical for any thread count.
    // target ranges of the target-sorted lists
    auto runs_of = [](const std::vector<std::array<int32_t, 2>>& L) {
        std::vector<size_t> r;
        for (size_t k = 0; k < L.size(); ++k)
            if (k == 0 || L[k][0] != L[k - 1][0]) r.push_back(k);
        r.push_back(L.size());
        return r;
    };
    // nearfield (src/fmm.jl:3-40)
    if (R.kinds.empty()) {
        const std::vector<size_t> rd = runs_of(R.direct);
        const int64_t nr = (int64_t)rd.size() - 1;
#pragma omp parallel for schedule(dynamic, 1)
        for (int64_t g = 0; g < nr; ++g)
            for (size_t k = rd[g]; k < rd[g + 1]; ++k) {
                const auto& pr = R.direct[k];
                const OBranch& s = R.st.branches[pr[1]];
                const OBranch& t = R.tt.branches[pr[0]];
                user_direct(R.kind, R.tt.buffer.data(), t.first, t.n_bodies, R.st.buffer.data(), dpb, s.first, s.n_bodies,
                            o.PS, o.GS, o.HS);
            }
    }
    // several systems: each source system in turn over the whole direct list, with its own kernel
    for (size_t is = 0; is < R.kinds.size(); ++is) {
        const std::vector<size_t> rd = runs_of(R.direct);
        const int64_t nr = (int64_t)rd.size() - 1;
#pragma omp parallel for schedule(dynamic, 1)
        for (int64_t g = 0; g < nr; ++g)
            for (size_t k = rd[g]; k < rd[g + 1]; ++k) {
                const auto& pr = R.direct[k];
                const OBranch& t = R.tt.branches[pr[0]];
                int64_t s0, sn;
                system_range(R, R.st.branches[pr[1]], is, s0, sn);
                user_direct(R.kinds[is], R.tt.buffer.data(), t.first, t.n_bodies, R.st.buffer.data(), dpb, s0, sn,
                            o.PS, o.GS, o.HS);
            }
    }
    // upward pass: P2M (src/bodytomultipole.jl:19-26) then M2M (src/fmm.jl:262-284)
    const int64_t nsl = (int64_t)R.st.leaf_index.size();
#pragma omp parallel
    {
        std::vector<double> harm(4 * n_harmonics(p + 2), 0.0);
#pragma omp for schedule(dynamic, 16)
        for (int64_t k = 0; k < nsl; ++k) {
            const int64_t il = R.st.leaf_index[k];
            const OBranch& b = R.st.branches[il];
            Exp M{&R.st.expansions[(size_t)il * 4 * Hn]};
            if (R.kinds.empty())
                for (int64_t i = b.first; i < b.first + b.n_bodies; ++i)
                    p2m_body(R.kind, M, Exp{harm.data()}, &R.st.buffer[(size_t)i * dpb], b.source_center, p);
            for (size_t is = 0; is < R.kinds.size(); ++is) {   // systems in order (src/fmm.jl:331-345)
                int64_t s0, sn;
                system_range(R, b, is, s0, sn);
                for (int64_t i = s0; i < s0 + sn; ++i)
                    p2m_body(R.kinds[is], M, Exp{harm.data()}, &R.st.buffer[(size_t)i * dpb], b.source_center, p);
            }
        }
    }
    for (int64_t lv = (int64_t)R.st.levels.size() - 1; lv >= 0; --lv) {
        const int64_t b0 = R.st.levels[lv][0], b1 = R.st.levels[lv][1];
#pragma omp parallel
        {
            Scratch S;
            S.ensure(p);
#pragma omp for schedule(dynamic, 4)
            for (int64_t ib = b0; ib <= b1; ++ib) {
                const OBranch& b = R.st.branches[ib];
                if (b.n_branches != 0)
                    for (int64_t c = b.child_first; c < b.child_first + b.n_branches; ++c)
                        multipole_to_multipole(Exp{&R.st.expansions[(size_t)ib * 4 * Hn]}, b.source_center,
                                               Exp{&R.st.expansions[(size_t)c * 4 * Hn]}, R.st.branches[c].source_center, S,
                                               p, LH);
            }
        }
    }
    // horizontal pass (src/fmm.jl:455-490)
    R.m2l_P.assign(R.m2l.size(), 0);
    R.m2l_margin.assign(R.m2l.size(), INFINITY);
    std::vector<char> ok_k(R.m2l.size(), 1);
    {
        const std::vector<size_t> rm = runs_of(R.m2l);
        const int64_t nr = (int64_t)rm.size() - 1;
#pragma omp parallel
        {
            Scratch S;
            S.ensure(p);
#pragma omp for schedule(dynamic, 1)
            for (int64_t g = 0; g < nr; ++g)
                for (size_t k = rm[g]; k < rm[g + 1]; ++k) {
                    int32_t it = R.m2l[k][0], js = R.m2l[k][1];
                    bool ok = true;
                    int P = multipole_to_local(Exp{&R.tt.expansions[(size_t)it * 4 * Hn]}, geom(R.tt.branches[it]),
                                               Exp{&R.st.expansions[(size_t)js * 4 * Hn]}, geom(R.st.branches[js]), S, p, LH,
                                               o.em, ok);
                    R.m2l_P[k] = P;
                    R.m2l_margin[k] = last_m2l_margin;
                    ok_k[k] = ok ? 1 : 0;
                }
        }
    }
    R.Pmax = 0;
    for (size_t k = 0; k < R.m2l.size(); ++k) {
        R.Pmax = std::max(R.m2l_P[k], R.Pmax);
        R.error_success = R.error_success && ok_k[k];
    }
    if (!R.error_success) R.Pmax += 1;
    // downward pass: L2L (src/fmm.jl:602-622) then L2P (src/evaluate_expansions.jl:1-43)
    for (size_t lv = 1; lv < R.tt.levels.size(); ++lv) {
        const int64_t c0 = R.tt.levels[lv][0], c1 = R.tt.levels[lv][1];
#pragma omp parallel
        {
            Scratch S;
            S.ensure(p);
#pragma omp for schedule(dynamic, 4)
            for (int64_t c = c0; c <= c1; ++c) {
                const OBranch& ch = R.tt.branches[c];
                const int64_t ib = ch.i_parent;
                local_to_local(Exp{&R.tt.expansions[(size_t)c * 4 * Hn]}, ch.target_center,
                               Exp{&R.tt.expansions[(size_t)ib * 4 * Hn]}, R.tt.branches[ib].target_center, S, p, LH);
            }
        }
    }
    const int64_t ntl = (int64_t)R.tt.leaf_index.size();
#pragma omp parallel
    {
        std::vector<double> harm(4 * n_harmonics(p + 2), 0.0);
        std::vector<double> gnm;
#pragma omp for schedule(dynamic, 16)
        for (int64_t k = 0; k < ntl; ++k) {
            const int64_t il = R.tt.leaf_index[k];
            const OBranch& b = R.tt.branches[il];
            Exp L{&R.tt.expansions[(size_t)il * 4 * Hn]};
            for (int64_t i = b.first; i < b.first + b.n_bodies; ++i) {
                double* t = &R.tt.buffer[(size_t)i * 16];
                double dx[3] = {t[0] - b.target_center[0], t[1] - b.target_center[1], t[2] - b.target_center[2]};
                double phi, v[3], h[9];
                evaluate_local(dx, Exp{harm.data()}, gnm, L, p, LH, o.PS, o.GS, o.HS, phi, v, h);
                if (o.PS) t[3] += phi;
                if (o.GS) { t[4] += v[0]; t[5] += v[1]; t[6] += v[2]; }
                if (o.HS) for (int k2 = 0; k2 < 9; ++k2) t[7 + k2] += h[k2];
            }
        }
    }
}


// ------------------------------------------------------------ FastGaussSeidel
// src/solve.jl:1-1022 for one system that is both target and source. The
// influence is the scalar potential (test/solve_test.jl:11-13) or, for panels, the
// gradient's component along the target panel's unit normal. Differences from
// the reference, both deliberate: the mean squared residual is taken over every
// body of every leaf (the reference's residual! of :995-1021 reuses one leaf-sized
// vector, so its value is that of the last leaves written), and the iteration count
// returned is the number of Gauss-Seidel sweeps performed.
struct FGSolver {
    Run R;
    int influence = 0;                                    // 0 potential, 1 normal gradient (panels)
    std::vector<std::array<int32_t, 2>> sorted, full_direct;   // by source; direct + self pairs by target
    std::vector<int64_t> leaf_of_branch, leaf_start, lpos;     // leaf-order positions (strengths_by_leaf)
    std::vector<std::vector<double>> self, nonself;       // column-major (rows x cols), per leaf
    std::vector<int64_t> ns_rows;
    std::vector<std::array<int64_t, 2>> index_map;        // [first, last) of the leaf's pairs in `sorted`
    std::vector<int64_t> ns_rhs_off;                      // offset of the leaf's rows in `old_infl`
    std::vector<double> strengths, rhs, old_infl, extra;
    std::vector<double> mse_history;
    int sweeps = 0;
};

static void panel_unit_normal(int kind, const double* s, double n[3]);

// influence of a unit-strength source body js on target body it (sorted indices)
static double unit_influence(const FGSolver& F, int64_t it, int64_t js) {
    const Run& R = F.R;
    const int dpb = data_per_body(R.kind);
    double tb[16] = {0};
    for (int d = 0; d < 3; ++d) tb[d] = R.tt.buffer[(size_t)it * 16 + d];
    std::vector<double> sb(R.st.buffer.begin() + (size_t)js * dpb, R.st.buffer.begin() + (size_t)(js + 1) * dpb);
    sb[4] = 1.0;                                          // set_unit_strength! (:62-70)
    user_direct(R.kind, tb, 0, 1, sb.data(), dpb, 0, 1, true, F.influence == 1, false);
    if (F.influence == 0) return tb[3];
    double nrm[3];
    panel_unit_normal(R.kind, &R.st.buffer[(size_t)it * dpb], nrm);
    return tb[4] * nrm[0] + tb[5] * nrm[1] + tb[6] * nrm[2];
}

static double body_influence(const FGSolver& F, const double* t, int64_t it) {   // influence! on a filled target row
    if (F.influence == 0) return t[3];
    double nrm[3];
    panel_unit_normal(F.R.kind, &F.R.st.buffer[(size_t)it * data_per_body(F.R.kind)], nrm);
    return t[4] * nrm[0] + t[5] * nrm[1] + t[6] * nrm[2];
}

static void fgs_setup(FGSolver& F) {
    Run& R = F.R;
    const Options& o = R.opt;
    const int dpb = data_per_body(R.kind);
    build_tree(R.tt, R.nt, 16, R.tfull.data(), true, o.ls_target, o.method, o.shrink, o.p);
    build_tree(R.st, R.ns, dpb, R.sfull.data(), false, o.ls_source, o.method, o.shrink, o.p);
    // lists in discovery order, farfield + nearfield, no self pairs (:569-571)
    ListBuilder L{R.tt, R.st, o.ls_source, o.theta, true, true, false, o.method, {}, {}};
    L.run(0, 0);
    R.m2l = L.m2l;
    R.direct = L.direct;
    // sort_by_source (src/interaction_list.jl:215-245): stable counting sort
    {
        const size_t nb = R.st.branches.size();
        std::vector<int64_t> cnt(nb + 1, 0);
        for (auto& e : R.direct) cnt[e[1] + 1] += 1;
        for (size_t i = 1; i <= nb; ++i) cnt[i] += cnt[i - 1];
        F.sorted.assign(R.direct.size(), {0, 0});
        for (auto& e : R.direct) F.sorted[cnt[e[1]]++] = e;
    }
    // add_self_interactions (:515-536)
    {
        auto full = R.direct;
        for (int64_t b : R.st.leaf_index) full.push_back({(int32_t)b, (int32_t)b});
        F.full_direct = sort_by_target(full, R.tt.branches.size());
    }
    const size_t nl = R.st.leaf_index.size();
    F.leaf_of_branch.assign(R.st.branches.size(), -1);
    F.leaf_start.assign(nl + 1, 0);
    F.lpos.assign(R.ns, 0);
    for (size_t k = 0; k < nl; ++k) {
        const OBranch& b = R.st.branches[R.st.leaf_index[k]];
        F.leaf_of_branch[R.st.leaf_index[k]] = (int64_t)k;
        F.leaf_start[k + 1] = F.leaf_start[k] + b.n_bodies;
        for (int64_t i = 0; i < b.n_bodies; ++i) F.lpos[b.first + i] = F.leaf_start[k] + i;   // map_by_leaf (:470-495)
    }
    // index_by_source (:292-371)
    F.index_map.assign(nl, {0, 0});
    for (size_t a = 0; a < F.sorted.size();) {
        size_t e = a;
        while (e < F.sorted.size() && F.sorted[e][1] == F.sorted[a][1]) ++e;
        F.index_map[F.leaf_of_branch[F.sorted[a][1]]] = {(int64_t)a, (int64_t)e};
        a = e;
    }
    // nonself_influence_matrices (:94-290)
    F.nonself.assign(nl, {});
    F.ns_rows.assign(nl, 0);
    F.ns_rhs_off.assign(nl + 1, 0);
    for (size_t k = 0; k < nl; ++k) {
        const OBranch& sb = R.st.branches[R.st.leaf_index[k]];
        int64_t rows = 0;
        for (int64_t q = F.index_map[k][0]; q < F.index_map[k][1]; ++q) rows += R.tt.branches[F.sorted[q][0]].n_bodies;
        F.ns_rows[k] = rows;
        F.ns_rhs_off[k + 1] = F.ns_rhs_off[k] + rows;
        auto& M = F.nonself[k];
        M.assign((size_t)rows * sb.n_bodies, 0.0);
        int64_t r0 = 0;
        for (int64_t q = F.index_map[k][0]; q < F.index_map[k][1]; ++q) {
            const OBranch& tb = R.tt.branches[F.sorted[q][0]];
            for (int64_t c = 0; c < sb.n_bodies; ++c)
                for (int64_t r = 0; r < tb.n_bodies; ++r)
                    M[(size_t)c * rows + r0 + r] = unit_influence(F, tb.first + r, sb.first + c);
            r0 += tb.n_bodies;
        }
    }
    // self_influence_matrices (:378-468)
    F.self.assign(nl, {});
    for (size_t k = 0; k < nl; ++k) {
        const OBranch& b = R.st.branches[R.st.leaf_index[k]];
        auto& M = F.self[k];
        M.assign((size_t)b.n_bodies * b.n_bodies, 0.0);
        for (int64_t c = 0; c < b.n_bodies; ++c)
            for (int64_t r = 0; r < b.n_bodies; ++r) M[(size_t)c * b.n_bodies + r] = unit_influence(F, b.first + r, b.first + c);
    }
    F.strengths.assign(R.ns, 0.0);
    F.rhs.assign(R.ns, 0.0);
    F.extra.assign(R.ns, 0.0);
    F.old_infl.assign(F.ns_rhs_off[nl], 0.0);
}

// x = A \ b by LU with partial pivoting (Julia's `\` on a square matrix)
static void lu_solve(std::vector<double> A, int64_t n, const double* b, double* x) {
    std::vector<int64_t> piv(n);
    for (int64_t i = 0; i < n; ++i) piv[i] = i;
    for (int64_t k = 0; k < n; ++k) {
        int64_t p = k;
        for (int64_t i = k + 1; i < n; ++i)
            if (std::fabs(A[(size_t)k * n + i]) > std::fabs(A[(size_t)k * n + p])) p = i;
        if (p != k) {
            for (int64_t j = 0; j < n; ++j) std::swap(A[(size_t)j * n + k], A[(size_t)j * n + p]);
            std::swap(piv[k], piv[p]);
        }
        const double d = A[(size_t)k * n + k];
        for (int64_t i = k + 1; i < n; ++i) A[(size_t)k * n + i] /= d;
        for (int64_t j = k + 1; j < n; ++j) {
            const double f = A[(size_t)j * n + k];
            for (int64_t i = k + 1; i < n; ++i) A[(size_t)j * n + i] -= A[(size_t)k * n + i] * f;
        }
    }
    std::vector<double> y(n);
    for (int64_t i = 0; i < n; ++i) {
        double v = b[piv[i]];
        for (int64_t j = 0; j < i; ++j) v -= A[(size_t)j * n + i] * y[j];
        y[i] = v;
    }
    for (int64_t i = n - 1; i >= 0; --i) {
        double v = y[i];
        for (int64_t j = i + 1; j < n; ++j) v -= A[(size_t)j * n + i] * x[j];
        x[i] = v / A[(size_t)i * n + i];
    }
}

// update_nonself_influence! for one source leaf (:740-795)
static void fgs_nonself_update(FGSolver& F, size_t k) {
    const Run& R = F.R;
    const OBranch& sb = R.st.branches[R.st.leaf_index[k]];
    const int64_t rows = F.ns_rows[k];
    if (rows == 0) return;
    const double* s = &F.strengths[F.leaf_start[k]];
    double* old = &F.old_infl[F.ns_rhs_off[k]];
    std::vector<double> now(rows, 0.0);
    for (int64_t c = 0; c < sb.n_bodies; ++c)
        for (int64_t r = 0; r < rows; ++r) now[r] += F.nonself[k][(size_t)c * rows + r] * s[c];
    int64_t r0 = 0;
    for (int64_t q = F.index_map[k][0]; q < F.index_map[k][1]; ++q) {
        const OBranch& tb = R.tt.branches[F.sorted[q][0]];
        for (int64_t r = 0; r < tb.n_bodies; ++r) {
            const int64_t d = F.lpos[tb.first + r];
            F.rhs[d] += old[r0 + r];
            F.rhs[d] -= now[r0 + r];
        }
        r0 += tb.n_bodies;
    }
    for (int64_t r = 0; r < rows; ++r) old[r] = now[r];
}

// farfield-only fmm! on the solver's trees and M2L list (:839-850): potential and gradient
static void fgs_farfield(FGSolver& F) {
    Run& R = F.R;
    const int p = R.opt.p, Hn = n_harmonics(p), dpb = data_per_body(R.kind);
    const bool LH = has_vector_potential(R.kind);
    tables().ensure(p);
    for (int64_t i = 0; i < R.nt; ++i)
        for (int k = 3; k < 16; ++k) R.tt.buffer[(size_t)i * 16 + k] = 0.0;
    std::fill(R.tt.expansions.begin(), R.tt.expansions.end(), 0.0);
    std::fill(R.st.expansions.begin(), R.st.expansions.end(), 0.0);
    Scratch S;
    S.ensure(p);
    std::vector<double> harm(4 * n_harmonics(p + 2), 0.0);
    for (int64_t il : R.st.leaf_index) {
        const OBranch& b = R.st.branches[il];
        Exp M{&R.st.expansions[(size_t)il * 4 * Hn]};
        for (int64_t i = b.first; i < b.first + b.n_bodies; ++i) {
            const double* x = &R.st.buffer[(size_t)i * dpb];
            double dx[3] = {x[0] - b.source_center[0], x[1] - b.source_center[1], x[2] - b.source_center[2]};
            if (is_panel(R.kind)) p2m_panel(R.kind, M, Exp{harm.data()}, x, b.source_center, p);
            else p2m_point_source(M, Exp{harm.data()}, dx, x[4], p);
        }
    }
    for (int64_t ib = (int64_t)R.st.branches.size() - 1; ib >= 0; --ib) {
        const OBranch& b = R.st.branches[ib];
        if (b.n_branches != 0)
            for (int64_t c = b.child_first; c < b.child_first + b.n_branches; ++c)
                multipole_to_multipole(Exp{&R.st.expansions[(size_t)ib * 4 * Hn]}, b.source_center,
                                       Exp{&R.st.expansions[(size_t)c * 4 * Hn]}, R.st.branches[c].source_center, S, p, LH);
    }
    for (auto& pr : R.m2l) {
        bool ok = true;
        multipole_to_local(Exp{&R.tt.expansions[(size_t)pr[0] * 4 * Hn]}, geom(R.tt.branches[pr[0]]),
                           Exp{&R.st.expansions[(size_t)pr[1] * 4 * Hn]}, geom(R.st.branches[pr[1]]), S, p, LH, ErrorMethod(), ok);
    }
    for (size_t ib = 0; ib < R.tt.branches.size(); ++ib) {
        const OBranch& b = R.tt.branches[ib];
        if (b.n_branches > 0)
            for (int64_t c = b.child_first; c < b.child_first + b.n_branches; ++c)
                local_to_local(Exp{&R.tt.expansions[(size_t)c * 4 * Hn]}, R.tt.branches[c].target_center,
                               Exp{&R.tt.expansions[ib * 4 * Hn]}, b.target_center, S, p, LH);
    }
    std::vector<double> gnm;
    for (int64_t il : R.tt.leaf_index) {
        const OBranch& b = R.tt.branches[il];
        Exp Lx{&R.tt.expansions[(size_t)il * 4 * Hn]};
        for (int64_t i = b.first; i < b.first + b.n_bodies; ++i) {
            double* t = &R.tt.buffer[(size_t)i * 16];
            double dx[3] = {t[0] - b.target_center[0], t[1] - b.target_center[1], t[2] - b.target_center[2]};
            double phi, v[3], h[9];
            evaluate_local(dx, Exp{harm.data()}, gnm, Lx, p, LH, true, true, false, phi, v, h);
            t[3] += phi; t[4] += v[0]; t[5] += v[1]; t[6] += v[2];
        }
    }
}

// unit normal of a panel body from its buffer row (vertices from row 5 on)
static void panel_unit_normal(int kind, const double* s, double n[3]) {
    (void)kind;
    double v[3][3];
    for (int k = 0; k < 3; ++k)
        for (int a = 0; a < 3; ++a) v[k][a] = s[5 + 3 * k + a];
    panel_normal(v, n);
}

// solve! (:797-946); ext = the external influence per body (original order)
static void fgs_solve(FGSolver& F, const double* ext, int max_iterations, double tolerance) {
    Run& R = F.R;
    const int dpb = data_per_body(R.kind);
    const size_t nl = R.st.leaf_index.size();
    for (int64_t i = 0; i < R.ns; ++i) F.extra[F.lpos[i]] = -ext[R.st.sort_index[i]];   // influence! (:960-993)
    F.rhs = F.extra;
    for (int64_t i = 0; i < R.ns; ++i) F.strengths[F.lpos[i]] = R.st.buffer[(size_t)i * dpb + 4];   // update_by_leaf! (:658-688)
    std::fill(F.old_infl.begin(), F.old_infl.end(), 0.0);
    for (size_t k = 0; k < nl; ++k) fgs_nonself_update(F, k);
    F.mse_history.clear();
    F.sweeps = 0;
    for (int it = 1; it <= max_iterations; ++it) {
        fgs_farfield(F);
        for (int64_t i = 0; i < R.ns; ++i) F.extra[F.lpos[i]] = -body_influence(F, &R.tt.buffer[(size_t)i * 16], i);
        for (int64_t i = 0; i < R.ns; ++i) F.rhs[i] += F.extra[i];
        double mse = 0.0;
        for (size_t k = 0; k < nl; ++k) {
            const int64_t n = F.leaf_start[k + 1] - F.leaf_start[k];
            const double* s = &F.strengths[F.leaf_start[k]];
            for (int64_t r = 0; r < n; ++r) {
                double v = -F.rhs[F.leaf_start[k] + r];
                for (int64_t c = 0; c < n; ++c) v += F.self[k][(size_t)c * n + r] * s[c];
                mse += v * v;
            }
        }
        mse /= (double)std::max<int64_t>(1, R.ns);
        F.mse_history.push_back(mse);
        if (mse <= tolerance) break;
        for (size_t k = 0; k < nl; ++k) {
            const int64_t n = F.leaf_start[k + 1] - F.leaf_start[k];
            lu_solve(F.self[k], n, &F.rhs[F.leaf_start[k]], &F.strengths[F.leaf_start[k]]);
            fgs_nonself_update(F, k);
        }
        ++F.sweeps;
        for (int64_t i = 0; i < R.ns; ++i) R.st.buffer[(size_t)i * dpb + 4] = F.strengths[F.lpos[i]];   // update_by_leaf!
        for (int64_t i = 0; i < R.ns; ++i) F.rhs[i] -= F.extra[i];
    }
    // final nearfield with the full direct list on top of the last farfield (:917-931)
    for (auto& pr : F.full_direct) {
        const OBranch& s = R.st.branches[pr[1]];
        const OBranch& t = R.tt.branches[pr[0]];
        user_direct(R.kind, R.tt.buffer.data(), t.first, t.n_bodies, R.st.buffer.data(), dpb, s.first, s.n_bodies, true, true, false);
    }
}

}  // namespace orc

// ======================================================================
// Flat C API used by tests/ (ctypes) and bench.py's cpu_baseline leg.
// ======================================================================
using namespace orc;

extern "C" {

// opts_i: [p, ls_source, ls_target, method, shrink, err_kind, err_relative, err_BE,
//          PS, GS, HS, farfield, nearfield, self_induced]; opts_d: [theta, eps]
// tfull: 16 x nt (positions in rows 1:3), sfull: data_per_body x ns, both original order.
void* orc_run_create(int kind, int64_t nt, const double* tfull, int64_t ns, const double* sfull, const int* oi, const double* od) {
    Run* R = new Run();
    R->kind = kind; R->nt = nt; R->ns = ns;
    R->tfull.assign(tfull, tfull + (size_t)16 * nt);
    R->sfull.assign(sfull, sfull + (size_t)data_per_body(kind) * ns);
    Options& o = R->opt;
    o.p = oi[0]; o.ls_source = oi[1]; o.ls_target = oi[2]; o.method = oi[3]; o.shrink = oi[4] != 0;
    o.em.kind = oi[5]; o.em.relative = oi[6] != 0; o.em.BE = oi[7] != 0;
    o.PS = oi[8] != 0; o.GS = oi[9] != 0; o.HS = oi[10] != 0;
    o.ff = oi[11] != 0; o.nf = oi[12] != 0; o.si = oi[13] != 0;
    o.theta = od[0]; o.em.eps = od[1];
    return R;
}

// several systems in one tree: n_sys kernels (0 source, 1 vortex, 2 dipole, 3 Gaussian vortex, 4..7 panels)
// with counts[s] bodies each, concatenated system-major in sfull (rows x sum(counts), rows the largest
// data_per_body of the kinds and at least 7; each system's rows first, zeros after them); the targets are
// the same bodies (tfull 16 x sum(counts)). The Lamb-Helmholtz terms are on when any system has a vector
// potential (src/compatibility.jl:292-298): the run's kernel tag is then a vortex kernel, else a dipole one.
// leaf_sizes: one per system, or null for the options' leaf size for each
void* orc_run_create_tuple(int n_sys, const int* kinds, const int64_t* counts, const int64_t* leaf_sizes, const double* tfull,
                           const double* sfull, const int* oi, const double* od) {
    std::vector<int64_t> off(n_sys + 1, 0);
    bool lh = false;
    int rows = 7;
    for (int s = 0; s < n_sys; ++s) {
        off[s + 1] = off[s] + counts[s];
        lh = lh || has_vector_potential(kinds[s]);
        rows = std::max(rows, data_per_body(kinds[s]));
    }
    Run* R = (Run*)orc_run_create(lh ? KIND_VORTEX : KIND_DIPOLE, off[n_sys], tfull, 0, sfull, oi, od);
    R->ns = off[n_sys];
    R->sfull.assign(sfull, sfull + (size_t)rows * R->ns);
    R->kinds.assign(kinds, kinds + n_sys);
    R->off = off;
    R->rows = rows;
    if (leaf_sizes) R->sys_ls.assign(leaf_sizes, leaf_sizes + n_sys);
    return R;
}

void orc_run_free(void* h) { delete (Run*)h; }
// OpenMP threads of the following calls (0: leave as is); returns the count in effect
int orc_threads(int n) {
    if (n > 0) omp_set_num_threads(n);
    return omp_get_max_threads();
}
void orc_run_fmm(void* h) { fmm_run(*(Run*)h, false); }

// tree only (no lists, no evaluation): which = 0 target tree, 1 source tree
void orc_run_trees(void* h) { build_trees(*(Run*)h); }

void orc_run_lists(void* h) {
    Run& R = *(Run*)h;
    R.m2l.clear(); R.direct.clear();
    if (R.tt.branches.empty() || R.st.branches.empty()) return;
    ListBuilder L{R.tt, R.st, R.opt.ls_source, R.opt.theta, R.opt.ff, R.opt.nf, R.opt.si, R.opt.method, {}, {}};
    L.run(0, 0);
    R.m2l = sort_by_target(L.m2l, R.tt.branches.size());
    R.direct = sort_by_target(L.direct, R.tt.branches.size());
}

static OTree& tree_of(Run& R, int which) { return which == 0 ? R.tt : R.st; }

int64_t orc_n_branches(void* h, int which) { return (int64_t)tree_of(*(Run*)h, which).branches.size(); }
int64_t orc_n_levels(void* h, int which) { return (int64_t)tree_of(*(Run*)h, which).levels.size(); }
int64_t orc_n_leaves(void* h, int which) { return (int64_t)tree_of(*(Run*)h, which).leaf_index.size(); }

// ints: nb x 6 [n_bodies, first, n_branches, child_first, i_parent, i_leaf]
// reals: nb x 17 [source_center(3), target_center(3), source_radius, target_radius,
//                 source_box(3), target_box(3), max_influence, pad, pad]
void orc_get_branches(void* h, int which, int64_t* ints, double* reals) {
    OTree& T = tree_of(*(Run*)h, which);
    for (size_t i = 0; i < T.branches.size(); ++i) {
        const OBranch& b = T.branches[i];
        int64_t* I = ints + 6 * i;
        I[0] = b.n_bodies; I[1] = b.first; I[2] = b.n_branches; I[3] = b.child_first; I[4] = b.i_parent; I[5] = b.i_leaf;
        double* D = reals + 17 * i;
        for (int d = 0; d < 3; ++d) {
            D[d] = b.source_center[d]; D[3 + d] = b.target_center[d];
            D[8 + d] = b.source_box[d]; D[11 + d] = b.target_box[d];
        }
        D[6] = b.source_radius; D[7] = b.target_radius; D[14] = b.max_influence; D[15] = 0; D[16] = 0;
    }
}

void orc_get_levels(void* h, int which, int64_t* out) {
    OTree& T = tree_of(*(Run*)h, which);
    for (size_t i = 0; i < T.levels.size(); ++i) { out[2 * i] = T.levels[i][0]; out[2 * i + 1] = T.levels[i][1]; }
}
void orc_get_leaf_index(void* h, int which, int64_t* out) {
    OTree& T = tree_of(*(Run*)h, which);
    std::memcpy(out, T.leaf_index.data(), T.leaf_index.size() * sizeof(int64_t));
}
void orc_get_sort_index(void* h, int which, int64_t* out) {
    OTree& T = tree_of(*(Run*)h, which);
    std::memcpy(out, T.sort_index.data(), T.sort_index.size() * sizeof(int64_t));
}
// sorted buffer of a tree (rows x N)
void orc_get_buffer(void* h, int which, double* out) {
    OTree& T = tree_of(*(Run*)h, which);
    std::memcpy(out, T.buffer.data(), T.buffer.size() * sizeof(double));
}
void orc_get_expansions(void* h, int which, double* out) {
    OTree& T = tree_of(*(Run*)h, which);
    std::memcpy(out, T.expansions.data(), T.expansions.size() * sizeof(double));
}
int64_t orc_n_list(void* h, int which) { Run& R = *(Run*)h; return (int64_t)(which == 0 ? R.m2l.size() : R.direct.size()); }
void orc_get_list(void* h, int which, int32_t* out) {
    Run& R = *(Run*)h;
    auto& L = which == 0 ? R.m2l : R.direct;
    for (size_t i = 0; i < L.size(); ++i) { out[2 * i] = L[i][0]; out[2 * i + 1] = L[i][1]; }
}
void orc_get_m2l_orders(void* h, int32_t* out) {
    Run& R = *(Run*)h;
    std::memcpy(out, R.m2l_P.data(), R.m2l_P.size() * sizeof(int32_t));
}
void orc_get_m2l_margins(void* h, double* out) {
    Run& R = *(Run*)h;
    std::memcpy(out, R.m2l_margin.data(), R.m2l_margin.size() * sizeof(double));
}
// target results in the ORIGINAL body order (16 x nt)
void orc_get_targets(void* h, double* out) {
    Run& R = *(Run*)h;
    for (int64_t i = 0; i < R.nt; ++i)
        std::memcpy(out + (size_t)R.tt.sort_index[i] * 16, &R.tt.buffer[(size_t)i * 16], 16 * sizeof(double));
}
int orc_pmax(void* h) { return ((Run*)h)->Pmax; }
int orc_error_success(void* h) { return ((Run*)h)->error_success ? 1 : 0; }

// direct! over all pairs (src/direct.jl:43-75); tbuf 16 x nt (positions set, rest accumulates)
void orc_direct(int kind, int64_t nt, double* tbuf, int64_t ns, const double* sbuf, int PS, int GS, int HS) {
    user_direct(kind, tbuf, 0, nt, sbuf, data_per_body(kind), 0, ns, PS != 0, GS != 0, HS != 0);
}
// direct over a subset of targets (index list) against all sources: for sampled-error checks
void orc_direct_sampled(int kind, int64_t nsamp, const int64_t* idx, const double* txyz, double* out16,
                        int64_t ns, const double* sbuf, int PS, int GS, int HS) {
    for (int64_t k = 0; k < nsamp; ++k) {
        double t[16] = {0};
        for (int d = 0; d < 3; ++d) t[d] = txyz[3 * idx[k] + d];
        user_direct(kind, t, 0, 1, sbuf, data_per_body(kind), 0, ns, PS != 0, GS != 0, HS != 0);
        std::memcpy(out16 + 16 * k, t, sizeof(t));
    }
}

// ---------------------------------------------- unit-level operator entry points
void orc_tables(int p, double* Hs, double* zetas, double* etas, double* Mt, double* Lt) {
    Tables& T = tables();
    T.ensure(p);
    if (Hs) std::memcpy(Hs, T.Hs_pi2.data(), length_Hs(p) * sizeof(double));
    if (zetas) std::memcpy(zetas, T.zetas.data(), length_Ts(p) * sizeof(double));
    if (etas) std::memcpy(etas, T.etas.data(), length_Ts(p) * sizeof(double));
    if (Mt) std::memcpy(Mt, T.Mt.data(), n_harmonics(p) * sizeof(double));
    if (Lt) std::memcpy(Lt, T.Lt.data(), n_harmonics(p) * sizeof(double));
}
void orc_hs_fresh(int p, double* out) {   // update_Hs_pi2! starting from [1.0]
    std::vector<double> Hs(length_Hs(p), 0.0);
    Hs[0] = 1.0;
    if (p >= 1) compute_Hs(Hs, PI / 2, 1, p);
    std::memcpy(out, Hs.data(), Hs.size() * sizeof(double));
}
void orc_pnms(int p, double s, double c, double* out) { auto v = compute_Pnms(p, s, c); std::memcpy(out, v.data(), v.size() * sizeof(double)); }
void orc_update_Ts(double beta, int p, double* out) {
    tables().ensure(p);
    update_Ts(out, tables().Hs_pi2.data(), beta, p);
}
void orc_eimphis(double phi, int P, double* out) { update_eimphis(out, phi, P); }
void orc_zeta_sign(double mag, int mp, int m, double* out) { zeta_sign(mag, mp, m, out[0], out[1]); }
void orc_eta_sign(double mag, int mp, int m, double* out) { eta_sign(mag, mp, m, out[0], out[1]); }
void orc_regular_harmonics(double rho, double th, double ph, int p, double* out) { regular_harmonics(Exp{out}, rho, th, ph, p); }
void orc_irregular_harmonics(double rho, double th, double ph, int p, double* out) { irregular_harmonics(Exp{out}, rho, th, ph, p); }
void orc_c2s(double x, double y, double z, double* out) { cartesian_to_spherical(x, y, z, out[0], out[1], out[2]); }
double orc_factorial(int k) { return factorial_to_double(k); }

void orc_rotate_multipole_y(const double* in, double theta, int p, int LH, double* out) {
    tables().ensure(p);
    std::vector<double> Ts(length_Ts(p)), tmp(in, in + 4 * n_harmonics(p));
    update_Ts(Ts.data(), tables().Hs_pi2.data(), theta, p);
    rotate_multipole_y_full(Exp{out}, Exp{tmp.data()}, Ts.data(), tables().zetas.data(), p, LH != 0);
}
void orc_rotate_local_y(const double* in, double theta, int p, int LH, double* out) {
    tables().ensure(p);
    std::vector<double> Ts(length_Ts(p)), tmp(in, in + 4 * n_harmonics(p));
    update_Ts(Ts.data(), tables().Hs_pi2.data(), theta, p);
    rotate_local_y_full(Exp{out}, Exp{tmp.data()}, Ts.data(), tables().etas.data(), p, LH != 0);
}
void orc_rotate_z(const double* in, double phi, int p, int LH, double* out) {
    std::vector<double> e(2 * (p + 1)), tmp(in, in + 4 * n_harmonics(p));
    rotate_z(Exp{out}, Exp{tmp.data()}, e.data(), phi, p, LH != 0);
}
void orc_translate_multipole_z(const double* in, double t, int p, int LH, double* out) {
    std::vector<double> tmp(in, in + 4 * n_harmonics(p));
    translate_multipole_z(Exp{out}, Exp{tmp.data()}, t, p, LH != 0);
}
void orc_translate_m2l_z(const double* in, double t, int p, int LH, double* out) {
    std::vector<double> tmp(in, in + 4 * n_harmonics(p));
    translate_multipole_to_local_z(Exp{out}, Exp{tmp.data()}, t, p, LH != 0);
}
void orc_translate_local_z(const double* in, double t, int p, int LH, double* out) {
    std::vector<double> tmp(in, in + 4 * n_harmonics(p));
    translate_local_z(Exp{out}, Exp{tmp.data()}, t, p, LH != 0);
}
void orc_transform_lh_multipole(double* e, double r, int p) { transform_lh_multipole(Exp{e}, r, p); }
void orc_transform_lh_local(double* e, double r, int p) { transform_lh_local(Exp{e}, r, p); }
void orc_m2m(const double* tc, const double* sc, const double* in, int p, int LH, double* out) {
    Scratch s;
    std::vector<double> tmp(in, in + 4 * n_harmonics(p));
    multipole_to_multipole(Exp{out}, tc, Exp{tmp.data()}, sc, s, p, LH != 0);
}
void orc_l2l(const double* tc, const double* sc, const double* in, int p, int LH, double* out) {
    Scratch s;
    std::vector<double> tmp(in, in + 4 * n_harmonics(p));
    local_to_local(Exp{out}, tc, Exp{tmp.data()}, sc, s, p, LH != 0);
}
// tgeom/sgeom: [source_center(3), target_center(3), source_radius, target_radius, source_box(3), target_box(3), max_influence]
int orc_m2l(const double* tgeom, const double* sgeom, const double* in, int p, int LH, int err_kind, int err_rel,
            double eps, int BE, double* out, int* success) {
    auto G = [](const double* g) {
        BranchGeom b;
        for (int d = 0; d < 3; ++d) { b.source_center[d] = g[d]; b.target_center[d] = g[3 + d]; b.source_box[d] = g[8 + d]; b.target_box[d] = g[11 + d]; }
        b.source_radius = g[6]; b.target_radius = g[7]; b.max_influence = g[14];
        return b;
    };
    ErrorMethod em;
    em.kind = err_kind; em.relative = err_rel != 0; em.eps = eps; em.BE = BE != 0;
    Scratch s;
    std::vector<double> tmp(in, in + 4 * n_harmonics(p));
    bool ok = true;
    int P = multipole_to_local(Exp{out}, G(tgeom), Exp{tmp.data()}, G(sgeom), s, p, LH != 0, em, ok);
    *success = ok ? 1 : 0;
    return P;
}
void orc_l2p(const double* dx, const double* local, int p, int LH, int PS, int GS, int HS, double* out13) {
    std::vector<double> harm(4 * n_harmonics(p + 2), 0.0), gnm;
    std::vector<double> tmp(local, local + 4 * n_harmonics(p));
    double phi, v[3], h[9];
    evaluate_local(dx, Exp{harm.data()}, gnm, Exp{tmp.data()}, p, LH != 0, PS != 0, GS != 0, HS != 0, phi, v, h);
    out13[0] = phi;
    for (int k = 0; k < 3; ++k) out13[1 + k] = v[k];
    for (int k = 0; k < 9; ++k) out13[4 + k] = h[k];
}
// panel row s = [centroid(3), radius, strength, v1, v2, v3, (v4)] of kind 4..7; accumulates into out
void orc_p2m_panel(int kind, const double* s, const double* center, int p, double* out) {
    std::vector<double> harm(4 * n_harmonics(p + 2), 0.0);
    p2m_panel(kind, Exp{out}, Exp{harm.data()}, s, center, p);
}

// kind 0 source (strength[0]), 1 vortex (strength[0..2]), 2 dipole (strength[0..2]); accumulates into out
void orc_p2m(int kind, const double* dx, const double* strength, int p, double* out) {
    std::vector<double> harm(4 * n_harmonics(p + 2), 0.0);
    if (kind == 0) p2m_point_source(Exp{out}, Exp{harm.data()}, dx, strength[0], p);
    else if (kind == 1) p2m_point_vortex(Exp{out}, Exp{harm.data()}, dx, strength, p);
    else p2m_point_dipole(Exp{out}, Exp{harm.data()}, dx, strength, p);
}


// FastGaussSeidel: oi = [p, leaf_size, shrink_recenter, influence], od = [multipole_acceptance]
void* orc_fgs_create(int kind, int64_t n, const double* sfull, const int* oi, const double* od) {
    FGSolver* F = new FGSolver();
    Run& R = F->R;
    R.kind = kind; R.nt = n; R.ns = n;
    const int dpb = data_per_body(kind);
    R.sfull.assign(sfull, sfull + (size_t)dpb * n);
    R.tfull.assign((size_t)16 * n, 0.0);
    for (int64_t i = 0; i < n; ++i)
        for (int d = 0; d < 3; ++d) R.tfull[(size_t)i * 16 + d] = sfull[(size_t)i * dpb + d];
    R.opt.p = oi[0]; R.opt.ls_source = R.opt.ls_target = oi[1]; R.opt.shrink = oi[2] != 0;
    R.opt.method = METHOD_BARBA; R.opt.theta = od[0];
    R.opt.PS = true; R.opt.GS = true; R.opt.HS = false;
    F->influence = oi[3];
    fgs_setup(*F);
    return F;
}
void orc_fgs_free(void* h) { delete (FGSolver*)h; }
int64_t orc_fgs_n_leaves(void* h) { return (int64_t)((FGSolver*)h)->R.st.leaf_index.size(); }
// per leaf: [branch, n_bodies, nonself rows, first sorted body]
void orc_fgs_leaf_info(void* h, int64_t* out) {
    FGSolver& F = *(FGSolver*)h;
    for (size_t k = 0; k < F.R.st.leaf_index.size(); ++k) {
        const OBranch& b = F.R.st.branches[F.R.st.leaf_index[k]];
        out[4 * k] = F.R.st.leaf_index[k]; out[4 * k + 1] = b.n_bodies; out[4 * k + 2] = F.ns_rows[k]; out[4 * k + 3] = b.first;
    }
}
void orc_fgs_self(void* h, int64_t k, double* out) { auto& M = ((FGSolver*)h)->self[k]; std::memcpy(out, M.data(), M.size() * 8); }
void orc_fgs_nonself(void* h, int64_t k, double* out) { auto& M = ((FGSolver*)h)->nonself[k]; std::memcpy(out, M.data(), M.size() * 8); }
// original index of the body of every nonself row of leaf k, and of the leaf's own bodies
void orc_fgs_nonself_rows(void* h, int64_t k, int64_t* out) {
    FGSolver& F = *(FGSolver*)h;
    int64_t r = 0;
    for (int64_t q = F.index_map[k][0]; q < F.index_map[k][1]; ++q) {
        const OBranch& tb = F.R.tt.branches[F.sorted[q][0]];
        for (int64_t i = 0; i < tb.n_bodies; ++i) out[r++] = F.R.tt.sort_index[tb.first + i];
    }
}
void orc_fgs_leaf_bodies(void* h, int64_t k, int64_t* out) {
    FGSolver& F = *(FGSolver*)h;
    const OBranch& b = F.R.st.branches[F.R.st.leaf_index[k]];
    for (int64_t i = 0; i < b.n_bodies; ++i) out[i] = F.R.st.sort_index[b.first + i];
}
// solve!: strengths and the 16 x n target buffer back in original order
int orc_fgs_solve(void* h, const double* ext, int max_iterations, double tolerance, double* strengths, double* targets,
                  double* mse_history) {
    FGSolver& F = *(FGSolver*)h;
    fgs_solve(F, ext, max_iterations, tolerance);
    const int dpb = data_per_body(F.R.kind);
    for (int64_t i = 0; i < F.R.ns; ++i) {
        strengths[F.R.st.sort_index[i]] = F.R.st.buffer[(size_t)i * dpb + 4];
        std::memcpy(targets + (size_t)F.R.tt.sort_index[i] * 16, &F.R.tt.buffer[(size_t)i * 16], 16 * sizeof(double));
    }
    for (size_t k = 0; k < F.mse_history.size(); ++k) mse_history[k] = F.mse_history[k];
    return F.sweeps;
}
}  // extern "C"
