// FastGaussSeidel on the GPU (src/solve.jl): one system that is both target and
// source (point sources, or source panels), influence = the scalar potential
// (test/solve_test.jl:11-13) or, for panels, the gradient along the target
// panel's unit normal.
//
// Set-up (FastGaussSeidel, :538-645): the evaluation's trees and Barba lists
// (farfield + nearfield, no self pairs) are built by the caller; here the direct
// list in discovery order is sorted by source (sort_by_source), every source
// leaf gets its non-self influence matrix (rows: the bodies of the target leaves
// that take it in their direct list, in that order; columns: its bodies) and
// every leaf its self matrix, all in HBM, column-major as in the reference. The
// self matrices are inverted once (Gauss-Jordan with partial pivoting, one
// workgroup per leaf, in LDS) so that each block solve of a sweep is one
// matrix-vector product; the reference factorises at every solve (`mat \ rhs`),
// which differs by rounding only.
//
// solve! (:797-946): right-hand side and strengths are kept in leaf order
// (strengths_by_leaf). Each iteration runs the farfield-only evaluation on the
// current strengths, moves its influence into the right-hand side, measures the
// mean squared residual of the self blocks over all bodies, and, above the
// tolerance, sweeps the leaves in order: block solve, then the leaf's non-self
// influence on its neighbours (update_nonself_influence!, :740-795). The sweep keeps
// the reference's order (leaf k sees the updates of exactly the leaves < k) while
// solving independent leaves together: leaves are grouped into levels of the
// dependency order and each level is one launch. After the iterations the nearfield
// of the new strengths, self pairs included, is added to the last farfield (:917-931).
#include "fmmhip_internal.hpp"
#include <cstdlib>
#include "panels.hpp"
#include "scan3.hpp"
#include "radix_sort.hpp"
#include <algorithm>
#include <vector>

namespace fmmhip {

static constexpr int FGS_LDS_LEAF = 101;    // self blocks of up to this many bodies are inverted in LDS (2 n^2 doubles <= 160 KiB)
static constexpr int FGS_MAX_LEAF = 4096;   // a sweep keeps a leaf's right-hand side and strengths in LDS (2 n doubles <= 64 KiB)
static constexpr int FGS_STAGE = 101;       // panels staged in LDS at a time by the influence-block kernel
static constexpr double FGS_INV4PI = 1.0 / (4.0 * 3.141592653589793);

struct FgsState {
    int influence = 0;
    int p = 0;
    int64_t nl = 0, rows = 0;
    int max_leaf = 0;              // bodies of the largest leaf
    Lists full;                    // direct list with self pairs, for the final nearfield
    DBuf<int64_t> leaf_start;      // nl + 1: leaf-order offsets (strengths_by_leaf)
    DBuf<int64_t> lpos;            // sorted body -> leaf-order position
    DBuf<int64_t> self_off;        // nl + 1: offsets of the n^2 self blocks
    DBuf<double> self_A, self_inv;
    DBuf<uint32_t> by_src_s;       // direct list sorted by source (sort_by_source)
    DBuf<int32_t> by_src_t;
    DBuf<int64_t> seg;             // 2 nl: [first, last) of each leaf's pairs in the by-source list
    DBuf<int64_t> pair_rows;       // n_direct + 1: exclusive prefix of target leaf sizes
    DBuf<int64_t> leaf_row0;       // nl + 1: first non-self row of each leaf
    DBuf<int64_t> ns_off;          // nl + 1: offsets of the non-self blocks (rows x n)
    DBuf<int64_t> row_body;        // rows: sorted target body of each non-self row
    DBuf<double> ns_M, old_infl;
    DBuf<double> strengths, rhs, extra, normals, leaf_sq;
    std::vector<int64_t> leaf_start_h;
    // sweep schedule: leaf k's level is one above the highest level of the leaves before it
    // that it exchanges influence with, so one level's leaves are solved concurrently and
    // every leaf still sees exactly the updates of the leaves before it
    std::vector<int64_t> level_ptr;   // host: levels + 1
    DBuf<int32_t> level_leaves;       // leaves grouped by level, in order within a level
    DBuf<int32_t> all_leaves;         // 0 .. nl-1
};

void free_fgs(FgsState* f) { delete f; }

static inline unsigned grid_for(int64_t n, int block) { return (unsigned)std::max<int64_t>(1, (n + block - 1) / block); }

template <class T, class O>
static void excl_scan(Context& C, const T* in, O* out, int64_t n) {
    excl_scan3<T, O>(C, in, out, n);
}

// ---------------------------------------------------------------- set-up
__global__ void k_fgs_leaf_sizes(const Branch* __restrict__ br, const int64_t* __restrict__ leaf_index, int64_t nl,
                                 int64_t* __restrict__ n, int64_t* __restrict__ n2) {
    const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (k > nl) return;
    const int64_t m = k < nl ? br[leaf_index[k]].n_bodies : 0;
    n[k] = m;
    n2[k] = m * m;
}

__global__ void k_fgs_lpos(const Branch* __restrict__ br, const int64_t* __restrict__ leaf_index, int64_t nl,
                           const int64_t* __restrict__ leaf_start, int64_t* __restrict__ lpos) {
    const Branch& b = br[leaf_index[blockIdx.x]];
    for (int64_t i = threadIdx.x; i < b.n_bodies; i += blockDim.x) lpos[b.first + i] = leaf_start[blockIdx.x] + i;
}

// sizes of the target leaves of the by-source list (rows of the non-self blocks); segment bounds per source leaf
__global__ void k_fgs_pairs(const uint32_t* __restrict__ s, const int32_t* __restrict__ t, int64_t np,
                            const Branch* __restrict__ tb, const Branch* __restrict__ sb, int64_t* __restrict__ rows,
                            int64_t* __restrict__ seg) {
    const int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (q > np) return;
    if (q == np) { rows[q] = 0; return; }
    rows[q] = tb[t[q]].n_bodies;
    const int64_t leaf = sb[s[q]].leaf;
    if (q == 0 || s[q] != s[q - 1]) seg[2 * leaf] = q;
    if (q == np - 1 || s[q] != s[q + 1]) seg[2 * leaf + 1] = q + 1;
}

__global__ void k_fgs_leaf_rows(const int64_t* __restrict__ seg, const int64_t* __restrict__ pair_rows,
                                const int64_t* __restrict__ leaf_n, int64_t nl, int64_t* __restrict__ rows,
                                int64_t* __restrict__ entries) {
    const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (k > nl) return;
    if (k == nl) { rows[k] = 0; entries[k] = 0; return; }
    const int64_t r = pair_rows[seg[2 * k + 1]] - pair_rows[seg[2 * k]];
    rows[k] = r;
    entries[k] = r * leaf_n[k];
}

__global__ void k_fgs_row_body(const int32_t* __restrict__ t, int64_t np, const Branch* __restrict__ tb,
                               const int64_t* __restrict__ pair_rows, int64_t* __restrict__ row_body) {
    const int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (q >= np) return;
    const Branch& b = tb[t[q]];
    for (int64_t i = 0; i < b.n_bodies; ++i) row_body[pair_rows[q] + i] = b.first + i;
}

// unit normal of every sorted panel from its first three vertices (get_normal, src/compatibility.jl:65-72)
__global__ void k_fgs_normals(const double* __restrict__ rows, int64_t n, double* __restrict__ nrm) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    double v[3][3];
    for (int k = 0; k < 3; ++k)
        for (int a = 0; a < 3; ++a) v[k][a] = rows[(int64_t)(1 + 3 * k + a) * n + i];
    const double a0 = v[1][0] - v[0][0], a1 = v[1][1] - v[0][1], a2 = v[1][2] - v[0][2];
    const double b0 = v[2][0] - v[0][0], b1 = v[2][1] - v[0][1], b2 = v[2][2] - v[0][2];
    double n0 = a1 * b2 - a2 * b1, n1 = a2 * b0 - a0 * b2, n2 = a0 * b1 - a1 * b0;
    const double nn = sqrt(n0 * n0 + n1 * n1 + n2 * n2);
    nrm[3 * i] = n0 / nn; nrm[3 * i + 1] = n1 / nn; nrm[3 * i + 2] = n2 / nn;
}

// Influence blocks. One workgroup per leaf: its source bodies (staged panels for
// panel kernels) against the rows of the block. SELF: the leaf's own bodies as
// rows (n x n); otherwise the non-self rows. Entry (row, col) is the influence at
// the row's body of the column's body at unit strength (set_unit_strength!,
// :62-70), column-major.
template <int NV, bool SELF>
__global__ __launch_bounds__(256) void k_fgs_blocks(const Branch* __restrict__ br, const int64_t* __restrict__ leaf_index,
                                                    const double* __restrict__ x, const double* __restrict__ y,
                                                    const double* __restrict__ z, const double* __restrict__ srows, int64_t ns,
                                                    const int64_t* __restrict__ off, const int64_t* __restrict__ leaf_row0,
                                                    const int64_t* __restrict__ row_body, const double* __restrict__ nrm,
                                                    int influence, double* __restrict__ out) {
    __shared__ PanelStage<(NV > 0 ? NV : 3)> tile[FGS_STAGE];
    const Branch& b = br[leaf_index[blockIdx.x]];
    const int64_t n = b.n_bodies;
    const int64_t rows = SELF ? n : leaf_row0[blockIdx.x + 1] - leaf_row0[blockIdx.x];
    if (rows == 0) return;
    double* M = out + off[blockIdx.x];
    // columns in chunks of FGS_STAGE (panels staged once per chunk); every entry is computed on its own
    for (int64_t c0 = 0; c0 < n; c0 += FGS_STAGE) {
        const int64_t nc = n - c0 < FGS_STAGE ? n - c0 : FGS_STAGE;
        if (NV > 0) {
            for (int64_t c = threadIdx.x; c < nc; c += blockDim.x) {
                stage_panel<(NV > 0 ? NV : 3)>(x, y, z, srows, ns, b.first + c0 + c, tile[c]);
                tile[c].strength = 1.0;
            }
            __syncthreads();
        }
        for (int64_t e = threadIdx.x; e < rows * nc; e += blockDim.x) {
            const int64_t r = e % rows, cl = e / rows, c = c0 + cl;
            const int64_t it = SELF ? b.first + r : row_body[leaf_row0[blockIdx.x] + r];
            const int64_t js = b.first + c;
            double v = 0.0;
            if (NV == 0) {   // point source at unit strength (test/gravitational.jl:80-104)
                const double dx = x[it] - x[js], dy = y[it] - y[js], dz = z[it] - z[js];
                const double r2 = dx * dx + dy * dy + dz * dz;
                if (r2 > 0) v = 1.0 / sqrt(r2) * FGS_INV4PI;
            } else {
                double a[13] = {0.0};
                if (influence == 0) panel_pair<(NV > 0 ? NV : 3), false, true, false, false>(tile[cl], x[it], y[it], z[it], a, 1.0);
                else panel_pair<(NV > 0 ? NV : 3), false, false, true, false>(tile[cl], x[it], y[it], z[it], a, 1.0);
                v = influence == 0 ? a[0] : a[1] * nrm[3 * it] + a[2] * nrm[3 * it + 1] + a[3] * nrm[3 * it + 2];
            }
            M[c * rows + r] = v;
        }
        if (NV > 0) __syncthreads();   // the next chunk overwrites the stage
    }
}

// inverse of every self block: Gauss-Jordan with partial pivoting on [A | I]. HBM = false: blocks of up to
// FGS_LDS_LEAF bodies, worked in LDS; HBM = true: the larger blocks, worked in device memory (work holds
// 2 n^2 doubles per block at twice the block's offset). Each launch skips the other's blocks; both run
// the same steps in the same order.
template <bool HBM>
__global__ __launch_bounds__(256) void k_fgs_invert(const int64_t* __restrict__ leaf_start, const int64_t* __restrict__ off,
                                                    const double* __restrict__ A, double* __restrict__ Ainv, double* work) {
    extern __shared__ double sh[];
    __shared__ int piv;
    const int n = (int)(leaf_start[blockIdx.x + 1] - leaf_start[blockIdx.x]);
    if (n == 0 || (n > FGS_LDS_LEAF) != HBM) return;
    double* a = HBM ? work + 2 * off[blockIdx.x] : sh;   // row-major n x n
    double* b = a + (int64_t)n * n;                      // row-major n x n
    const double* src = A + off[blockIdx.x];
    for (int e = threadIdx.x; e < n * n; e += blockDim.x) {
        const int r = e % n, c = e / n;   // column-major source
        a[r * n + c] = src[e];
        b[r * n + c] = (r == c) ? 1.0 : 0.0;
    }
    __syncthreads();
    for (int k = 0; k < n; ++k) {
        if (threadIdx.x < 64) {   // pivot: the largest |a(r, k)|, r >= k, first on ties
            double best = -1.0;
            int bi = k;
            for (int r = k + (int)threadIdx.x; r < n; r += 64) {
                const double v = fabs(a[r * n + k]);
                if (v > best) { best = v; bi = r; }
            }
            for (int w = 32; w > 0; w >>= 1) {
                const double ob = __shfl_xor(best, w, 64);
                const int oi = __shfl_xor(bi, w, 64);
                if (ob > best || (ob == best && oi < bi)) { best = ob; bi = oi; }
            }
            if (threadIdx.x == 0) piv = bi;
        }
        __syncthreads();
        const int p = piv;
        if (p != k)
            for (int c = threadIdx.x; c < n; c += blockDim.x) {
                double t = a[k * n + c]; a[k * n + c] = a[p * n + c]; a[p * n + c] = t;
                t = b[k * n + c]; b[k * n + c] = b[p * n + c]; b[p * n + c] = t;
            }
        __syncthreads();
        const double d = a[k * n + k];
        __syncthreads();
        for (int c = threadIdx.x; c < n; c += blockDim.x) { a[k * n + c] /= d; b[k * n + c] /= d; }
        __syncthreads();
        for (int e = threadIdx.x; e < n * n; e += blockDim.x) {
            const int r = e / n, c = e % n;
            if (r == k) continue;
            const double f = a[r * n + k];
            if (c != k) a[r * n + c] -= f * a[k * n + c];
            b[r * n + c] -= f * b[k * n + c];
        }
        __syncthreads();
        for (int r = threadIdx.x; r < n; r += blockDim.x)
            if (r != k) a[r * n + k] = 0.0;
        __syncthreads();
    }
    double* dst = Ainv + off[blockIdx.x];
    for (int e = threadIdx.x; e < n * n; e += blockDim.x) {
        const int r = e % n, c = e / n;
        dst[e] = b[r * n + c];
    }
}

// ---------------------------------------------------------------- solve
__global__ void k_fgs_from_bodies(const int64_t* __restrict__ lpos, const double* __restrict__ v, const int64_t* __restrict__ sort_index,
                                  int64_t n, double scale, int original_order, double* __restrict__ out) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < n) out[lpos[i]] = scale * v[original_order ? sort_index[i] : i];
}

__global__ void k_fgs_to_sorted(const int64_t* __restrict__ lpos, const double* __restrict__ s, int64_t n, double* __restrict__ dst) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < n) dst[i] = s[lpos[i]];
}

// extra = -influence of the farfield evaluation (influence!, :960-993), rhs += extra
__global__ void k_fgs_extra(const int64_t* __restrict__ lpos, const double* __restrict__ tout, int64_t n, int influence,
                            const double* __restrict__ nrm, double* __restrict__ extra, double* __restrict__ rhs) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double v = influence == 0 ? tout[i]
                                    : tout[n + i] * nrm[3 * i] + tout[2 * n + i] * nrm[3 * i + 1] + tout[3 * n + i] * nrm[3 * i + 2];
    extra[lpos[i]] = -v;
    rhs[lpos[i]] += -v;
}

__global__ void k_fgs_axpy(const double* __restrict__ x, int64_t n, double a, double* __restrict__ y) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < n) y[i] += a * x[i];
}

// squared residual of each self block, A s - rhs (residual!, :995-1021)
__global__ __launch_bounds__(256) void k_fgs_residual(const int64_t* __restrict__ leaf_start, const int64_t* __restrict__ off,
                                                      const double* __restrict__ A, const double* __restrict__ s,
                                                      const double* __restrict__ rhs, double* __restrict__ leaf_sq) {
    __shared__ double part[4];
    const int64_t s0 = leaf_start[blockIdx.x];
    const int n = (int)(leaf_start[blockIdx.x + 1] - s0);
    const double* M = A + off[blockIdx.x];
    double acc = 0.0;
    for (int r = threadIdx.x; r < n; r += blockDim.x) {
        double v = -rhs[s0 + r];
        for (int c = 0; c < n; ++c) v += M[c * n + r] * s[s0 + c];
        acc += v * v;
    }
    for (int w = 32; w > 0; w >>= 1) acc += __shfl_xor(acc, w, 64);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) leaf_sq[blockIdx.x] = part[0] + part[1] + part[2] + part[3];
}

// One level of a Gauss-Seidel sweep: one workgroup per leaf of the level. SOLVE:
// s_k = A_k^-1 rhs_k; then the leaf's non-self influence replaces its previous one
// in the right-hand side of its neighbours (update_nonself_influence!, :740-795).
// Leaves of one level never update each other's right-hand side; two of them may
// update the same neighbour, hence the atomic adds. SOLVE = 0 only moves every
// leaf's influence into the right-hand side (the update before the iterations,
// :872-874).
template <bool SOLVE>
__global__ __launch_bounds__(256) void k_fgs_level(const int32_t* __restrict__ leaves, const int64_t* __restrict__ leaf_start,
                                                   const int64_t* __restrict__ self_off, const double* __restrict__ Ainv,
                                                   const int64_t* __restrict__ leaf_row0, const int64_t* __restrict__ ns_off,
                                                   const double* __restrict__ M, const int64_t* __restrict__ row_body,
                                                   const int64_t* __restrict__ lpos, double* __restrict__ s, double* rhs,
                                                   double* __restrict__ old_infl, int cap) {
    extern __shared__ double xr[];   // 2 cap doubles (cap = the largest leaf): strengths, then the right-hand side
    double* xs = xr;
    double* rk = xr + cap;
    const int64_t k = leaves[blockIdx.x];
    const int64_t s0 = leaf_start[k];
    const int n = (int)(leaf_start[k + 1] - s0);
    if (SOLVE) {
        const double* B = Ainv + self_off[k];
        for (int c = threadIdx.x; c < n; c += blockDim.x) rk[c] = rhs[s0 + c];
        __syncthreads();
        for (int r = threadIdx.x; r < n; r += blockDim.x) {
            double v = 0.0;
#pragma unroll 8
            for (int c = 0; c < n; ++c) v += B[c * n + r] * rk[c];
            xs[r] = v;
            s[s0 + r] = v;
        }
    } else {
        for (int r = threadIdx.x; r < n; r += blockDim.x) xs[r] = s[s0 + r];
    }
    __syncthreads();
    const int64_t r0 = leaf_row0[k], rows = leaf_row0[k + 1] - r0;
    const double* Mk = M + ns_off[k];
    for (int64_t r = threadIdx.x; r < rows; r += blockDim.x) {
        double now = 0.0;
#pragma unroll 8
        for (int c = 0; c < n; ++c) now += Mk[c * rows + r] * xs[c];
        atomicAdd(rhs + lpos[row_body[r0 + r]], old_infl[r0 + r] - now);
        old_infl[r0 + r] = now;
    }
}

// ---------------------------------------------------------------- host
template <int NV, bool SELF>
static void launch_blocks(Context& C, const Tree& st, FgsState& F, const int64_t* off, double* out) {
    k_fgs_blocks<NV, SELF><<<(unsigned)F.nl, 256, 0, C.stream>>>(st.branches.data(), st.leaf_index.data(), st.x.data(), st.y.data(),
                                                                 st.z.data(), st.strength.data(), st.n, off, F.leaf_row0.data(),
                                                                 F.row_body.data(), F.normals.data(), F.influence, out);
    LAUNCH_CHECK();
}

template <bool SELF>
static void build_blocks(Context& C, int kernel, const Tree& st, FgsState& F, const int64_t* off, double* out) {
    switch (kernel) {
        case FMMHIP_POINT_SOURCE: launch_blocks<0, SELF>(C, st, F, off, out); break;
        case FMMHIP_PANEL_SOURCE_TRI: launch_blocks<3, SELF>(C, st, F, off, out); break;
        default: launch_blocks<4, SELF>(C, st, F, off, out); break;
    }
}

void fgs_setup(Eval& E, const fmmhip_options& o, int influence) {
    Context& C = *E.ctx;
    delete E.fgs;
    E.fgs = new FgsState();
    FgsState& F = *E.fgs;
    F.influence = influence;
    F.p = o.expansion_order;
    Tree& st = E.st;
    const int64_t ns = st.n;
    F.nl = st.n_leaves;
    const int64_t nl = F.nl;
    {   // the solver's rows and columns share one body order: the target and source trees must sort alike
        std::vector<int64_t> a(std::max<int64_t>(1, ns)), b(std::max<int64_t>(1, ns));
        if (ns) {
            HIP_CHECK(hipMemcpy(a.data(), E.tt.sort_index.data(), sizeof(int64_t) * ns, hipMemcpyDeviceToHost));
            HIP_CHECK(hipMemcpy(b.data(), st.sort_index.data(), sizeof(int64_t) * ns, hipMemcpyDeviceToHost));
        }
        if (E.tt.n != ns || a != b) throw Error(FMMHIP_EINVAL, "FastGaussSeidel needs identical target and source trees");
    }
    // full direct list (self pairs included) for the final nearfield
    build_lists(C, E.tt, st, F.full, o.leaf_size_source, o.multipole_acceptance, FMMHIP_BARBA, false, true, true, false);
    // leaf-order layout
    DBuf<int64_t> leaf_n, leaf_n2;
    leaf_n.resize(nl + 1); leaf_n2.resize(nl + 1);
    F.leaf_start.resize(nl + 1); F.self_off.resize(nl + 1); F.lpos.resize(std::max<int64_t>(1, ns));
    k_fgs_leaf_sizes<<<grid_for(nl + 1, 256), 256, 0, C.stream>>>(st.branches.data(), st.leaf_index.data(), nl, leaf_n.data(),
                                                                 leaf_n2.data());
    LAUNCH_CHECK();
    excl_scan(C, leaf_n.data(), F.leaf_start.data(), nl + 1);
    excl_scan(C, leaf_n2.data(), F.self_off.data(), nl + 1);
    F.leaf_start_h.resize(nl + 1);
    HIP_CHECK(hipMemcpyAsync(F.leaf_start_h.data(), F.leaf_start.data(), sizeof(int64_t) * (nl + 1), hipMemcpyDeviceToHost, C.stream));
    HIP_CHECK(hipStreamSynchronize(C.stream));
    int64_t max_n = 0;
    for (int64_t k = 0; k < nl; ++k) max_n = std::max(max_n, F.leaf_start_h[k + 1] - F.leaf_start_h[k]);
    if (max_n > FGS_MAX_LEAF) throw Error(FMMHIP_EINVAL, "FastGaussSeidel supports leaves of at most 4096 bodies");
    F.max_leaf = (int)max_n;
    {
        const int lsh = (int)(sizeof(double) * 2 * (size_t)std::max<int64_t>(1, max_n));
        HIP_CHECK(hipFuncSetAttribute((const void*)k_fgs_level<true>, hipFuncAttributeMaxDynamicSharedMemorySize, lsh));
        HIP_CHECK(hipFuncSetAttribute((const void*)k_fgs_level<false>, hipFuncAttributeMaxDynamicSharedMemorySize, lsh));
    }
    k_fgs_lpos<<<(unsigned)std::max<int64_t>(1, nl), 256, 0, C.stream>>>(st.branches.data(), st.leaf_index.data(), nl,
                                                                        F.leaf_start.data(), F.lpos.data());
    LAUNCH_CHECK();
    // sort_by_source of the discovery-order direct list (src/interaction_list.jl:215-245): stable radix sort
    const int64_t np = E.lists.n_direct;
    F.by_src_s.resize(std::max<int64_t>(1, np)); F.by_src_t.resize(std::max<int64_t>(1, np));
    if (np) {
        DBuf<uint32_t> keys;
        keys.resize(np);
        HIP_CHECK(hipMemcpyAsync(keys.data(), E.lists.dfs_s.data(), sizeof(uint32_t) * np, hipMemcpyDeviceToDevice, C.stream));
        radix_sort<uint32_t>(C, keys.data(), F.by_src_s.data(), (const uint32_t*)E.lists.dfs_t.data(),
                             (uint32_t*)F.by_src_t.data(), np, 32);
        HIP_CHECK(hipStreamSynchronize(C.stream));
    }
    F.seg.resize(2 * std::max<int64_t>(1, nl));
    HIP_CHECK(hipMemsetAsync(F.seg.data(), 0, sizeof(int64_t) * 2 * std::max<int64_t>(1, nl), C.stream));
    DBuf<int64_t> prow;
    prow.resize(np + 1);
    F.pair_rows.resize(np + 1);
    k_fgs_pairs<<<grid_for(np + 1, 256), 256, 0, C.stream>>>(F.by_src_s.data(), F.by_src_t.data(), np, E.tt.branches.data(),
                                                             st.branches.data(), prow.data(), F.seg.data());
    LAUNCH_CHECK();
    excl_scan(C, prow.data(), F.pair_rows.data(), np + 1);
    DBuf<int64_t> lrows, lent;
    lrows.resize(nl + 1); lent.resize(nl + 1);
    F.leaf_row0.resize(nl + 1); F.ns_off.resize(nl + 1);
    k_fgs_leaf_rows<<<grid_for(nl + 1, 256), 256, 0, C.stream>>>(F.seg.data(), F.pair_rows.data(), leaf_n.data(), nl, lrows.data(),
                                                                lent.data());
    LAUNCH_CHECK();
    excl_scan(C, lrows.data(), F.leaf_row0.data(), nl + 1);
    excl_scan(C, lent.data(), F.ns_off.data(), nl + 1);
    F.rows = read_back(C, F.leaf_row0.data() + nl);
    const int64_t n_ns = read_back(C, F.ns_off.data() + nl), n_self = read_back(C, F.self_off.data() + nl);
    F.row_body.resize(std::max<int64_t>(1, F.rows));
    if (np)
        k_fgs_row_body<<<grid_for(np, 256), 256, 0, C.stream>>>(F.by_src_t.data(), np, E.tt.branches.data(), F.pair_rows.data(),
                                                               F.row_body.data());
    // sweep schedule (host, once): neighbours from the direct list, in both directions
    {
        std::vector<uint32_t> ps(std::max<int64_t>(1, np));
        std::vector<int32_t> pt(std::max<int64_t>(1, np));
        std::vector<Branch> br(st.n_branches);
        if (np) {
            HIP_CHECK(hipMemcpy(ps.data(), F.by_src_s.data(), sizeof(uint32_t) * np, hipMemcpyDeviceToHost));
            HIP_CHECK(hipMemcpy(pt.data(), F.by_src_t.data(), sizeof(int32_t) * np, hipMemcpyDeviceToHost));
        }
        HIP_CHECK(hipMemcpy(br.data(), st.branches.data(), sizeof(Branch) * br.size(), hipMemcpyDeviceToHost));
        std::vector<std::vector<int32_t>> nb(nl);
        for (int64_t q = 0; q < np; ++q) {
            const int32_t a = (int32_t)br[ps[q]].leaf, b = (int32_t)br[pt[q]].leaf;   // trees are identical
            nb[a].push_back(b);
            nb[b].push_back(a);
        }
        std::vector<int64_t> level(nl, 0);
        int64_t nlev = 0;
        for (int64_t k = 0; k < nl; ++k) {
            int64_t l = 0;
            for (int32_t j : nb[k])
                if (j < k) l = std::max(l, level[j] + 1);
            level[k] = l;
            nlev = std::max(nlev, l + 1);
        }
        F.level_ptr.assign(nlev + 1, 0);
        for (int64_t k = 0; k < nl; ++k) F.level_ptr[level[k] + 1] += 1;
        for (int64_t l = 0; l < nlev; ++l) F.level_ptr[l + 1] += F.level_ptr[l];
        std::vector<int32_t> order(std::max<int64_t>(1, nl)), all(std::max<int64_t>(1, nl)), fill(F.level_ptr.begin(), F.level_ptr.end() - 1);
        for (int64_t k = 0; k < nl; ++k) { order[fill[level[k]]++] = (int32_t)k; all[k] = (int32_t)k; }
        F.level_leaves.resize(std::max<int64_t>(1, nl));
        F.all_leaves.resize(std::max<int64_t>(1, nl));
        HIP_CHECK(hipMemcpy(F.level_leaves.data(), order.data(), sizeof(int32_t) * order.size(), hipMemcpyHostToDevice));
        HIP_CHECK(hipMemcpy(F.all_leaves.data(), all.data(), sizeof(int32_t) * all.size(), hipMemcpyHostToDevice));
    }
    F.normals.resize(std::max<int64_t>(3, 3 * ns));
    if (influence == 1) {
        k_fgs_normals<<<grid_for(ns, 256), 256, 0, C.stream>>>(st.strength.data(), ns, F.normals.data());
        LAUNCH_CHECK();
    }
    // influence blocks and the inverses of the self blocks
    F.self_A.resize(std::max<int64_t>(1, n_self)); F.self_inv.resize(std::max<int64_t>(1, n_self));
    F.ns_M.resize(std::max<int64_t>(1, n_ns));
    if (nl) {
        build_blocks<true>(C, E.kernel, st, F, F.self_off.data(), F.self_A.data());
        build_blocks<false>(C, E.kernel, st, F, F.ns_off.data(), F.ns_M.data());
        const int64_t m = std::min<int64_t>(max_n, FGS_LDS_LEAF);
        const size_t sh = sizeof(double) * 2 * (size_t)m * (size_t)m;
        HIP_CHECK(hipFuncSetAttribute((const void*)k_fgs_invert<false>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh));
        k_fgs_invert<false><<<(unsigned)nl, 256, sh, C.stream>>>(F.leaf_start.data(), F.self_off.data(), F.self_A.data(),
                                                                 F.self_inv.data(), nullptr);
        LAUNCH_CHECK();
        if (max_n > FGS_LDS_LEAF) {   // larger blocks: the same elimination in device memory
            DBuf<double> work;
            work.resize(2 * (size_t)n_self);
            k_fgs_invert<true><<<(unsigned)nl, 256, 0, C.stream>>>(F.leaf_start.data(), F.self_off.data(), F.self_A.data(),
                                                                   F.self_inv.data(), work.data());
            LAUNCH_CHECK();
            HIP_CHECK(hipStreamSynchronize(C.stream));   // work is freed on return
        }
    }
    F.old_infl.resize(std::max<int64_t>(1, F.rows));
    F.strengths.resize(std::max<int64_t>(1, ns)); F.rhs.resize(std::max<int64_t>(1, ns)); F.extra.resize(std::max<int64_t>(1, ns));
    F.leaf_sq.resize(std::max<int64_t>(1, nl));
    HIP_CHECK(hipStreamSynchronize(C.stream));
}

// farfield-only evaluation on the current strengths (fmm! with an empty direct list, :839-850):
// potential and gradient into E.tout (sorted)
// farfield-only evaluation of the current strengths (fixed order: no read-back, no host
// synchronisation, zeroing by kernels, so the sequence can be captured in a graph)
static void fgs_farfield(Eval& E, FgsState& F) {
    Context& C = *E.ctx;
    const int p = F.p;
    fill_zero(C, E.tout.data(), E.tout.size());
    fill_zero(C, E.tt.expansions.data(), E.tt.expansions.size());
    fill_zero(C, E.st.expansions.data(), E.st.expansions.size());
    ensure_tables(C, p);
    run_p2m(C, E.kernel, E.st, p);
    run_m2m(C, E.st, p, false);
    run_m2l(C, E.tt, E.st, E.tt, E.lists, p, false, FMMHIP_ERR_NONE, false, 0.0, true, FMMHIP_FP64, nullptr, nullptr);
    run_l2l(C, E.tt, p, false);
    run_l2p(C, E.tt, E.tout.data(), E.nt, p, false, true, true, false);
}

void fgs_solve(Eval& E, const double* ext_dev, int max_iterations, double tolerance, double* strengths_host, int32_t* sweeps,
               double* mse_host) {
    Context& C = *E.ctx;
    FgsState& F = *E.fgs;
    Tree& st = E.st;
    const int64_t n = st.n, nl = F.nl;
    E.tout.resize((size_t)13 * std::max<int64_t>(1, E.nt));
    // right-hand side = -external influence, strengths from the source buffer (update_by_leaf!, :658-688)
    k_fgs_from_bodies<<<grid_for(n, 256), 256, 0, C.stream>>>(F.lpos.data(), ext_dev, st.sort_index.data(), n, -1.0, 1, F.rhs.data());
    k_fgs_from_bodies<<<grid_for(n, 256), 256, 0, C.stream>>>(F.lpos.data(), st.strength.data(), st.sort_index.data(), n, 1.0, 0,
                                                             F.strengths.data());
    LAUNCH_CHECK();
    HIP_CHECK(hipMemsetAsync(F.old_infl.data(), 0, F.old_infl.size() * sizeof(double), C.stream));
    const int cap = std::max(1, F.max_leaf);
    const size_t level_sh = sizeof(double) * 2 * (size_t)cap;
    if (nl)
        k_fgs_level<false><<<(unsigned)nl, 256, level_sh, C.stream>>>(F.all_leaves.data(), F.leaf_start.data(), F.self_off.data(),
                                                                     F.self_inv.data(), F.leaf_row0.data(), F.ns_off.data(),
                                                                     F.ns_M.data(), F.row_body.data(), F.lpos.data(),
                                                                     F.strengths.data(), F.rhs.data(), F.old_infl.data(), cap);
    LAUNCH_CHECK();
    std::vector<double> sq(std::max<int64_t>(1, nl));
    // one iteration: farfield of the current strengths, the right-hand side's farfield term, the residual
    auto part_a = [&]() {
        fgs_farfield(E, F);
        k_fgs_extra<<<grid_for(n, 256), 256, 0, C.stream>>>(F.lpos.data(), E.tout.data(), n, F.influence, F.normals.data(),
                                                           F.extra.data(), F.rhs.data());
        LAUNCH_CHECK();
        if (nl) {
            k_fgs_residual<<<(unsigned)nl, 256, 0, C.stream>>>(F.leaf_start.data(), F.self_off.data(), F.self_A.data(),
                                                               F.strengths.data(), F.rhs.data(), F.leaf_sq.data());
            LAUNCH_CHECK();
        }
    };
    // the sweep in the reference's leaf order, then the strengths back into the tree's order
    auto part_b = [&]() {
        for (size_t l = 0; l + 1 < F.level_ptr.size(); ++l) {
            const int64_t a = F.level_ptr[l], cnt = F.level_ptr[l + 1] - a;
            k_fgs_level<true><<<(unsigned)cnt, 256, level_sh, C.stream>>>(F.level_leaves.data() + a, F.leaf_start.data(),
                                                                          F.self_off.data(), F.self_inv.data(), F.leaf_row0.data(),
                                                                          F.ns_off.data(), F.ns_M.data(), F.row_body.data(),
                                                                          F.lpos.data(), F.strengths.data(), F.rhs.data(),
                                                                          F.old_infl.data(), cap);
        }
        LAUNCH_CHECK();
        k_fgs_to_sorted<<<grid_for(n, 256), 256, 0, C.stream>>>(F.lpos.data(), F.strengths.data(), n, st.strength.data());
        k_fgs_axpy<<<grid_for(n, 256), 256, 0, C.stream>>>(F.extra.data(), n, -1.0, F.rhs.data());
        LAUNCH_CHECK();
    };
    auto residual = [&]() {
        double mse = 0.0;
        if (nl) {
            HIP_CHECK(hipMemcpyAsync(sq.data(), F.leaf_sq.data(), sizeof(double) * nl, hipMemcpyDeviceToHost, C.stream));
            HIP_CHECK(hipStreamSynchronize(C.stream));
            for (int64_t k = 0; k < nl; ++k) mse += sq[k];
        }
        return mse / (double)std::max<int64_t>(1, n);
    };
    // Iterations after the first replay one captured graph (sweep, then the next farfield and residual):
    // its buffers are all sized by the first, eager iteration, and nothing in it synchronises with the
    // host. FMMHIP_GRAPH=0 launches every kernel directly instead.
    const char* genv = std::getenv("FMMHIP_GRAPH");
    bool use_graph = C.graphs && !(genv && genv[0] == '0');
    hipGraphExec_t gexec = nullptr;
    // captures part_b + part_a; any failure ends the capture (the stream must not stay in capture mode)
    // and the solve continues with direct launches
    auto capture = [&]() -> bool {
        if (hipStreamBeginCapture(C.stream, hipStreamCaptureModeRelaxed) != hipSuccess) return false;
        bool ok = true;
        try {
            part_b();
            part_a();
        } catch (const Error&) {
            ok = false;
        }
        hipGraph_t g = nullptr;
        if (hipStreamEndCapture(C.stream, &g) != hipSuccess) ok = false;
        if (ok && hipGraphInstantiate(&gexec, g, nullptr, nullptr, 0) != hipSuccess) {
            gexec = nullptr;
            ok = false;
        }
        if (g) (void)hipGraphDestroy(g);
        (void)hipGetLastError();
        return ok;
    };
    int done = 0;
    if (max_iterations >= 1) part_a();
    for (int it = 1; it <= max_iterations; ++it) {
        const double mse = residual();
        if (mse_host) mse_host[it - 1] = mse;
        if (mse <= tolerance) break;
        ++done;
        if (it == max_iterations) {
            part_b();
            break;
        }
        if (!use_graph) {
            part_b();
            part_a();
            continue;
        }
        if (!gexec && !capture()) {
            use_graph = false;   // nothing of the failed capture ran: this iteration's work directly
            part_b();
            part_a();
            continue;
        }
        HIP_CHECK(hipGraphLaunch(gexec, C.stream));
    }
    if (gexec) HIP_CHECK(hipGraphExecDestroy(gexec));
    if (max_iterations < 1) HIP_CHECK(hipMemsetAsync(E.tout.data(), 0, E.tout.size() * sizeof(double), C.stream));
    // final nearfield with the self pairs, on top of the last farfield (:917-931)
    run_p2p(C, E.kernel, E.tt, st, F.full, E.tout.data(), E.nt, true, true, false, FMMHIP_FP64);
    E.tout_aos.resize((size_t)16 * std::max<int64_t>(1, E.nt));
    gather_outputs(C, E.tt, E.tout.data(), E.tout_aos.data(), E.nt);
    E.ran = true;
    if (strengths_host) {
        DBuf<double> tmp;
        tmp.resize(std::max<int64_t>(1, n));
        k_fgs_from_bodies<<<grid_for(n, 256), 256, 0, C.stream>>>(st.sort_index.data(), st.strength.data(), nullptr, n, 1.0, 0,
                                                                 tmp.data());
        LAUNCH_CHECK();
        HIP_CHECK(hipMemcpyAsync(strengths_host, tmp.data(), sizeof(double) * n, hipMemcpyDeviceToHost, C.stream));
    }
    HIP_CHECK(hipStreamSynchronize(C.stream));
    if (sweeps) *sweeps = done;
}

// ---------------------------------------------------------------- inspection
int64_t fgs_n_leaves(const Eval& E) { return E.fgs ? E.fgs->nl : 0; }

void fgs_leaf_info(const Eval& E, int64_t* out) {
    const FgsState& F = *E.fgs;
    Context& C = *E.ctx;
    const int64_t nl = F.nl;
    std::vector<int64_t> r0(nl + 1), li(std::max<int64_t>(1, nl));
    std::vector<Branch> br(E.st.n_branches);
    HIP_CHECK(hipMemcpy(r0.data(), F.leaf_row0.data(), sizeof(int64_t) * (nl + 1), hipMemcpyDeviceToHost));
    if (nl) HIP_CHECK(hipMemcpy(li.data(), E.st.leaf_index.data(), sizeof(int64_t) * nl, hipMemcpyDeviceToHost));
    HIP_CHECK(hipMemcpy(br.data(), E.st.branches.data(), sizeof(Branch) * br.size(), hipMemcpyDeviceToHost));
    (void)C;
    for (int64_t k = 0; k < nl; ++k) {
        out[4 * k] = li[k];
        out[4 * k + 1] = F.leaf_start_h[k + 1] - F.leaf_start_h[k];
        out[4 * k + 2] = r0[k + 1] - r0[k];
        out[4 * k + 3] = br[li[k]].first;
    }
}

void fgs_matrix(const Eval& E, int which, int64_t k, double* out) {
    const FgsState& F = *E.fgs;
    if (k < 0 || k >= F.nl) throw Error(FMMHIP_EINVAL, "leaf out of range");
    int64_t off[2], r0[2];
    const int64_t n = F.leaf_start_h[k + 1] - F.leaf_start_h[k];
    if (which == 0) {
        HIP_CHECK(hipMemcpy(off, F.self_off.data() + k, sizeof(off), hipMemcpyDeviceToHost));
        HIP_CHECK(hipMemcpy(out, F.self_A.data() + off[0], sizeof(double) * n * n, hipMemcpyDeviceToHost));
    } else {
        HIP_CHECK(hipMemcpy(off, F.ns_off.data() + k, sizeof(off), hipMemcpyDeviceToHost));
        HIP_CHECK(hipMemcpy(r0, F.leaf_row0.data() + k, sizeof(r0), hipMemcpyDeviceToHost));
        if (r0[1] > r0[0]) HIP_CHECK(hipMemcpy(out, F.ns_M.data() + off[0], sizeof(double) * (r0[1] - r0[0]) * n, hipMemcpyDeviceToHost));
    }
}

void fgs_nonself_rows(const Eval& E, int64_t k, int64_t* out) {
    const FgsState& F = *E.fgs;
    if (k < 0 || k >= F.nl) throw Error(FMMHIP_EINVAL, "leaf out of range");
    int64_t r0[2];
    HIP_CHECK(hipMemcpy(r0, F.leaf_row0.data() + k, sizeof(r0), hipMemcpyDeviceToHost));
    const int64_t rows = r0[1] - r0[0];
    if (rows == 0) return;
    std::vector<int64_t> body(rows), sidx(E.tt.n);
    HIP_CHECK(hipMemcpy(body.data(), F.row_body.data() + r0[0], sizeof(int64_t) * rows, hipMemcpyDeviceToHost));
    HIP_CHECK(hipMemcpy(sidx.data(), E.tt.sort_index.data(), sizeof(int64_t) * E.tt.n, hipMemcpyDeviceToHost));
    for (int64_t r = 0; r < rows; ++r) out[r] = sidx[body[r]];
}

}  // namespace fmmhip
