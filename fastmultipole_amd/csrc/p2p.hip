// Nearfield P2P on the GPU for the built-in point kernels:
//   Point{Source}: Laplace 1/r potential + gradient  (test/gravitational.jl:80-104)
//   Point{Vortex}: singular vortex velocity + velocity gradient (test/vortex.jl:83-128)
//   Point{Dipole}: phi = q.d/(4 pi r^3), the potential of the reference's dipole
//                  expansion (src/bodytomultipole.jl:431-463), with gradient and
//                  Hessian slots in the expansion evaluation's convention
//   Gaussian-core vortex: the singular vortex kernel times the core factor
//                  g(r/sigma) = erf(rho/sqrt 2) - sqrt(2/pi) rho exp(-rho^2/2)
// driven by the direct list sorted by target (src/fmm.jl:18-40).
//
// A workgroup evaluates one target leaf at a time, taking leaves from a queue
// (persistent launch of as many workgroups as the device holds). Targets live in
// registers, two per thread, with a group of t lanes per target pair when the
// leaf is small, each lane taking a strided share of the sources; the leaf's
// source leaves form one stream staged through LDS in 256-body tiles and
// broadcast to the whole workgroup. The partial sums of a group are folded with
// wavefront shuffles.
//
// FP64 form: IEEE double arithmetic with 1/r from a refined hardware rsqrt.
// Mixed form: every pair in FP32 with positions stored relative to the target
// leaf centre as a hi+lo pair of floats (differences keep ~2^-48 relative
// accuracy), each thread's per-tile FP32 partial sum added to its running total
// with Kahan compensation.
// Strengths are pre-scaled by 1/(4 pi).
#include "fmmhip_internal.hpp"
#include <map>
#include <mutex>
#include <cstdlib>
#include <algorithm>

namespace fmmhip {

static constexpr int P2P_BLOCK = 256;
static constexpr double INV4PI = 1.0 / (4.0 * 3.141592653589793);

__device__ inline double rinv_f64(double r2) {
    const double y = __builtin_amdgcn_rsq(r2);      // ~23-bit estimate
    // one third-order step: y (1 + e/2 + 3e^2/8), e = 1 - r2 y^2 (|e| ~ 2^-23 -> 2^-69)
    const double e = fma(-r2 * y, y, 1.0);
    return fma(y * e, fma(e, 0.375, 0.5), y);
}

// Kahan-compensated add of a partial sum into (s, c)
__device__ inline void kahan_add(float& s, float& c, float x) {
    float y = x - c;
    float t = s + y;
    c = (t - s) - y;
    s = t;
}

// sources per LDS slot: strengths (pre-scaled by 1/(4 pi)) and, for the Gaussian
// core, sigma
template <int K> struct KernelTraits {
    static constexpr int NS = K == FMMHIP_POINT_SOURCE ? 1 : (K == FMMHIP_VORTEX_GAUSSIAN || K == FMMHIP_POINT_SOURCE_VORTEX ? 4 : 3);
    static constexpr bool HESS = K != FMMHIP_POINT_SOURCE;   // the reference's source kernel has no Hessian term
    static constexpr bool POT = K == FMMHIP_POINT_SOURCE || K == FMMHIP_POINT_DIPOLE || K == FMMHIP_POINT_SOURCE_VORTEX;
    // outputs [phi, gradient (3), Hessian (9)]; in FP64 the point vortex keeps three
    // more sums, sum_j Gamma_j / r_j^3, whose antisymmetric part of the Hessian is
    // added once per target (fold_acc) instead of once per pair. FP32 sums keep the
    // per-pair form: the deferred sums cancel against the rest of the Hessian.
    static constexpr int NOUT = K == FMMHIP_POINT_SOURCE ? 4 : 13;
    template <class R> static constexpr bool defer() { return K == FMMHIP_POINT_VORTEX && sizeof(R) == 8; }
    template <class R> static constexpr int na() { return defer<R>() ? 16 : NOUT; }
    // accumulators a launch with these outputs writes (the others stay zero and are not folded)
    template <class R, bool PS, bool GS, bool HS> static constexpr bool used(int k) {
        return k == 0 ? (PS && POT) : k < 4 ? GS : k < 13 ? (HS && HESS) : (HS && defer<R>());
    }
};

template <class R> __device__ inline R erf_(R x);
template <> __device__ inline double erf_<double>(double x) { return erf(x); }
template <> __device__ inline float erf_<float>(float x) { return erff(x); }
template <class R> __device__ inline R exp_(R x);
template <> __device__ inline double exp_<double>(double x) { return exp(x); }
template <> __device__ inline float exp_<float>(float x) { return expf(x); }

// one body pair: d = x_target - x_source, ri = 1/|d|, g = staged source values;
// accumulates [phi, gradient slot (3), Hessian slot (9)] into a
template <int K, bool PS, bool GS, bool HS, class R>
__device__ inline void pair_eval(R dx, R dy, R dz, R r2, R ri, const R* g, R* a) {
    const R ri2 = ri * ri;
    if (K == FMMHIP_POINT_SOURCE) {
        const R q = g[0];
        if (PS) a[0] += q * ri;
        if (GS) { const R f = q * ri * ri2; a[1] += dx * f; a[2] += dy * f; a[3] += dz * f; }
    } else if (K == FMMHIP_POINT_SOURCE_VORTEX) {
        // the point source's terms (strength g[0], with the gradient's derivative), then the singular vortex's (g[1..3])
        const R q = g[0], gx = g[1], gy = g[2], gz = g[3];
        const R d3 = ri * ri2;
        if (PS) a[0] += q * ri;
        const R cx = dz * gy - dy * gz, cy = dx * gz - dz * gx, cz = dy * gx - dx * gy;
        if (GS) {
            const R f = q * d3;
            a[1] += dx * f; a[2] += dy * f; a[3] += dz * f;
            a[1] += cx * d3; a[2] += cy * d3; a[3] += cz * d3;
        }
        if (HS) {
            const R d5 = d3 * ri2, t3 = R(-3) * d5;
            // the source's gradient derivative q (delta_ij / r^3 - 3 d_i d_j / r^5)
            const R qd = q * d3, qx = q * t3 * dx, qy = q * t3 * dy, qz = q * t3 * dz;
            a[4] += qx * dx + qd; a[5] += qx * dy;      a[6] += qx * dz;
            a[7] += qy * dx;      a[8] += qy * dy + qd; a[9] += qy * dz;
            a[10] += qz * dx;     a[11] += qz * dy;     a[12] += qz * dz + qd;
            const R ax = t3 * dx, ay = t3 * dy, az = t3 * dz;
            const R g3x = gx * d3, g3y = gy * d3, g3z = gz * d3;
            a[4] += ax * cx;       a[5] += ax * cy + g3z; a[6] += ax * cz - g3y;
            a[7] += ay * cx - g3z; a[8] += ay * cy;       a[9] += ay * cz + g3x;
            a[10] += az * cx + g3y; a[11] += az * cy - g3x; a[12] += az * cz;
        }
    } else if (K == FMMHIP_POINT_DIPOLE) {
        const R qx = g[0], qy = g[1], qz = g[2];
        const R qd = qx * dx + qy * dy + qz * dz;
        const R ri3 = ri * ri2, ri5 = ri3 * ri2;
        if (PS) a[0] += qd * ri3;
        const R t = R(3) * qd * ri5;
        if (GS) { a[1] += t * dx - qx * ri3; a[2] += t * dy - qy * ri3; a[3] += t * dz - qz * ri3; }
        if (HS) {
            const R u = R(-15) * qd * ri5 * ri2, f3 = R(3) * ri5;
            const R d[3] = {dx, dy, dz}, q[3] = {qx, qy, qz};
#pragma unroll
            for (int i = 0; i < 3; ++i)
#pragma unroll
                for (int j = 0; j < 3; ++j)
                    a[4 + 3 * i + j] += f3 * (q[j] * d[i] + q[i] * d[j]) + u * d[i] * d[j] + (i == j ? t : R(0));
        }
    } else {
        const R gx = g[0], gy = g[1], gz = g[2];
        R gf = R(1), gp = R(0), sri = R(0);
        if (K == FMMHIP_VORTEX_GAUSSIAN) {
            const R sig = g[3];
            if (sig > R(0)) {
                const R r = r2 * ri, rho = r / sig;
                if (rho < (sizeof(R) == 8 ? R(9.5) : R(7))) {
                    const R c2pi = R(0.79788456080286535588), e = exp_<R>(R(-0.5) * rho * rho);
                    gf = erf_<R>(rho * R(0.70710678118654752440)) - c2pi * rho * e;
                    gp = c2pi * rho * rho * e;
                    sri = gp / (sig * r);
                }
            }
        }
        const R d3 = ri * ri2;
        const R cx = dz * gy - dy * gz, cy = dx * gz - dz * gx, cz = dy * gx - dx * gy;
        if (GS) { a[1] += gf * cx * d3; a[2] += gf * cy * d3; a[3] += gf * cz * d3; }
        if (HS) {
            const R d5 = d3 * ri2;
            if (KernelTraits<K>::template defer<R>()) {
                // sum_j d5 d_i c_j; the factor -3 is applied once per target (fold_acc)
                const R ax = d5 * dx, ay = d5 * dy, az = d5 * dz;
                a[4] += ax * cx; a[5] += ax * cy; a[6] += ax * cz;
                a[7] += ay * cx; a[8] += ay * cy; a[9] += ay * cz;
                a[10] += az * cx; a[11] += az * cy; a[12] += az * cz;
                a[13] += gx * d3; a[14] += gy * d3; a[15] += gz * d3;
            } else {
                const R t3 = R(-3) * d5;
                const R ax = t3 * dx, ay = t3 * dy, az = t3 * dz;
                const R g3x = gx * d3, g3y = gy * d3, g3z = gz * d3;
                R h[9] = {ax * cx,       ax * cy + g3z, ax * cz - g3y,
                          ay * cx - g3z, ay * cy,       ay * cz + g3x,
                          az * cx + g3y, az * cy - g3x, az * cz};
                if (K == FMMHIP_VORTEX_GAUSSIAN) {
                    const R c[3] = {cx * d3, cy * d3, cz * d3}, d[3] = {dx, dy, dz};
#pragma unroll
                    for (int i = 0; i < 3; ++i)
#pragma unroll
                        for (int j = 0; j < 3; ++j) h[3 * i + j] = gf * h[3 * i + j] + sri * d[i] * c[j];
                }
#pragma unroll
                for (int k = 0; k < 9; ++k) a[4 + k] += h[k];
            }
        }
    }
}

// completes the deferred form of the point vortex: the symmetric sums were kept
// without their factor -3, and S = sum_j Gamma_j / r_j^3 is added to the Hessian
// slots as the per-pair terms H(0,1) += S_z, H(0,2) -= S_y, H(1,0) -= S_z,
// H(1,2) += S_x, H(2,0) += S_y, H(2,1) -= S_x would have
template <int K, bool HS, class R>
__device__ inline void fold_acc(R* a) {
    if constexpr (KernelTraits<K>::template defer<R>() && HS) {
#pragma unroll
        for (int k = 4; k < 13; ++k) a[k] *= R(-3);
        a[5] += a[15]; a[6] -= a[14]; a[7] -= a[15];
        a[9] += a[13]; a[10] += a[14]; a[11] -= a[13];
    }
}

// stages source j of the sorted arrays (strength SoA [dim][ns], radius) into slot values
// g[0], g[stride], ...
template <int K, class R>
__device__ inline void stage_source(const double* __restrict__ sq, const double* __restrict__ srad, int64_t ns, int64_t j, R* g,
                                    int stride = 1) {
    g[0] = (R)(sq[j] * INV4PI);
    if (KernelTraits<K>::NS >= 3) { g[stride] = (R)(sq[ns + j] * INV4PI); g[2 * stride] = (R)(sq[2 * ns + j] * INV4PI); }
    if (K == FMMHIP_VORTEX_GAUSSIAN) g[3 * stride] = (R)srad[j];
    if (K == FMMHIP_POINT_SOURCE_VORTEX) g[3 * stride] = (R)(sq[3 * ns + j] * INV4PI);
}

// Source body at position f of the flattened stream of a target leaf's source
// leaves (direct-list entries l0..l1-1, dir_cum = exclusive prefix of their sizes):
// tiles span source leaves, so small leaves do not cost a barrier each.
__device__ inline int64_t stream_entry(const int64_t* __restrict__ dir_cum, int64_t l0, int64_t l1, int64_t f) {
    int64_t lo = l0, hi = l1 - 1;
    while (lo < hi) {
        const int64_t mid = (lo + hi + 1) >> 1;
        if (dir_cum[mid] <= f) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

__device__ inline int64_t stream_body(const int32_t* __restrict__ dir_s, const int64_t* __restrict__ dir_cum,
                                      const Branch* __restrict__ sbr, int64_t l0, int64_t l1, int64_t f) {
    const int64_t li = stream_entry(dir_cum, l0, l1, f);
    return sbr[dir_s[li]].first + (f - dir_cum[li]);
}

// Lanes of a workgroup over a target leaf's work units (target pairs, or targets): each unit is
// taken by a group of t adjacent lanes of one wavefront, 64/t groups per wavefront (the lanes
// left over idle), with t the largest group size that covers the leaf in one pass. Any t, not only
// powers of two, so leaves of every size keep most lanes busy; a group's partial sums are folded
// by shuffle-down steps that stay inside the group.
struct LeafLanes {
    int t;         // lanes per unit
    int gpw;       // units per wavefront
    int per_pass;  // units per workgroup pass
    int gl;        // this lane's unit within its wavefront (>= gpw: idle lane)
    int lane_k;    // this lane's rank within its unit's group
};
__device__ inline LeafLanes leaf_lanes(int64_t units) {
    constexpr int NW = P2P_BLOCK / 64;
    LeafLanes L;
    L.t = 1;
    for (int c = 64; c > 1; --c)
        if ((int64_t)NW * (64 / c) >= units) {
            L.t = c;
            break;
        }
    L.gpw = 64 / L.t;
    L.per_pass = NW * L.gpw;
    const int ln = (int)(threadIdx.x & 63);
    L.gl = ln / L.t;
    L.lane_k = ln % L.t;
    return L;
}
// this lane's unit in the pass starting at c0
__device__ inline int64_t leaf_unit(const LeafLanes& L, int64_t c0) {
    return c0 + (int64_t)(threadIdx.x >> 6) * L.gpw + L.gl;
}
// sum of a group's values, left in its first lane
__device__ inline double group_fold(double v, const LeafLanes& L) {
    for (int w = 1; w < L.t; w <<= 1) {
        const double o = __shfl_down(v, w, 64);
        if (L.lane_k + w < L.t) v += o;
    }
    return v;
}

// a 64-bit value that is the same in every lane (read from LDS), as a wave-uniform value
__device__ inline int64_t uniform_i64(int64_t v) {
    return (int64_t)(((uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)(v >> 32)) << 32) |
                     (uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)v));
}

// ---------------------------------------------------------------- FP64 form
// 1/r of a pair for the FP64 kernels, 0 for a coincident pair (r = 0 contributes nothing).
// Every term of the vortex and dipole kernels and of the source kernel's gradient carries at
// least 1/r^2, so there clearing the high word is enough: what is left is at most a denormal
// whose square is 0, one select instead of two. The source potential (q/r) clears both words.
template <bool FULL>
__device__ inline double rinv_guarded(double r2) {
    const double ri = rinv_f64(r2);
    if (FULL) return r2 > 0.0 ? ri : 0.0;
    const uint64_t bits = (uint64_t)__double_as_longlong(ri);
    const uint32_t hi = r2 > 0.0 ? (uint32_t)(bits >> 32) : 0u;
    return __longlong_as_double((long long)(((uint64_t)hi << 32) | (bits & 0xffffffffull)));
}

// One source of an LDS tile against a lane's two targets. A tile holds rows [x, y, z, g...]
// of P2P_BLOCK doubles, so the reads of a source are one address with constant offsets.
template <int KERNEL, bool PS, bool GS, bool HS>
__device__ inline void p2p_source_f64(const double* __restrict__ q, double xa, double ya, double za, double xb, double yb,
                                      double zb, double* a, double* b) {
    constexpr int NS = KernelTraits<KERNEL>::NS;
    const double sxs = q[0], sys = q[P2P_BLOCK], szs = q[2 * P2P_BLOCK];
    double gs[NS];
#pragma unroll
    for (int k = 0; k < NS; ++k) gs[k] = q[(3 + k) * P2P_BLOCK];
    const double dxa = xa - sxs, dya = ya - sys, dza = za - szs;
    const double r2a = dxa * dxa + dya * dya + dza * dza;
    const double dxb = xb - sxs, dyb = yb - sys, dzb = zb - szs;
    const double r2b = dxb * dxb + dyb * dyb + dzb * dzb;
    if (KERNEL == FMMHIP_VORTEX_GAUSSIAN) {
        if (r2a > 0.0) pair_eval<KERNEL, PS, GS, HS, double>(dxa, dya, dza, r2a, rinv_f64(r2a), gs, a);
        if (r2b > 0.0) pair_eval<KERNEL, PS, GS, HS, double>(dxb, dyb, dzb, r2b, rinv_f64(r2b), gs, b);
    } else {
        constexpr bool FULL = KERNEL == FMMHIP_POINT_SOURCE && PS;
        pair_eval<KERNEL, PS, GS, HS, double>(dxa, dya, dza, r2a, rinv_guarded<FULL>(r2a), gs, a);
        pair_eval<KERNEL, PS, GS, HS, double>(dxb, dyb, dzb, r2b, rinv_guarded<FULL>(r2b), gs, b);
    }
}

// Each thread evaluates two adjacent targets of the leaf against its share of
// the staged sources, so a source read from LDS serves two pair evaluations.
template <int KERNEL, bool PS, bool GS, bool HS>
__global__ __launch_bounds__(P2P_BLOCK) void k_p2p_f64(const int32_t* __restrict__ dir_targets,
                                                       const int64_t* __restrict__ dir_offsets,
                                                       const int32_t* __restrict__ dir_s, const int64_t* __restrict__ dir_cum,
                                                       const Branch* __restrict__ tbr,
                                                       const Branch* __restrict__ sbr, const double* __restrict__ tx,
                                                       const double* __restrict__ ty, const double* __restrict__ tz,
                                                       const double* __restrict__ sx, const double* __restrict__ sy,
                                                       const double* __restrict__ sz, const double* __restrict__ sq,
                                                       const double* __restrict__ srad, int64_t ns, double* __restrict__ tout,
                                                       int64_t nt, int64_t blk0,
                                                       unsigned long long* __restrict__ queue, int64_t n_queue) {
    constexpr int NS = KernelTraits<KERNEL>::NS;
    constexpr int NA = KernelTraits<KERNEL>::template na<double>(), NOUT = KernelTraits<KERNEL>::NOUT;
    __shared__ double tile[3 + NS][P2P_BLOCK];   // rows x, y, z, staged source values
    // one target leaf per workgroup; with a queue (persistent launch) each workgroup takes leaves
    // from it until none are left
    __shared__ int64_t s_next;
    const int tid = threadIdx.x;
    int64_t bi = blk0 + blockIdx.x;
    for (;;) {
        if (queue) {
            __syncthreads();   // the previous leaf is done with the LDS tiles and s_next
            if (tid == 0) s_next = (int64_t)atomicAdd(queue, 1ull);
            __syncthreads();
            const int64_t next = uniform_i64(s_next);
            if (next >= n_queue) break;
            bi = blk0 + next;
        }
        const Branch T = tbr[dir_targets[bi]];
        const int64_t l0 = dir_offsets[bi], l1 = dir_offsets[bi + 1];
        const int64_t base = dir_cum[l0], end = dir_cum[l1];   // the target leaf's source bodies as one stream
        const int64_t ngroups = (T.n_bodies + 1) / 2;           // target pairs
        const LeafLanes LL = leaf_lanes(ngroups);
        const int tpt = LL.t, lane_k = LL.lane_k;
        for (int64_t c0 = 0; c0 < ngroups; c0 += LL.per_pass) {
            const int64_t g = leaf_unit(LL, c0);
            const bool own = LL.gl < LL.gpw;
            const bool act0 = own && 2 * g < T.n_bodies, act1 = own && 2 * g + 1 < T.n_bodies;
            const int64_t i0 = T.first + (act0 ? 2 * g : 0), i1 = act1 ? i0 + 1 : i0;
            const double xa = tx[i0], ya = ty[i0], za = tz[i0];
            const double xb = tx[i1], yb = ty[i1], zb = tz[i1];
            double a[NA], b[NA];
    #pragma unroll
            for (int k = 0; k < NA; ++k) { a[k] = 0.0; b[k] = 0.0; }
            for (int64_t f0 = base; f0 < end; f0 += P2P_BLOCK) {
                const int cnt = (int)std::min<int64_t>(P2P_BLOCK, end - f0);
                __syncthreads();
                if (tid < cnt) {
                    const int64_t j = stream_body(dir_s, dir_cum, sbr, l0, l1, f0 + tid);
                    tile[0][tid] = sx[j]; tile[1][tid] = sy[j]; tile[2][tid] = sz[j];
                    stage_source<KERNEL, double>(sq, srad, ns, j, &tile[3][tid], P2P_BLOCK);
                }
                __syncthreads();
                if (act0) {
                    // the lane takes sources lane_k, lane_k + t, ... (t lanes per target pair) in
                    // uniform rounds, so only the LDS address is per-lane loop work
                    const int rounds = cnt / tpt;
                    const double* p = &tile[0][lane_k];
                    for (int r = 0; r < rounds; ++r, p += tpt) p2p_source_f64<KERNEL, PS, GS, HS>(p, xa, ya, za, xb, yb, zb, a, b);
                    if (lane_k < cnt - rounds * tpt) p2p_source_f64<KERNEL, PS, GS, HS>(p, xa, ya, za, xb, yb, zb, a, b);
                }
            }
            // fold the partial sums of a target pair's lanes
    #pragma unroll
            for (int k = 0; k < NA; ++k)
                if (KernelTraits<KERNEL>::template used<double, PS, GS, HS>(k)) { a[k] = group_fold(a[k], LL); b[k] = group_fold(b[k], LL); }
            fold_acc<KERNEL, HS, double>(a);
            fold_acc<KERNEL, HS, double>(b);
            if (lane_k == 0) {
    #pragma unroll
                for (int h = 0; h < 2; ++h) {
                    if (!(h ? act1 : act0)) continue;
                    const int64_t it = h ? i1 : i0;
                    const double* r = h ? b : a;
                    if (PS && KernelTraits<KERNEL>::POT) tout[it] += r[0];
                    if (GS) { tout[nt + it] += r[1]; tout[2 * nt + it] += r[2]; tout[3 * nt + it] += r[3]; }
                    if (HS && KernelTraits<KERNEL>::HESS)
                        for (int k = 4; k < NOUT; ++k) tout[(int64_t)k * nt + it] += r[k];
                }
            }
        }
        if (!queue) break;
    }
}

// --------------------------------------------------------------- mixed form
// One source of an LDS tile (rows hi x, y, z, lo x, y, z, staged values) against a lane's two
// targets (hi th, lo tl relative to the leaf centre); the pair's partial sums go to pa, pb.
template <int KERNEL, bool PS, bool GS, bool HS>
__device__ inline void p2p_source_f32(const float* __restrict__ q, const float (&th)[2][3], const float (&tl)[2][3], float* pa,
                                      float* pb) {
    constexpr int NS = KernelTraits<KERNEL>::NS;
    const float shx = q[0], shy = q[P2P_BLOCK], shz = q[2 * P2P_BLOCK];
    const float slx = q[3 * P2P_BLOCK], sly = q[4 * P2P_BLOCK], slz = q[5 * P2P_BLOCK];
    float gs[NS];
#pragma unroll
    for (int k = 0; k < NS; ++k) gs[k] = q[(6 + k) * P2P_BLOCK];
    const float dxa = (th[0][0] - shx) + (tl[0][0] - slx), dya = (th[0][1] - shy) + (tl[0][1] - sly),
                dza = (th[0][2] - shz) + (tl[0][2] - slz);
    const float dxb = (th[1][0] - shx) + (tl[1][0] - slx), dyb = (th[1][1] - shy) + (tl[1][1] - sly),
                dzb = (th[1][2] - shz) + (tl[1][2] - slz);
    const float r2a = dxa * dxa + dya * dya + dza * dza, r2b = dxb * dxb + dyb * dyb + dzb * dzb;
    if (KERNEL == FMMHIP_VORTEX_GAUSSIAN) {
        if (r2a > 0.f) pair_eval<KERNEL, PS, GS, HS, float>(dxa, dya, dza, r2a, __builtin_amdgcn_rsqf(r2a), gs, pa);
        if (r2b > 0.f) pair_eval<KERNEL, PS, GS, HS, float>(dxb, dyb, dzb, r2b, __builtin_amdgcn_rsqf(r2b), gs, pb);
    } else {   // a coincident pair (r = 0) contributes nothing: 1/r is selected to 0
        const float ria = r2a > 0.f ? __builtin_amdgcn_rsqf(r2a) : 0.f;
        const float rib = r2b > 0.f ? __builtin_amdgcn_rsqf(r2b) : 0.f;
        pair_eval<KERNEL, PS, GS, HS, float>(dxa, dya, dza, r2a, ria, gs, pa);
        pair_eval<KERNEL, PS, GS, HS, float>(dxb, dyb, dzb, r2b, rib, gs, pb);
    }
}


template <int KERNEL, bool PS, bool GS, bool HS>
__global__ __launch_bounds__(P2P_BLOCK) void k_p2p_f32(const int32_t* __restrict__ dir_targets,
                                                       const int64_t* __restrict__ dir_offsets,
                                                       const int32_t* __restrict__ dir_s, const int64_t* __restrict__ dir_cum,
                                                       const Branch* __restrict__ tbr,
                                                       const Branch* __restrict__ sbr, const double* __restrict__ tx,
                                                       const double* __restrict__ ty, const double* __restrict__ tz,
                                                       const double* __restrict__ sx, const double* __restrict__ sy,
                                                       const double* __restrict__ sz, const double* __restrict__ sq,
                                                       const double* __restrict__ srad, int64_t ns, double* __restrict__ tout,
                                                       int64_t nt, int64_t blk0,
                                                       unsigned long long* __restrict__ queue, int64_t n_queue) {
    constexpr int NS = KernelTraits<KERNEL>::NS;
    constexpr int NA = KernelTraits<KERNEL>::template na<float>(), NOUT = KernelTraits<KERNEL>::NOUT;
    // rows: source position hi (3) and lo (3) parts relative to the target leaf centre, staged values
    __shared__ float tile[6 + NS][P2P_BLOCK];
    // one target leaf per workgroup; with a queue (persistent launch) each workgroup takes leaves
    // from it until none are left
    __shared__ int64_t s_next;
    const int tid = threadIdx.x;
    int64_t bi = blk0 + blockIdx.x;
    for (;;) {
        if (queue) {
            __syncthreads();   // the previous leaf is done with the LDS tiles and s_next
            if (tid == 0) s_next = (int64_t)atomicAdd(queue, 1ull);
            __syncthreads();
            const int64_t next = uniform_i64(s_next);
            if (next >= n_queue) break;
            bi = blk0 + next;
        }
        const Branch T = tbr[dir_targets[bi]];
        const int64_t l0 = dir_offsets[bi], l1 = dir_offsets[bi + 1];
        const int64_t base = dir_cum[l0], end = dir_cum[l1];
        const double cx0 = T.target_center[0], cy0 = T.target_center[1], cz0 = T.target_center[2];
        const int64_t ngroups = (T.n_bodies + 1) / 2;
        const LeafLanes LL = leaf_lanes(ngroups);
        const int tpt = LL.t, lane_k = LL.lane_k;
        for (int64_t c0 = 0; c0 < ngroups; c0 += LL.per_pass) {
            const int64_t g = leaf_unit(LL, c0);
            const bool own = LL.gl < LL.gpw;
            const bool act0 = own && 2 * g < T.n_bodies, act1 = own && 2 * g + 1 < T.n_bodies;
            const int64_t i0 = T.first + (act0 ? 2 * g : 0), i1 = act1 ? i0 + 1 : i0;
            float th[2][3], tl[2][3];
    #pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int64_t it = h ? i1 : i0;
                const double r[3] = {tx[it] - cx0, ty[it] - cy0, tz[it] - cz0};
    #pragma unroll
                for (int d = 0; d < 3; ++d) { th[h][d] = (float)r[d]; tl[h][d] = (float)(r[d] - (double)th[h][d]); }
            }
            float sum[2][NA], cmp[2][NA];
    #pragma unroll
            for (int k = 0; k < NA; ++k) { sum[0][k] = sum[1][k] = 0.f; cmp[0][k] = cmp[1][k] = 0.f; }
            for (int64_t f0 = base; f0 < end; f0 += P2P_BLOCK) {
                const int cnt = (int)std::min<int64_t>(P2P_BLOCK, end - f0);
                __syncthreads();
                if (tid < cnt) {
                    const int64_t j = stream_body(dir_s, dir_cum, sbr, l0, l1, f0 + tid);
                    const double px = sx[j] - cx0, py = sy[j] - cy0, pz = sz[j] - cz0;
                    const float hx = (float)px, hy = (float)py, hz = (float)pz;
                    tile[0][tid] = hx; tile[1][tid] = hy; tile[2][tid] = hz;
                    tile[3][tid] = (float)(px - (double)hx); tile[4][tid] = (float)(py - (double)hy);
                    tile[5][tid] = (float)(pz - (double)hz);
                    stage_source<KERNEL, float>(sq, srad, ns, j, &tile[6][tid], P2P_BLOCK);
                }
                __syncthreads();
                if (act0) {
                    float pa[NA], pb[NA];
    #pragma unroll
                    for (int k = 0; k < NA; ++k) { pa[k] = 0.f; pb[k] = 0.f; }
                    // sources lane_k, lane_k + t, ... in uniform rounds (only the LDS address is per lane)
                    const int rounds = cnt / tpt;
                    const float* p = &tile[0][lane_k];
                    for (int r = 0; r < rounds; ++r, p += tpt) p2p_source_f32<KERNEL, PS, GS, HS>(p, th, tl, pa, pb);
                    if (lane_k < cnt - rounds * tpt) p2p_source_f32<KERNEL, PS, GS, HS>(p, th, tl, pa, pb);
    #pragma unroll
                    for (int k = 0; k < NA; ++k) { kahan_add(sum[0][k], cmp[0][k], pa[k]); kahan_add(sum[1][k], cmp[1][k], pb[k]); }
                }
            }
            double acc[2][NA];
    #pragma unroll
            for (int h = 0; h < 2; ++h)
    #pragma unroll
                for (int k = 0; k < NA; ++k) acc[h][k] = (double)sum[h][k] - (double)cmp[h][k];
    #pragma unroll
            for (int k = 0; k < NA; ++k)
                if (KernelTraits<KERNEL>::template used<float, PS, GS, HS>(k)) {
                    acc[0][k] = group_fold(acc[0][k], LL);
                    acc[1][k] = group_fold(acc[1][k], LL);
                }
            if (lane_k == 0) {
    #pragma unroll
                for (int h = 0; h < 2; ++h) {
                    if (!(h ? act1 : act0)) continue;
                    const int64_t it = h ? i1 : i0;
                    if (PS && KernelTraits<KERNEL>::POT) tout[it] += acc[h][0];
                    if (GS) { tout[nt + it] += acc[h][1]; tout[2 * nt + it] += acc[h][2]; tout[3 * nt + it] += acc[h][3]; }
                    if (HS && KernelTraits<KERNEL>::HESS)
                        for (int k = 4; k < NOUT; ++k) tout[(int64_t)k * nt + it] += acc[h][k];
                }
            }
        }
        if (!queue) break;
    }
}

__global__ void k_zero_queue(unsigned long long* q) { *q = 0ull; }

// Persistent launch: as many workgroups as the device holds at once (from the kernel's own
// register and LDS use), each taking target leaves from the queue in list order until it is
// empty, so no workgroup waits for a slot behind the last ones and no slot idles at the end.
// Every workgroup exits once the queue index passes the leaf count.
// slots > 0: at most that many workgroups per CU; slots < 0: that many fewer than fit (at least
// one); the rest of the CU stays free for kernels of other streams.
template <class... P, class... A>
static void launch_queued(Context& C, int slots, void (*kfn)(P...), int64_t nblk, A... args) {
    int per_cu = 0;
    HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void*>(kfn), P2P_BLOCK, 0));
    if (!C.n_cu) HIP_CHECK(hipDeviceGetAttribute(&C.n_cu, hipDeviceAttributeMultiprocessorCount, C.device));
    per_cu = std::max(1, slots > 0 ? std::min(per_cu, slots) : per_cu + slots);
    const int64_t grid = std::min<int64_t>(nblk, (int64_t)per_cu * C.n_cu);
    kfn<<<(unsigned)grid, P2P_BLOCK, 0, C.stream>>>(args...);
}

template <int K>
static void launch_p2p(Context& C, const Tree& tt, const Tree& st, const Lists& L, double* tout, int64_t nt, bool PS, bool GS,
                       bool HS, int precision, int64_t b0, int64_t b1, int slots, SourceView sv) {
    const int64_t nblk = b1 - b0;
    if (nblk <= 0) return;
    const Branch* sbr = sv.branches ? sv.branches : st.branches.data();
    const int64_t* cum = sv.dir_cum ? sv.dir_cum : L.dir_cum.data();
    C.p2p_queue.resize(1);
    unsigned long long* queue = C.p2p_queue.data();
    k_zero_queue<<<1, 1, 0, C.stream>>>(queue);
    LAUNCH_CHECK();
#define P2P_ARGS                                                                                                         \
    L.dir_targets.data(), L.dir_offsets.data(), L.dir_s.data(), cum, tt.branches.data(), sbr,                            \
        tt.x.data(),                                                                                                     \
        tt.y.data(), tt.z.data(), st.x.data(), st.y.data(), st.z.data(), st.strength.data(), st.radius.data(), st.n,    \
        tout, nt, b0, queue, nblk
#define P2P_DISPATCH(KFN)                                                                                               \
    do {                                                                                                                \
        int code = (PS ? 4 : 0) | (GS ? 2 : 0) | (HS ? 1 : 0);                                                          \
        switch (code) {                                                                                                 \
            case 0: break;                                                                                              \
            case 1: launch_queued(C, slots, KFN<K, false, false, true>, nblk, P2P_ARGS); break;                               \
            case 2: launch_queued(C, slots, KFN<K, false, true, false>, nblk, P2P_ARGS); break;                               \
            case 3: launch_queued(C, slots, KFN<K, false, true, true>, nblk, P2P_ARGS); break;                                \
            case 4: launch_queued(C, slots, KFN<K, true, false, false>, nblk, P2P_ARGS); break;                               \
            case 5: launch_queued(C, slots, KFN<K, true, false, true>, nblk, P2P_ARGS); break;                                \
            case 6: launch_queued(C, slots, KFN<K, true, true, false>, nblk, P2P_ARGS); break;                                \
            default: launch_queued(C, slots, KFN<K, true, true, true>, nblk, P2P_ARGS); break;                                \
        }                                                                                                               \
    } while (0)
    if (precision == FMMHIP_FP64) P2P_DISPATCH(k_p2p_f64);
    else P2P_DISPATCH(k_p2p_f32);
    LAUNCH_CHECK();
#undef P2P_DISPATCH
#undef P2P_ARGS
}

void run_p2p_panels(Context& C, int kernel, const Tree& tt, const Tree& st, const Lists& L, double* tout, int64_t nt, bool PS,
                    bool GS, bool HS, int64_t b0, int64_t b1, SourceView sv);
// targets b0 .. b1 - 1 of the direct list's CSR, none inside another
static void run_p2p_range(Context& C, int kernel, const Tree& tt, const Tree& st, const Lists& L, double* tout, int64_t nt,
                          bool PS, bool GS, bool HS, int precision, int64_t b0, int64_t b1, int slots, SourceView sv);
void run_direct_panels(Context& C, int kernel, int64_t nt, const double* t_xyz, int64_t ns, const double* s_buf,
                       double* tout_aos, bool PS, bool GS, bool HS);

void run_p2p(Context& C, int kernel, const Tree& tt, const Tree& st, const Lists& L, double* tout, int64_t nt, bool PS, bool GS,
             bool HS, int precision, int64_t b0, int64_t b1, int slots, SourceView sv) {
    if (b1 < 0) b1 = L.n_direct_targets;
    if (!L.dir_level_split.empty()) {   // nested targets: one launch per level, ancestors first
        const std::vector<int64_t>& s = L.dir_level_split;
        for (size_t l = 0; l + 1 < s.size(); ++l) {
            const int64_t a = std::max(b0, s[l]), b = std::min(b1, s[l + 1]);
            if (a < b) run_p2p_range(C, kernel, tt, st, L, tout, nt, PS, GS, HS, precision, a, b, slots, sv);
        }
        return;
    }
    run_p2p_range(C, kernel, tt, st, L, tout, nt, PS, GS, HS, precision, b0, b1, slots, sv);
}

static void run_p2p_range(Context& C, int kernel, const Tree& tt, const Tree& st, const Lists& L, double* tout, int64_t nt,
                          bool PS, bool GS, bool HS, int precision, int64_t b0, int64_t b1, int slots, SourceView sv) {
    if (is_panel(kernel)) { run_p2p_panels(C, kernel, tt, st, L, tout, nt, PS, GS, HS, b0, b1, sv); return; }
    switch (kernel) {
        case FMMHIP_POINT_SOURCE: launch_p2p<FMMHIP_POINT_SOURCE>(C, tt, st, L, tout, nt, PS, GS, HS, precision, b0, b1, slots, sv); break;
        case FMMHIP_POINT_VORTEX: launch_p2p<FMMHIP_POINT_VORTEX>(C, tt, st, L, tout, nt, PS, GS, HS, precision, b0, b1, slots, sv); break;
        case FMMHIP_POINT_DIPOLE: launch_p2p<FMMHIP_POINT_DIPOLE>(C, tt, st, L, tout, nt, PS, GS, HS, precision, b0, b1, slots, sv); break;
        case FMMHIP_POINT_SOURCE_VORTEX:
            launch_p2p<FMMHIP_POINT_SOURCE_VORTEX>(C, tt, st, L, tout, nt, PS, GS, HS, precision, b0, b1, slots, sv);
            break;
        default: launch_p2p<FMMHIP_VORTEX_GAUSSIAN>(C, tt, st, L, tout, nt, PS, GS, HS, precision, b0, b1, slots, sv); break;
    }
}

// ------------------------------------------------------ direct!, all pairs
// folded sums of one target -> its 16-entry AoS record (accumulated)
template <int KERNEL, class T>
__device__ inline void direct_store(double* o, const T* a, int PS, int GS, int HS) {
    if (KERNEL == FMMHIP_POINT_SOURCE) {
        o[3] += a[0];          // the reference kernel ignores the switches (test/gravitational.jl:92-98)
        o[4] += a[1]; o[5] += a[2]; o[6] += a[3];
    } else {
        if (PS && KernelTraits<KERNEL>::POT) o[3] += a[0];
        if (GS) { o[4] += a[1]; o[5] += a[2]; o[6] += a[3]; }
        if (HS) for (int k = 4; k < 13; ++k) o[3 + k] += a[k];
    }
}

// original-order AoS inputs; output accumulated into a 16 x nt AoS buffer
template <int KERNEL, typename R>
__global__ __launch_bounds__(P2P_BLOCK) void k_direct_all(const double* __restrict__ txyz, int64_t nt,
                                                          const double* __restrict__ sbuf, int rows, int64_t ns,
                                                          double* __restrict__ out, int PS, int GS, int HS) {
    constexpr int NS = KernelTraits<KERNEL>::NS;
    __shared__ double lx[P2P_BLOCK], ly[P2P_BLOCK], lz[P2P_BLOCK];
    __shared__ R lg[P2P_BLOCK][NS];
    const int64_t it = blockIdx.x * (int64_t)P2P_BLOCK + threadIdx.x;
    const bool active = it < nt;
    const double xt = active ? txyz[3 * it] : 0.0, yt = active ? txyz[3 * it + 1] : 0.0, zt = active ? txyz[3 * it + 2] : 0.0;
    constexpr int NA = KernelTraits<KERNEL>::template na<R>();
    R a[NA];
    for (int k = 0; k < NA; ++k) a[k] = R(0);
    for (int64_t s0 = 0; s0 < ns; s0 += P2P_BLOCK) {
        const int64_t cnt = std::min<int64_t>(P2P_BLOCK, ns - s0);
        __syncthreads();
        if ((int64_t)threadIdx.x < cnt) {
            const double* b = sbuf + (s0 + threadIdx.x) * rows;
            lx[threadIdx.x] = b[0]; ly[threadIdx.x] = b[1]; lz[threadIdx.x] = b[2];
            lg[threadIdx.x][0] = (R)(b[4] * INV4PI);
            if (NS >= 3) { lg[threadIdx.x][1] = (R)(b[5] * INV4PI); lg[threadIdx.x][2] = (R)(b[6] * INV4PI); }
            if (KERNEL == FMMHIP_VORTEX_GAUSSIAN) lg[threadIdx.x][3] = (R)b[3];
            if (KERNEL == FMMHIP_POINT_SOURCE_VORTEX) lg[threadIdx.x][NS - 1] = (R)(b[7] * INV4PI);
        }
        __syncthreads();
        if (!active) continue;
        for (int s = 0; s < cnt; ++s) {
            const R dx = (R)(xt - lx[s]), dy = (R)(yt - ly[s]), dz = (R)(zt - lz[s]);
            const R r2 = dx * dx + dy * dy + dz * dz;
            if (!(r2 > R(0))) continue;
            R ri;
            if constexpr (sizeof(R) == 8) ri = rinv_f64(r2); else ri = __builtin_amdgcn_rsqf(r2);
            pair_eval<KERNEL, true, true, true, R>(dx, dy, dz, r2, ri, lg[s], a);
        }
    }
    if (!active) return;
    fold_acc<KERNEL, true, R>(a);
    direct_store<KERNEL>(out + it * 16, a, PS, GS, HS);
}

// Few targets against many sources: the sources are split into S chunks and workgroup (b, c)
// evaluates target block b against chunk c, writing its folded sums to part[(c * NOUT + k) * nt + it]
// (every entry for it < nt written, nothing accumulated); k_direct_reduce then adds the chunks'
// sums per target in chunk order. One workgroup per target block would leave most of the device
// idle at, e.g., 10,000 sampled targets.
template <int KERNEL, typename R>
__global__ __launch_bounds__(P2P_BLOCK) void k_direct_part(const double* __restrict__ txyz, int64_t nt,
                                                           const double* __restrict__ sbuf, int rows, int64_t ns,
                                                           int64_t chunk, double* __restrict__ part) {
    constexpr int NS = KernelTraits<KERNEL>::NS;
    constexpr int NOUT = KernelTraits<KERNEL>::NOUT;
    __shared__ double lx[P2P_BLOCK], ly[P2P_BLOCK], lz[P2P_BLOCK];
    __shared__ R lg[P2P_BLOCK][NS];
    const int64_t it = blockIdx.x * (int64_t)P2P_BLOCK + threadIdx.x;
    const bool active = it < nt;
    const double xt = active ? txyz[3 * it] : 0.0, yt = active ? txyz[3 * it + 1] : 0.0, zt = active ? txyz[3 * it + 2] : 0.0;
    constexpr int NA = KernelTraits<KERNEL>::template na<R>();
    R a[NA];
    for (int k = 0; k < NA; ++k) a[k] = R(0);
    const int64_t c0 = (int64_t)blockIdx.y * chunk, c1 = std::min<int64_t>(ns, c0 + chunk);
    for (int64_t s0 = c0; s0 < c1; s0 += P2P_BLOCK) {
        const int64_t cnt = std::min<int64_t>(P2P_BLOCK, c1 - s0);
        __syncthreads();
        if ((int64_t)threadIdx.x < cnt) {
            const double* b = sbuf + (s0 + threadIdx.x) * rows;
            lx[threadIdx.x] = b[0]; ly[threadIdx.x] = b[1]; lz[threadIdx.x] = b[2];
            lg[threadIdx.x][0] = (R)(b[4] * INV4PI);
            if (NS >= 3) { lg[threadIdx.x][1] = (R)(b[5] * INV4PI); lg[threadIdx.x][2] = (R)(b[6] * INV4PI); }
            if (KERNEL == FMMHIP_VORTEX_GAUSSIAN) lg[threadIdx.x][3] = (R)b[3];
            if (KERNEL == FMMHIP_POINT_SOURCE_VORTEX) lg[threadIdx.x][NS - 1] = (R)(b[7] * INV4PI);
        }
        __syncthreads();
        if (!active) continue;
        for (int s = 0; s < cnt; ++s) {
            const R dx = (R)(xt - lx[s]), dy = (R)(yt - ly[s]), dz = (R)(zt - lz[s]);
            const R r2 = dx * dx + dy * dy + dz * dz;
            if (!(r2 > R(0))) continue;
            R ri;
            if constexpr (sizeof(R) == 8) ri = rinv_f64(r2); else ri = __builtin_amdgcn_rsqf(r2);
            pair_eval<KERNEL, true, true, true, R>(dx, dy, dz, r2, ri, lg[s], a);
        }
    }
    if (!active) return;
    fold_acc<KERNEL, true, R>(a);
    double* pc = part + (int64_t)blockIdx.y * NOUT * nt + it;
#pragma unroll
    for (int k = 0; k < NOUT; ++k) pc[(int64_t)k * nt] = (double)a[k];
}

template <int KERNEL>
__global__ __launch_bounds__(P2P_BLOCK) void k_direct_reduce(const double* __restrict__ part, int S, int64_t nt,
                                                             double* __restrict__ out, int PS, int GS, int HS) {
    constexpr int NOUT = KernelTraits<KERNEL>::NOUT;
    const int64_t it = blockIdx.x * (int64_t)P2P_BLOCK + threadIdx.x;
    if (it >= nt) return;
    double a[NOUT];
#pragma unroll
    for (int k = 0; k < NOUT; ++k) a[k] = 0.0;
    for (int c = 0; c < S; ++c) {
        const double* pc = part + (int64_t)c * NOUT * nt + it;
#pragma unroll
        for (int k = 0; k < NOUT; ++k) a[k] += pc[(int64_t)k * nt];
    }
    direct_store<KERNEL>(out + it * 16, a, PS, GS, HS);
}

// source chunks of the split direct! sum: 1 when the target blocks alone fill the device
// (>= 1024 workgroups); otherwise about 2048 workgroups in all, each chunk at least four source
// tiles. S * nt <= 2048 * P2P_BLOCK, so the partial sums need at most 13 * 8 * 524288 bytes.
int direct_split(int64_t nt, int64_t ns) {
    const int64_t tb = (nt + P2P_BLOCK - 1) / P2P_BLOCK;
    if (tb >= 1024 || nt == 0) return 1;
    int64_t S = 2048 / tb;
    S = std::min<int64_t>(S, std::max<int64_t>(1, ns / (4 * P2P_BLOCK)));
    return (int)std::max<int64_t>(1, S);
}

size_t direct_part_bytes(int kernel, int64_t nt, int64_t ns) {
    if (is_panel(kernel)) return 0;
    const int S = direct_split(nt, ns);
    if (S <= 1) return 0;
    const int nout = kernel == FMMHIP_POINT_SOURCE ? KernelTraits<FMMHIP_POINT_SOURCE>::NOUT : 13;
    return sizeof(double) * (size_t)S * (size_t)nout * (size_t)nt;
}

// Static workspace of the allocation-free direct! entry point (fmmhip_direct_noalloc): the largest
// chunk-sum buffer direct_split can ask for, 13 sums x 2048 x P2P_BLOCK targets. Every GPU the
// library runs on holds one copy until the process ends. In host code `scratch` names a host-side
// stand-in, not device memory: kernels and copies take the address hipGetSymbolAddress returns for
// the current device, which direct_scratch looks up once per device and keeps.
static constexpr size_t DIRECT_SCRATCH_DOUBLES = (size_t)13 * 2048 * P2P_BLOCK;
static __device__ __align__(256) double scratch[DIRECT_SCRATCH_DOUBLES];

size_t direct_scratch_bytes() { return sizeof(double) * DIRECT_SCRATCH_DOUBLES; }

double* direct_scratch(int device) {
    static std::mutex mu;
    static std::map<int, double*> ptrs;
    std::lock_guard<std::mutex> lk(mu);
    auto it = ptrs.find(device);
    if (it != ptrs.end()) return it->second;
    int cur = -1;
    HIP_CHECK(hipGetDevice(&cur));
    if (cur != device) HIP_CHECK(hipSetDevice(device));
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (cur != device) HIP_CHECK(hipSetDevice(cur));
    HIP_CHECK(e);
    if (!p) throw Error(FMMHIP_EDEVICE, "hipGetSymbolAddress returned no address for the direct! workspace");
    ptrs[device] = (double*)p;
    return (double*)p;
}

template <int K>
static void launch_direct_all(Context& C, int64_t nt, const double* t_xyz, int64_t ns, const double* s_buf, double* tout_aos,
                              bool PS, bool GS, bool HS, int precision, double* part) {
    const unsigned grid = (unsigned)((nt + P2P_BLOCK - 1) / P2P_BLOCK);
    const int rows = data_per_body(K);
    const int S = direct_split(nt, ns);
    if (S > 1) {
        int64_t chunk = (ns + S - 1) / S;
        chunk = (chunk + P2P_BLOCK - 1) / P2P_BLOCK * P2P_BLOCK;
        const int Sr = (int)((ns + chunk - 1) / chunk);   // chunks that hold sources
        const dim3 g2(grid, (unsigned)Sr);
        if (precision == FMMHIP_FP64)
            k_direct_part<K, double><<<g2, P2P_BLOCK, 0, C.stream>>>(t_xyz, nt, s_buf, rows, ns, chunk, part);
        else
            k_direct_part<K, float><<<g2, P2P_BLOCK, 0, C.stream>>>(t_xyz, nt, s_buf, rows, ns, chunk, part);
        LAUNCH_CHECK();
        k_direct_reduce<K><<<grid, P2P_BLOCK, 0, C.stream>>>(part, Sr, nt, tout_aos, PS, GS, HS);
    } else if (precision == FMMHIP_FP64)
        k_direct_all<K, double><<<grid, P2P_BLOCK, 0, C.stream>>>(t_xyz, nt, s_buf, rows, ns, tout_aos, PS, GS, HS);
    else
        k_direct_all<K, float><<<grid, P2P_BLOCK, 0, C.stream>>>(t_xyz, nt, s_buf, rows, ns, tout_aos, PS, GS, HS);
    LAUNCH_CHECK();
}

void run_direct_all(Context& C, int kernel, int64_t nt, const double* t_xyz, int64_t ns, const double* s_buf, double* tout_aos,
                    bool PS, bool GS, bool HS, int precision, double* part) {
    if (nt == 0 || ns == 0) return;
    if (is_panel(kernel)) { run_direct_panels(C, kernel, nt, t_xyz, ns, s_buf, tout_aos, PS, GS, HS); return; }
    const size_t need = direct_part_bytes(kernel, nt, ns);
    if (need && !part) {   // the context's workspace for the chunk sums
        C.direct_part.resize(need / sizeof(double));
        part = C.direct_part.data();
    }
    switch (kernel) {
        case FMMHIP_POINT_SOURCE:
            launch_direct_all<FMMHIP_POINT_SOURCE>(C, nt, t_xyz, ns, s_buf, tout_aos, PS, GS, HS, precision, part);
            break;
        case FMMHIP_POINT_VORTEX:
            launch_direct_all<FMMHIP_POINT_VORTEX>(C, nt, t_xyz, ns, s_buf, tout_aos, PS, GS, HS, precision, part);
            break;
        case FMMHIP_POINT_DIPOLE:
            launch_direct_all<FMMHIP_POINT_DIPOLE>(C, nt, t_xyz, ns, s_buf, tout_aos, PS, GS, HS, precision, part);
            break;
        case FMMHIP_POINT_SOURCE_VORTEX:
            launch_direct_all<FMMHIP_POINT_SOURCE_VORTEX>(C, nt, t_xyz, ns, s_buf, tout_aos, PS, GS, HS, precision, part);
            break;
        default:
            launch_direct_all<FMMHIP_VORTEX_GAUSSIAN>(C, nt, t_xyz, ns, s_buf, tout_aos, PS, GS, HS, precision, part);
            break;
    }
}

}  // namespace fmmhip

#include "panels.hpp"

namespace fmmhip {


template <int NV, int MODE, bool PS, bool GS, bool HS>
__global__ __launch_bounds__(P2P_BLOCK) void k_p2p_panel(const int32_t* __restrict__ dir_targets,
                                                         const int64_t* __restrict__ dir_offsets,
                                                         const int32_t* __restrict__ dir_s, const int64_t* __restrict__ dir_cum,
                                                         const Branch* __restrict__ tbr, const Branch* __restrict__ sbr,
                                                         const double* __restrict__ tx, const double* __restrict__ ty,
                                                         const double* __restrict__ tz, const double* __restrict__ sx,
                                                         const double* __restrict__ sy, const double* __restrict__ sz,
                                                         const double* __restrict__ rows, int64_t ns, double* __restrict__ tout,
                                                         int64_t nt, int64_t blk0) {
    __shared__ PanelStage<NV> tile[PANEL_TILE];
    const int64_t bi = blk0 + blockIdx.x;   // target leaf of this workgroup
    const Branch T = tbr[dir_targets[bi]];
    const int64_t l0 = dir_offsets[bi], l1 = dir_offsets[bi + 1];
    const int64_t base = dir_cum[l0], end = dir_cum[l1];
    const int tid = threadIdx.x;
    const LeafLanes LL = leaf_lanes(T.n_bodies);
    const int tpt = LL.t, lane_k = LL.lane_k;
    for (int64_t c0 = 0; c0 < T.n_bodies; c0 += LL.per_pass) {
        const int64_t ti = leaf_unit(LL, c0);
        const bool active = LL.gl < LL.gpw && ti < T.n_bodies;
        const int64_t it = T.first + (active ? ti : 0);
        const double xt = tx[it], yt = ty[it], zt = tz[it];
        double a[13];
#pragma unroll
        for (int k = 0; k < 13; ++k) a[k] = 0.0;
        for (int64_t f0 = base; f0 < end; f0 += PANEL_TILE) {
            const int64_t cnt = std::min<int64_t>(PANEL_TILE, end - f0);
            __syncthreads();
            if (tid < cnt)
                stage_panel<NV, MODE == 2 ? 2 : 1>(sx, sy, sz, rows, ns, stream_body(dir_s, dir_cum, sbr, l0, l1, f0 + tid), tile[tid]);
            __syncthreads();
            if (active)
                for (int s = lane_k; s < cnt; s += tpt) panel_pair_kind<NV, MODE, PS, GS, HS>(tile[s], xt, yt, zt, a);
        }
#pragma unroll
        for (int k = 0; k < 13; ++k) a[k] = group_fold(a[k], LL);
        if (active && lane_k == 0) {
            if (PS) tout[it] += a[0];
            if (GS) { tout[nt + it] += a[1]; tout[2 * nt + it] += a[2]; tout[3 * nt + it] += a[3]; }
            if (HS) for (int k = 4; k < 13; ++k) tout[(int64_t)k * nt + it] += a[k];
        }
    }
}

template <int NV, int MODE>
static void launch_p2p_panel(Context& C, const Tree& tt, const Tree& st, const Lists& L, double* tout, int64_t nt, bool PS,
                             bool GS, bool HS, int64_t b0, int64_t b1, SourceView sv) {
    const int64_t nblk = b1 - b0;
    if (nblk <= 0) return;
    const Branch* sbr = sv.branches ? sv.branches : st.branches.data();
    const int64_t* cum = sv.dir_cum ? sv.dir_cum : L.dir_cum.data();
#define PANEL_ARGS                                                                                                       \
    L.dir_targets.data(), L.dir_offsets.data(), L.dir_s.data(), cum, tt.branches.data(), sbr,                            \
        tt.x.data(), tt.y.data(), tt.z.data(), st.x.data(), st.y.data(), st.z.data(), st.strength.data(), st.n, tout, nt, b0
    const int code = (PS ? 4 : 0) | (GS ? 2 : 0) | (HS ? 1 : 0);
    switch (code) {
        case 0: break;
        case 1: k_p2p_panel<NV, MODE, false, false, true><<<(unsigned)nblk, P2P_BLOCK, 0, C.stream>>>(PANEL_ARGS); break;
        case 2: k_p2p_panel<NV, MODE, false, true, false><<<(unsigned)nblk, P2P_BLOCK, 0, C.stream>>>(PANEL_ARGS); break;
        case 3: k_p2p_panel<NV, MODE, false, true, true><<<(unsigned)nblk, P2P_BLOCK, 0, C.stream>>>(PANEL_ARGS); break;
        case 4: k_p2p_panel<NV, MODE, true, false, false><<<(unsigned)nblk, P2P_BLOCK, 0, C.stream>>>(PANEL_ARGS); break;
        case 5: k_p2p_panel<NV, MODE, true, false, true><<<(unsigned)nblk, P2P_BLOCK, 0, C.stream>>>(PANEL_ARGS); break;
        case 6: k_p2p_panel<NV, MODE, true, true, false><<<(unsigned)nblk, P2P_BLOCK, 0, C.stream>>>(PANEL_ARGS); break;
        default: k_p2p_panel<NV, MODE, true, true, true><<<(unsigned)nblk, P2P_BLOCK, 0, C.stream>>>(PANEL_ARGS); break;
    }
#undef PANEL_ARGS
    LAUNCH_CHECK();
}

void run_p2p_panels(Context& C, int kernel, const Tree& tt, const Tree& st, const Lists& L, double* tout, int64_t nt, bool PS,
                    bool GS, bool HS, int64_t b0, int64_t b1, SourceView sv) {
    switch (kernel) {
        case FMMHIP_PANEL_SOURCE_TRI: launch_p2p_panel<3, 0>(C, tt, st, L, tout, nt, PS, GS, HS, b0, b1, sv); break;
        case FMMHIP_PANEL_DIPOLE_TRI: launch_p2p_panel<3, 1>(C, tt, st, L, tout, nt, PS, GS, HS, b0, b1, sv); break;
        case FMMHIP_PANEL_SOURCE_QUAD: launch_p2p_panel<4, 0>(C, tt, st, L, tout, nt, PS, GS, HS, b0, b1, sv); break;
        case FMMHIP_PANEL_SOURCE_DIPOLE_TRI: launch_p2p_panel<3, 2>(C, tt, st, L, tout, nt, PS, GS, HS, b0, b1, sv); break;
        case FMMHIP_PANEL_SOURCE_DIPOLE_QUAD: launch_p2p_panel<4, 2>(C, tt, st, L, tout, nt, PS, GS, HS, b0, b1, sv); break;
        default: launch_p2p_panel<4, 1>(C, tt, st, L, tout, nt, PS, GS, HS, b0, b1, sv); break;
    }
}

// direct! for panels: original-order AoS inputs, 16 x nt AoS output (+=)
template <int NV, int MODE>
__global__ __launch_bounds__(P2P_BLOCK) void k_direct_panels(const double* __restrict__ txyz, int64_t nt,
                                                             const double* __restrict__ sbuf, int64_t ns,
                                                             double* __restrict__ out, int PS, int GS, int HS) {
    __shared__ PanelStage<NV> tile[PANEL_TILE];
    constexpr int NSTR = MODE == 2 ? 2 : 1;
    constexpr int rows = 4 + NSTR + 3 * NV;
    const int64_t it = blockIdx.x * (int64_t)P2P_BLOCK + threadIdx.x;
    const bool active = it < nt;
    const double xt = active ? txyz[3 * it] : 0.0, yt = active ? txyz[3 * it + 1] : 0.0, zt = active ? txyz[3 * it + 2] : 0.0;
    double a[13];
    for (int k = 0; k < 13; ++k) a[k] = 0.0;
    for (int64_t s0 = 0; s0 < ns; s0 += PANEL_TILE) {
        const int64_t cnt = std::min<int64_t>(PANEL_TILE, ns - s0);
        __syncthreads();
        if ((int64_t)threadIdx.x < cnt) {
            const double* b = sbuf + (s0 + threadIdx.x) * rows;   // one panel's row as a 1-body SoA
            stage_panel<NV, NSTR>(b, b + 1, b + 2, b + 4, 1, 0, tile[threadIdx.x]);
        }
        __syncthreads();
        if (!active) continue;
        for (int s = 0; s < cnt; ++s) panel_pair_kind<NV, MODE, true, true, true>(tile[s], xt, yt, zt, a);
    }
    if (!active) return;
    double* o = out + it * 16;
    if (PS) o[3] += a[0];
    if (GS) { o[4] += a[1]; o[5] += a[2]; o[6] += a[3]; }
    if (HS) for (int k = 4; k < 13; ++k) o[3 + k] += a[k];
}

void run_direct_panels(Context& C, int kernel, int64_t nt, const double* t_xyz, int64_t ns, const double* s_buf,
                       double* tout_aos, bool PS, bool GS, bool HS) {
    const unsigned grid = (unsigned)((nt + P2P_BLOCK - 1) / P2P_BLOCK);
    switch (kernel) {
        case FMMHIP_PANEL_SOURCE_TRI: k_direct_panels<3, 0><<<grid, P2P_BLOCK, 0, C.stream>>>(t_xyz, nt, s_buf, ns, tout_aos, PS, GS, HS); break;
        case FMMHIP_PANEL_DIPOLE_TRI: k_direct_panels<3, 1><<<grid, P2P_BLOCK, 0, C.stream>>>(t_xyz, nt, s_buf, ns, tout_aos, PS, GS, HS); break;
        case FMMHIP_PANEL_SOURCE_QUAD: k_direct_panels<4, 0><<<grid, P2P_BLOCK, 0, C.stream>>>(t_xyz, nt, s_buf, ns, tout_aos, PS, GS, HS); break;
        case FMMHIP_PANEL_SOURCE_DIPOLE_TRI:
            k_direct_panels<3, 2><<<grid, P2P_BLOCK, 0, C.stream>>>(t_xyz, nt, s_buf, ns, tout_aos, PS, GS, HS);
            break;
        case FMMHIP_PANEL_SOURCE_DIPOLE_QUAD:
            k_direct_panels<4, 2><<<grid, P2P_BLOCK, 0, C.stream>>>(t_xyz, nt, s_buf, ns, tout_aos, PS, GS, HS);
            break;
        default: k_direct_panels<4, 1><<<grid, P2P_BLOCK, 0, C.stream>>>(t_xyz, nt, s_buf, ns, tout_aos, PS, GS, HS); break;
    }
    LAUNCH_CHECK();
}

}  // namespace fmmhip
