// Nearfield P2P on the GPU for the built-in point kernels:
//   Point{Source}: Laplace 1/r potential + gradient  (test/gravitational.jl:80-104)
//   Point{Vortex}: singular vortex velocity + velocity gradient (test/vortex.jl:83-128)
// driven by the direct list sorted by target (src/fmm.jl:18-40).
//
// One workgroup per target leaf. Targets live in registers (power-of-two
// threads per target when the leaf is small, each thread taking a strided share
// of the sources); each source leaf of the list is staged through LDS and
// broadcast to the whole workgroup. The partial sums of the threads of one
// target are folded with wavefront shuffles.
//
// FP64 form: IEEE double arithmetic with 1/r from a refined hardware rsqrt.
// Mixed form: FP32 arithmetic; positions are stored relative to the target
// leaf centre as a hi+lo pair of floats (differences of nearby bodies keep
// ~2^-48 relative accuracy), and each thread's per-source-leaf FP32 partial
// sum is added to its running total with Kahan compensation. Strengths are
// pre-scaled by 1/(4 pi).
#include "fmmhip_internal.hpp"
#include <algorithm>

namespace fmmhip {

static constexpr int P2P_BLOCK = 256;
static constexpr double INV4PI = 1.0 / (4.0 * 3.141592653589793);

__device__ inline double rinv_f64(double r2) {
    double y = __builtin_amdgcn_rsq(r2);            // ~23-bit estimate
    double h = 0.5 * r2;
    y = y * (1.5 - h * y * y);                       // Newton: ~46 bits
    y = y * (1.5 - h * y * y);                       // Newton: full double
    return y;
}

template <class R> struct Acc { R v[13]; };

// Kahan-compensated add of a partial sum into (s, c)
__device__ inline void kahan_add(float& s, float& c, float x) {
    float y = x - c;
    float t = s + y;
    c = (t - s) - y;
    s = t;
}

// ---------------------------------------------------------------- FP64 form
template <int KERNEL, bool PS, bool GS, bool HS>
__global__ __launch_bounds__(P2P_BLOCK) void k_p2p_f64(const int32_t* __restrict__ dir_targets,
                                                       const int64_t* __restrict__ dir_offsets,
                                                       const int32_t* __restrict__ dir_s, const Branch* __restrict__ tbr,
                                                       const Branch* __restrict__ sbr, const double* __restrict__ tx,
                                                       const double* __restrict__ ty, const double* __restrict__ tz,
                                                       const double* __restrict__ sx, const double* __restrict__ sy,
                                                       const double* __restrict__ sz, const double* __restrict__ sq,
                                                       int64_t ns, double* __restrict__ tout, int64_t nt) {
    __shared__ double lx[P2P_BLOCK], ly[P2P_BLOCK], lz[P2P_BLOCK], lg[3][P2P_BLOCK];
    const Branch T = tbr[dir_targets[blockIdx.x]];
    const int64_t l0 = dir_offsets[blockIdx.x], l1 = dir_offsets[blockIdx.x + 1];
    const int tid = threadIdx.x;
    int tpt = 1;   // threads per target
    while (tpt < 64 && T.n_bodies * (tpt * 2) <= P2P_BLOCK) tpt *= 2;
    const int per_chunk = P2P_BLOCK / tpt;
    for (int64_t c0 = 0; c0 < T.n_bodies; c0 += per_chunk) {
        const int64_t ti = c0 + tid / tpt;
        const int lane_k = tid % tpt;
        const bool active = ti < T.n_bodies;
        const int64_t it = T.first + (active ? ti : 0);
        const double xt = tx[it], yt = ty[it], zt = tz[it];
        Acc<double> a;
#pragma unroll
        for (int k = 0; k < 13; ++k) a.v[k] = 0.0;
        for (int64_t li = l0; li < l1; ++li) {
            const Branch S = sbr[dir_s[li]];
            for (int64_t s0 = 0; s0 < S.n_bodies; s0 += P2P_BLOCK) {
                const int64_t cnt = std::min<int64_t>(P2P_BLOCK, S.n_bodies - s0);
                __syncthreads();
                if (tid < cnt) {
                    const int64_t j = S.first + s0 + tid;
                    lx[tid] = sx[j]; ly[tid] = sy[j]; lz[tid] = sz[j];
                    if (KERNEL == FMMHIP_POINT_SOURCE) lg[0][tid] = sq[j] * INV4PI;
                    else { lg[0][tid] = sq[j] * INV4PI; lg[1][tid] = sq[ns + j] * INV4PI; lg[2][tid] = sq[2 * ns + j] * INV4PI; }
                }
                __syncthreads();
                if (active)
                    for (int s = lane_k; s < cnt; s += tpt) {
                        const double dx = xt - lx[s], dy = yt - ly[s], dz = zt - lz[s];
                        const double r2 = dx * dx + dy * dy + dz * dz;
                        if (r2 > 0.0) {
                            const double ri = rinv_f64(r2);
                            const double ri2 = ri * ri;
                            if (KERNEL == FMMHIP_POINT_SOURCE) {
                                const double q = lg[0][s];
                                if (PS) a.v[0] += q * ri;
                                if (GS) { const double f = q * ri * ri2; a.v[1] += dx * f; a.v[2] += dy * f; a.v[3] += dz * f; }
                            } else {
                                const double gx = lg[0][s], gy = lg[1][s], gz = lg[2][s];
                                const double d3 = ri * ri2;
                                const double cx = dz * gy - dy * gz, cy = dx * gz - dz * gx, cz = dy * gx - dx * gy;
                                if (GS) { a.v[1] += cx * d3; a.v[2] += cy * d3; a.v[3] += cz * d3; }
                                if (HS) {
                                    const double d5 = d3 * ri2, t3 = -3.0 * d5;
                                    const double ax = t3 * dx, ay = t3 * dy, az = t3 * dz;
                                    const double g3x = gx * d3, g3y = gy * d3, g3z = gz * d3;
                                    a.v[4] += ax * cx;         a.v[5] += ax * cy + g3z;   a.v[6] += ax * cz - g3y;
                                    a.v[7] += ay * cx - g3z;   a.v[8] += ay * cy;         a.v[9] += ay * cz + g3x;
                                    a.v[10] += az * cx + g3y;  a.v[11] += az * cy - g3x;  a.v[12] += az * cz;
                                }
                            }
                        }
                    }
            }
        }
        // fold the tpt partial sums of a target (adjacent lanes)
        for (int w = 1; w < tpt; w <<= 1)
#pragma unroll
            for (int k = 0; k < 13; ++k) a.v[k] += __shfl_xor(a.v[k], w, 64);
        if (active && lane_k == 0) {
            if (PS) tout[it] += a.v[0];
            if (GS) { tout[nt + it] += a.v[1]; tout[2 * nt + it] += a.v[2]; tout[3 * nt + it] += a.v[3]; }
            if (HS && KERNEL == FMMHIP_POINT_VORTEX)
                for (int k = 4; k < 13; ++k) tout[(int64_t)k * nt + it] += a.v[k];
        }
    }
}

// --------------------------------------------------------------- mixed form
template <int KERNEL, bool PS, bool GS, bool HS>
__global__ __launch_bounds__(P2P_BLOCK) void k_p2p_f32(const int32_t* __restrict__ dir_targets,
                                                       const int64_t* __restrict__ dir_offsets,
                                                       const int32_t* __restrict__ dir_s, const Branch* __restrict__ tbr,
                                                       const Branch* __restrict__ sbr, const double* __restrict__ tx,
                                                       const double* __restrict__ ty, const double* __restrict__ tz,
                                                       const double* __restrict__ sx, const double* __restrict__ sy,
                                                       const double* __restrict__ sz, const double* __restrict__ sq,
                                                       int64_t ns, double* __restrict__ tout, int64_t nt) {
    __shared__ float lh[3][P2P_BLOCK], ll[3][P2P_BLOCK], lg[3][P2P_BLOCK];
    const Branch T = tbr[dir_targets[blockIdx.x]];
    const int64_t l0 = dir_offsets[blockIdx.x], l1 = dir_offsets[blockIdx.x + 1];
    const int tid = threadIdx.x;
    const double cx0 = T.target_center[0], cy0 = T.target_center[1], cz0 = T.target_center[2];
    int tpt = 1;
    while (tpt < 64 && T.n_bodies * (tpt * 2) <= P2P_BLOCK) tpt *= 2;
    const int per_chunk = P2P_BLOCK / tpt;
    constexpr int NA = KERNEL == FMMHIP_POINT_SOURCE ? 4 : 13;
    for (int64_t c0 = 0; c0 < T.n_bodies; c0 += per_chunk) {
        const int64_t ti = c0 + tid / tpt;
        const int lane_k = tid % tpt;
        const bool active = ti < T.n_bodies;
        const int64_t it = T.first + (active ? ti : 0);
        const double rx = tx[it] - cx0, ry = ty[it] - cy0, rz = tz[it] - cz0;
        const float xh = (float)rx, yh = (float)ry, zh = (float)rz;
        const float xl = (float)(rx - (double)xh), yl = (float)(ry - (double)yh), zl = (float)(rz - (double)zh);
        float sum[NA], cmp[NA];
#pragma unroll
        for (int k = 0; k < NA; ++k) { sum[k] = 0.f; cmp[k] = 0.f; }
        for (int64_t li = l0; li < l1; ++li) {
            const Branch S = sbr[dir_s[li]];
            for (int64_t s0 = 0; s0 < S.n_bodies; s0 += P2P_BLOCK) {
                const int64_t cnt = std::min<int64_t>(P2P_BLOCK, S.n_bodies - s0);
                __syncthreads();
                if (tid < cnt) {
                    const int64_t j = S.first + s0 + tid;
                    const double px = sx[j] - cx0, py = sy[j] - cy0, pz = sz[j] - cz0;
                    const float hx = (float)px, hy = (float)py, hz = (float)pz;
                    lh[0][tid] = hx; lh[1][tid] = hy; lh[2][tid] = hz;
                    ll[0][tid] = (float)(px - (double)hx); ll[1][tid] = (float)(py - (double)hy); ll[2][tid] = (float)(pz - (double)hz);
                    if (KERNEL == FMMHIP_POINT_SOURCE) lg[0][tid] = (float)(sq[j] * INV4PI);
                    else {
                        lg[0][tid] = (float)(sq[j] * INV4PI);
                        lg[1][tid] = (float)(sq[ns + j] * INV4PI);
                        lg[2][tid] = (float)(sq[2 * ns + j] * INV4PI);
                    }
                }
                __syncthreads();
                if (active) {
                    float part[NA];
#pragma unroll
                    for (int k = 0; k < NA; ++k) part[k] = 0.f;
                    for (int s = lane_k; s < cnt; s += tpt) {
                        const float dx = (xh - lh[0][s]) + (xl - ll[0][s]);
                        const float dy = (yh - lh[1][s]) + (yl - ll[1][s]);
                        const float dz = (zh - lh[2][s]) + (zl - ll[2][s]);
                        const float r2 = dx * dx + dy * dy + dz * dz;
                        if (r2 > 0.f) {
                            const float ri = __builtin_amdgcn_rsqf(r2);
                            const float ri2 = ri * ri;
                            if (KERNEL == FMMHIP_POINT_SOURCE) {
                                const float q = lg[0][s];
                                if (PS) part[0] += q * ri;
                                if (GS) { const float f = q * ri * ri2; part[1] += dx * f; part[2] += dy * f; part[3] += dz * f; }
                            } else {
                                const float gx = lg[0][s], gy = lg[1][s], gz = lg[2][s];
                                const float d3 = ri * ri2;
                                const float cx = dz * gy - dy * gz, cy = dx * gz - dz * gx, cz = dy * gx - dx * gy;
                                if (GS) { part[1] += cx * d3; part[2] += cy * d3; part[3] += cz * d3; }
                                if (HS) {
                                    const float d5 = d3 * ri2, t3 = -3.f * d5;
                                    const float ax = t3 * dx, ay = t3 * dy, az = t3 * dz;
                                    const float g3x = gx * d3, g3y = gy * d3, g3z = gz * d3;
                                    part[4] += ax * cx;         part[5] += ax * cy + g3z;   part[6] += ax * cz - g3y;
                                    part[7] += ay * cx - g3z;   part[8] += ay * cy;         part[9] += ay * cz + g3x;
                                    part[10] += az * cx + g3y;  part[11] += az * cy - g3x;  part[12] += az * cz;
                                }
                            }
                        }
                    }
#pragma unroll
                    for (int k = 0; k < NA; ++k) kahan_add(sum[k], cmp[k], part[k]);
                }
            }
        }
        double acc[NA];
#pragma unroll
        for (int k = 0; k < NA; ++k) acc[k] = (double)sum[k] - (double)cmp[k];
        for (int w = 1; w < tpt; w <<= 1)
#pragma unroll
            for (int k = 0; k < NA; ++k) acc[k] += __shfl_xor(acc[k], w, 64);
        if (active && lane_k == 0) {
            if (PS && KERNEL == FMMHIP_POINT_SOURCE) tout[it] += acc[0];
            if (GS) { tout[nt + it] += acc[1]; tout[2 * nt + it] += acc[2]; tout[3 * nt + it] += acc[3]; }
            if (HS && KERNEL == FMMHIP_POINT_VORTEX)
                for (int k = 4; k < NA; ++k) tout[(int64_t)k * nt + it] += acc[k];
        }
    }
}

template <int K>
static void launch_p2p(Context& C, const Tree& tt, const Tree& st, const Lists& L, double* tout, int64_t nt, bool PS, bool GS,
                       bool HS, int precision) {
    const int64_t nblk = L.n_direct_targets;
    if (nblk == 0) return;
#define P2P_ARGS                                                                                                         \
    L.dir_targets.data(), L.dir_offsets.data(), L.dir_s.data(), tt.branches.data(), st.branches.data(), tt.x.data(),    \
        tt.y.data(), tt.z.data(), st.x.data(), st.y.data(), st.z.data(), st.strength.data(), st.n, tout, nt
#define P2P_DISPATCH(KFN)                                                                                               \
    do {                                                                                                                \
        int code = (PS ? 4 : 0) | (GS ? 2 : 0) | (HS ? 1 : 0);                                                          \
        switch (code) {                                                                                                 \
            case 0: break;                                                                                              \
            case 1: KFN<K, false, false, true><<<(unsigned)nblk, P2P_BLOCK, 0, C.stream>>>(P2P_ARGS); break;             \
            case 2: KFN<K, false, true, false><<<(unsigned)nblk, P2P_BLOCK, 0, C.stream>>>(P2P_ARGS); break;             \
            case 3: KFN<K, false, true, true><<<(unsigned)nblk, P2P_BLOCK, 0, C.stream>>>(P2P_ARGS); break;              \
            case 4: KFN<K, true, false, false><<<(unsigned)nblk, P2P_BLOCK, 0, C.stream>>>(P2P_ARGS); break;             \
            case 5: KFN<K, true, false, true><<<(unsigned)nblk, P2P_BLOCK, 0, C.stream>>>(P2P_ARGS); break;              \
            case 6: KFN<K, true, true, false><<<(unsigned)nblk, P2P_BLOCK, 0, C.stream>>>(P2P_ARGS); break;              \
            default: KFN<K, true, true, true><<<(unsigned)nblk, P2P_BLOCK, 0, C.stream>>>(P2P_ARGS); break;              \
        }                                                                                                               \
    } while (0)
    if (precision == FMMHIP_FP64) P2P_DISPATCH(k_p2p_f64);
    else P2P_DISPATCH(k_p2p_f32);
    LAUNCH_CHECK();
#undef P2P_DISPATCH
#undef P2P_ARGS
}

void run_p2p(Context& C, int kernel, const Tree& tt, const Tree& st, const Lists& L, double* tout, int64_t nt, bool PS, bool GS,
             bool HS, int precision) {
    if (kernel == FMMHIP_POINT_SOURCE) launch_p2p<FMMHIP_POINT_SOURCE>(C, tt, st, L, tout, nt, PS, GS, HS, precision);
    else launch_p2p<FMMHIP_POINT_VORTEX>(C, tt, st, L, tout, nt, PS, GS, HS, precision);
}

// ------------------------------------------------------ direct!, all pairs
// original-order AoS inputs; output accumulated into a 16 x nt AoS buffer
template <int KERNEL, typename R>
__global__ __launch_bounds__(P2P_BLOCK) void k_direct_all(const double* __restrict__ txyz, int64_t nt,
                                                          const double* __restrict__ sbuf, int rows, int64_t ns,
                                                          double* __restrict__ out, int PS, int GS, int HS) {
    __shared__ R lx[P2P_BLOCK], ly[P2P_BLOCK], lz[P2P_BLOCK], lg[3][P2P_BLOCK];
    const int64_t it = blockIdx.x * (int64_t)P2P_BLOCK + threadIdx.x;
    const bool active = it < nt;
    const double xt = active ? txyz[3 * it] : 0.0, yt = active ? txyz[3 * it + 1] : 0.0, zt = active ? txyz[3 * it + 2] : 0.0;
    double a[13];
    for (int k = 0; k < 13; ++k) a[k] = 0.0;
    for (int64_t s0 = 0; s0 < ns; s0 += P2P_BLOCK) {
        const int64_t cnt = std::min<int64_t>(P2P_BLOCK, ns - s0);
        __syncthreads();
        if ((int64_t)threadIdx.x < cnt) {
            const double* b = sbuf + (s0 + threadIdx.x) * rows;
            lx[threadIdx.x] = (R)b[0]; ly[threadIdx.x] = (R)b[1]; lz[threadIdx.x] = (R)b[2];
            lg[0][threadIdx.x] = (R)(b[4] * INV4PI);
            if (KERNEL == FMMHIP_POINT_VORTEX) { lg[1][threadIdx.x] = (R)(b[5] * INV4PI); lg[2][threadIdx.x] = (R)(b[6] * INV4PI); }
        }
        __syncthreads();
        if (!active) continue;
        for (int s = 0; s < cnt; ++s) {
            const R dx = (R)(xt - (double)lx[s]), dy = (R)(yt - (double)ly[s]), dz = (R)(zt - (double)lz[s]);
            const R r2 = dx * dx + dy * dy + dz * dz;
            if (!(r2 > R(0))) continue;
            R ri;
            if constexpr (sizeof(R) == 8) ri = rinv_f64(r2); else ri = __builtin_amdgcn_rsqf(r2);
            const R ri2 = ri * ri;
            if (KERNEL == FMMHIP_POINT_SOURCE) {
                const R q = lg[0][s];
                a[0] += q * ri;
                const R f = q * ri * ri2;
                a[1] += dx * f; a[2] += dy * f; a[3] += dz * f;
            } else {
                const R gx = lg[0][s], gy = lg[1][s], gz = lg[2][s];
                const R d3 = ri * ri2;
                const R cx = dz * gy - dy * gz, cy = dx * gz - dz * gx, cz = dy * gx - dx * gy;
                a[1] += cx * d3; a[2] += cy * d3; a[3] += cz * d3;
                if (HS) {
                    const R d5 = d3 * ri2, t3 = R(-3) * d5;
                    const R ax = t3 * dx, ay = t3 * dy, az = t3 * dz;
                    const R g3x = gx * d3, g3y = gy * d3, g3z = gz * d3;
                    a[4] += ax * cx;         a[5] += ax * cy + g3z;   a[6] += ax * cz - g3y;
                    a[7] += ay * cx - g3z;   a[8] += ay * cy;         a[9] += ay * cz + g3x;
                    a[10] += az * cx + g3y;  a[11] += az * cy - g3x;  a[12] += az * cz;
                }
            }
        }
    }
    if (!active) return;
    double* o = out + it * 16;
    if (KERNEL == FMMHIP_POINT_SOURCE) {
        o[3] += a[0];          // the reference kernel ignores the switches (test/gravitational.jl:92-98)
        o[4] += a[1]; o[5] += a[2]; o[6] += a[3];
    } else {
        if (GS) { o[4] += a[1]; o[5] += a[2]; o[6] += a[3]; }
        if (HS) for (int k = 4; k < 13; ++k) o[3 + k] += a[k];
    }
    (void)PS;
}

void run_direct_all(Context& C, int kernel, int64_t nt, const double* t_xyz, int64_t ns, const double* s_buf, double* tout_aos,
                    bool PS, bool GS, bool HS, int precision) {
    if (nt == 0 || ns == 0) return;
    const unsigned grid = (unsigned)((nt + P2P_BLOCK - 1) / P2P_BLOCK);
    const int rows = data_per_body(kernel);
    if (kernel == FMMHIP_POINT_SOURCE) {
        if (precision == FMMHIP_FP64) k_direct_all<FMMHIP_POINT_SOURCE, double><<<grid, P2P_BLOCK, 0, C.stream>>>(t_xyz, nt, s_buf, rows, ns, tout_aos, PS, GS, HS);
        else k_direct_all<FMMHIP_POINT_SOURCE, float><<<grid, P2P_BLOCK, 0, C.stream>>>(t_xyz, nt, s_buf, rows, ns, tout_aos, PS, GS, HS);
    } else {
        if (precision == FMMHIP_FP64) k_direct_all<FMMHIP_POINT_VORTEX, double><<<grid, P2P_BLOCK, 0, C.stream>>>(t_xyz, nt, s_buf, rows, ns, tout_aos, PS, GS, HS);
        else k_direct_all<FMMHIP_POINT_VORTEX, float><<<grid, P2P_BLOCK, 0, C.stream>>>(t_xyz, nt, s_buf, rows, ns, tout_aos, PS, GS, HS);
    }
    LAUNCH_CHECK();
}

}  // namespace fmmhip
