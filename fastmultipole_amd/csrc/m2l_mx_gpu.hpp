// Matrix-core M2L on gfx950: the backend of the engine in m2l_mx.hpp for one wavefront and its
// persistent launch. Included by one translation unit per order cap (m2l_mx_c*.hip), so the caps
// compile in parallel.
#pragma once
#include "expansions_common.hpp"
#include "m2l_mx.hpp"
#include <algorithm>
#include <climits>
#include <cstddef>
#include <cstdlib>

namespace fmmhip {

// ================================ M2L on matrix cores, register-resident (m2l_mx.hpp)
// Backend of the engine in m2l_mx.hpp for one gfx950 wavefront: a per-lane value is a register,
// neighbouring lanes are read with DPP (quad permutations, row rotations), other lanes with
// ds_bpermute (__shfl), and the product is v_mfma_f64_16x16x4f64.
template <int CTRL>
__device__ __forceinline__ double dpp_d(double x) {
    const long long u = __double_as_longlong(x);
    const int lo = __builtin_amdgcn_update_dpp((int)u, (int)u, CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp((int)(u >> 32), (int)(u >> 32), CTRL, 0xF, 0xF, false);
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}
static constexpr int DPP_QUAD(int a, int b, int c, int d) { return a | (b << 2) | (c << 4) | (d << 6); }
static constexpr int DPP_ROW_ROR(int n) { return 0x120 + n; }

struct MxGpu {
    using V = double;
    using I = int;
    using M = bool;
    typedef double V4 __attribute__((ext_vector_type(4)));
    using PV = const double*;
    static __device__ __forceinline__ I lane() { return (int)(threadIdx.x & 63); }
    static __device__ __forceinline__ V sel(M m, V a, V b) { return m ? a : b; }
    static __device__ __forceinline__ I seli(M m, I a, I b) { return m ? a : b; }
    static __device__ __forceinline__ I mini(I a, int b) { return min(a, b); }
    static __device__ __forceinline__ void sched_fence() { __builtin_amdgcn_sched_barrier(0); }
    static __device__ __forceinline__ I opaque(I a) {   // a value the compiler cannot move out of its scope
        asm volatile("" : "+v"(a));
        return a;
    }
    static __device__ __forceinline__ V fabs(V a) { return ::fabs(a); }
    static __device__ __forceinline__ V sqrt(V a) { return ::sqrt(a); }
    static __device__ __forceinline__ V fma(V a, V b, V c) { return ::fma(a, b, c); }
    static __device__ __forceinline__ V tod(I a) { return (double)a; }
    static __device__ __forceinline__ V perm1(V a) { return dpp_d<DPP_QUAD(1, 0, 3, 2)>(a); }
    static __device__ __forceinline__ V perm2(V a) { return dpp_d<DPP_QUAD(2, 3, 0, 1)>(a); }
    static __device__ __forceinline__ V perm3(V a) { return dpp_d<DPP_QUAD(3, 2, 1, 0)>(a); }
    static __device__ __forceinline__ V bperm(V a, I src) { return __shfl(a, src, 64); }
    static __device__ __forceinline__ V quadsum(V a) {
        const V b = a + __shfl_xor(a, 16, 64);
        return b + __shfl_xor(b, 32, 64);
    }
    template <int NC>
    static __device__ __forceinline__ V slotsum(V a) {   // sum over lanes of the row with the same lane % NC
        V b = a + dpp_d<DPP_ROW_ROR(8)>(a);
        b = b + dpp_d<DPP_ROW_ROR(4)>(b);
        if constexpr (NC == 2) b = b + dpp_d<DPP_ROW_ROR(2)>(b);
        return b;
    }
    template <int NC, int j>
    static __device__ __forceinline__ V qbcast(V a) {
        if constexpr (NC == 4) return dpp_d<DPP_QUAD(j, j, j, j)>(a);
        else return dpp_d<DPP_QUAD(j, j, 2 + j, 2 + j)>(a);
    }
    static __device__ __forceinline__ V4 zero4() { return V4{0.0, 0.0, 0.0, 0.0}; }
    static __device__ __forceinline__ V get4(const V4& a, int r) { return a[r]; }
    static __device__ __forceinline__ void set4(V4& a, int r, V x) { a[r] = x; }
    static __device__ __forceinline__ void mfma(V a, V b, V4& d) { d = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, d, 0, 0, 0); }
    static __device__ __forceinline__ bool any(M m) { return __ballot(m) != 0ull; }
    static __device__ __forceinline__ bool all(M m) { return __ballot(!m) == 0ull; }
    static __device__ __forceinline__ int maxi(I a, M m) {
        int v = m ? a : INT_MIN;
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
        return v;
    }
    static __device__ __forceinline__ V gld(const double* p, I i, M ok) {
        V r = 0.0;
        if (ok) r = p[i];
        return r;
    }
    static __device__ __forceinline__ I gldi(const int32_t* p, I i, M ok) {
        I r = 0;
        if (ok) r = p[i];
        return r;
    }
    static __device__ __forceinline__ PV ptr(const double* base, I js, int64_t stride) { return base + (int64_t)js * stride; }
    static __device__ __forceinline__ V gldp(PV p, I i, M ok) {
        V r = 0.0;
        if (ok) r = p[i];
        return r;
    }
    static __device__ __forceinline__ V lds(const double* p, I i) { return p[i]; }
    static __device__ __forceinline__ void lds_st(double* p, I i, V x, M ok) {
        if (ok) p[i] = x;
    }
    static __device__ __forceinline__ void lds_add(double* p, I i, V x, M ok) {
        if (ok) p[i] += x;
    }
    static __device__ __forceinline__ void sti(int32_t* p, I i, I x, M ok) {
        if (ok) p[i] = x;
    }
    static __device__ __forceinline__ void wsync() { wave_sync(); }
};

template <int CAP, bool LH>
__host__ __device__ constexpr int mx_wave_doubles() {
    return (LH ? 4 : 2) * nh(CAP) + 4 * (LH ? 4 : 8) * (CAP + 1) + (LH ? 4 : 8) * (2 * CAP + 1);
}
template <int CAP>
__host__ __device__ constexpr int mx_shared_doubles() {
    return mx::ALEN(CAP) + 6 * nh(CAP) + 4 * (CAP + 1);
}
template <int CAP, bool LH>
static size_t mx_lds_bytes(int nw) {
    return sizeof(double) * (size_t)(mx_shared_doubles<CAP>() + nw * mx_wave_doubles<CAP, LH>());
}

// Persistent: every workgroup copies the tables into LDS once, then each wavefront takes whole target
// boxes from a counter (flags[2]) and runs their interaction lists BP at a time; a box's local
// expansion is summed in the wavefront's LDS and added to the target's expansion once, by the one
// wavefront that owns the box (deterministic: the same sums in the same order on every run).
template <int CAP, bool LH, int NW, bool FULL>
__global__ __launch_bounds__(64 * NW) void k_m2l_mx(const int32_t* __restrict__ m2l_targets,
                                                    const int64_t* __restrict__ m2l_offsets,
                                                    const int32_t* __restrict__ m2l_s, const Branch* __restrict__ tbr,
                                                    const Branch* __restrict__ sbr, const double* __restrict__ sex,
                                                    double* __restrict__ tex, int n_boxes, int P, int PE, ErrParams E,
                                                    const int32_t* __restrict__ perm, int32_t* __restrict__ orders,
                                                    int32_t* __restrict__ flags, int32_t* __restrict__ box_ok,
                                                    TablePtrs tb) {
    using Eng = mx::Engine<MxGpu, CAP, LH, FULL>;
    constexpr int H = nh(CAP), NC = Eng::NC, BP = Eng::BP, NS = CAP + 1;
    extern __shared__ double lds[];
    double* A = lds;
    double* zin = A + mx::ALEN(CAP);
    double* zoh = zin + H;
    double* ein = zoh + H;
    double* eoh = ein + H;
    double* Mt = eoh + H;
    double* Lt = Mt + H;
    double* tf = Lt + H;
    for (int i = threadIdx.x; i < mx::ALEN(CAP); i += blockDim.x) A[i] = tb.Ad[i];
    for (int i = threadIdx.x; i < H; i += blockDim.x) {
        int n, m;
        idx_nm(i, n, m);
        const int b = len_Ts(n - 1);
        zin[i] = tb.zeta[b + m];
        zoh[i] = 0.5 * tb.zeta[b + m * (n + 1)];
        ein[i] = tb.eta[b + m];
        eoh[i] = 0.5 * tb.eta[b + m * (n + 1)];
        Mt[i] = tb.Mt[i];
        Lt[i] = tb.Lt[i];
    }
    double* i0 = tf + 2 * NS;
    for (int i = threadIdx.x; i < 2 * NS; i += blockDim.x) tf[i] = tb.tfac[i];
    for (int n = threadIdx.x; n < NS; n += blockDim.x) {
        i0[2 * n] = 1.0 / tb.Mt[nh(n - 1)];
        i0[2 * n + 1] = 1.0 / tb.Lt[nh(n - 1)];
    }
    __syncthreads();
    double* wl = lds + mx_shared_doubles<CAP>() + (size_t)wave_id() * mx_wave_doubles<CAP, LH>();
    const mx::Wave W{wl + NC * H, wl + NC * H + 2 * BP * NS, wl, wl + NC * H + 4 * BP * NS};
    const mx::Tabs T{A, zin, zoh, ein, eoh, Mt, Lt, tf, i0};
    const int64_t HE4 = 4 * (int64_t)nh(PE);
    const int lane = lane_id();
    int pmax = 0, okf = 1;
    for (;;) {
        int box = 0;
        if (lane == 0) box = atomicAdd(flags + 2, 1);
        box = __shfl(box, 0, 64);
        if (box >= n_boxes) break;
        for (int i = lane; i < NC * H; i += 64) W.acc[i] = 0.0;
        wave_sync();
        const int32_t it = m2l_targets[box];
        const Branch& Tb = tbr[it];
        mx::Box X;
        for (int d = 0; d < 3; ++d) {
            X.tc[d] = Tb.target_center[d];
            X.tbox[d] = Tb.target_box[d];
        }
        X.r_l = Tb.target_radius;
        const double eps = E.relative ? Tb.max_influence * E.eps : E.eps;
        X.limit = eps * 4 * PI_D;
        X.P = P;
        X.kind = E.kind;
        X.BE = E.BE;
        const int64_t l0 = m2l_offsets[box], l1 = m2l_offsets[box + 1];
        int okb = 1;
        for (int64_t b0 = l0; b0 < l1; b0 += BP) {
            const mx::Batch Bt{m2l_s, perm, (int)b0, (int)min((int64_t)BP, l1 - b0), orders};
            Eng::batch(T, W, X, Bt, (const double*)sbr, (int)(sizeof(Branch) / 8), (int)(offsetof(Branch, source_center) / 8),
                       sex, HE4, pmax, okb);
        }
        if (!okb) {
            okf = 0;
            if (box_ok && lane == 0) box_ok[box] = 0;
        }
        double* dst = tex + (int64_t)it * HE4;
        for (int i = lane; i < NC * nh(P); i += 64) dst[4 * (i / NC) + i % NC] += W.acc[i];
        wave_sync();
    }
    if (lane == 0) {
        atomicMax(flags, pmax);
        if (!okf) atomicAnd(flags + 1, 0);
    }
}

template <int CAP, bool LH, int NW, bool FULL>
static void launch_m2l_mx(Context& C, int p, int pe, const Lists& Ls, const Tree& st, Tree& tt_dst, const ErrParams& E,
                          const int32_t* perm, int32_t* flags, const TablePtrs& tb) {
    auto kfn = k_m2l_mx<CAP, LH, NW, FULL>;
    const size_t sh = mx_lds_bytes<CAP, LH>(NW);
    HIP_CHECK(hipFuncSetAttribute((const void*)kfn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh));
    int per_cu = 0;
    HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void*>(kfn), 64 * NW, sh));
    if (!C.n_cu) HIP_CHECK(hipDeviceGetAttribute(&C.n_cu, hipDeviceAttributeMultiprocessorCount, C.device));
    const int64_t want = (Ls.n_m2l_targets + NW - 1) / NW;
    const unsigned grid = (unsigned)std::max<int64_t>(1, std::min<int64_t>(want, (int64_t)std::max(per_cu, 1) * C.n_cu));
    kfn<<<grid, 64 * NW, sh, C.stream>>>(Ls.m2l_targets.data(), Ls.m2l_offsets.data(), Ls.m2l_s.data(), tt_dst.branches.data(),
                                         st.branches.data(), st.expansions.data(), tt_dst.expansions.data(),
                                         (int)Ls.n_m2l_targets, p, pe, E, perm, Ls.m2l_order.data(), flags,
                                         Ls.track_box_ok ? const_cast<int32_t*>(Ls.m2l_box_ok.data()) : nullptr, tb);
    LAUNCH_CHECK();
}


static constexpr int MX_NW = 4;   // wavefronts per workgroup (one per SIMD at the register use of the engine)

// the launch at order p <= CAP (CAP the smallest instantiated cap >= p)
template <int CAP>
void run_m2l_mx_cap(Context& C, int p, bool LH, const Lists& Ls, const Tree& st, Tree& tt_dst, const ErrParams& E,
                    const int32_t* perm, int32_t* flags, const TablePtrs& tb) {
    if (p == CAP) {   // degree loops of fixed length (34.1 against 34.4 ms on the 1M ring at cap 20)
        if (LH) launch_m2l_mx<CAP, true, MX_NW, true>(C, p, p, Ls, st, tt_dst, E, perm, flags, tb);
        else launch_m2l_mx<CAP, false, MX_NW, true>(C, p, p, Ls, st, tt_dst, E, perm, flags, tb);
    } else {
        if (LH) launch_m2l_mx<CAP, true, MX_NW, false>(C, p, p, Ls, st, tt_dst, E, perm, flags, tb);
        else launch_m2l_mx<CAP, false, MX_NW, false>(C, p, p, Ls, st, tt_dst, E, perm, flags, tb);
    }
}

}  // namespace fmmhip
