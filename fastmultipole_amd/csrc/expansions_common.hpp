// Definitions shared by the translation kernels of expansions.hip and the matrix-core M2L units
// (m2l_mx_c*.hip): expansion indexing, wavefront helpers, the device table pointers and the error
// method of the M2L.
#pragma once
#include "fmmhip_internal.hpp"

namespace fmmhip {

static constexpr double PI_D = 3.141592653589793;

__host__ __device__ constexpr int nh(int p) { return ((p + 1) * (p + 2)) >> 1; }
__host__ __device__ constexpr int len_Hs(int p) { return ((p + 1) * (p + 2) * (p + 3)) / 6; }
__host__ __device__ constexpr int len_Ts(int p) { return p < 0 ? 0 : ((p + 1) * (p + 2) * (2 * p + 3)) / 6; }

// idx -> (n, m) for the packed (n, m>=0) ordering
__device__ inline void idx_nm(int idx, int& n, int& m) {
    int k = (int)((sqrtf(8.0f * (float)idx + 1.0f) - 1.0f) * 0.5f);
    while (((k + 1) * (k + 2)) / 2 <= idx) ++k;
    while ((k * (k + 1)) / 2 > idx) --k;
    n = k;
    m = idx - (k * (k + 1)) / 2;
}

__device__ inline int lane_id() { return (int)(threadIdx.x & 63); }
__device__ inline int wave_id() { return (int)(threadIdx.x >> 6); }

// LDS ordering between lanes of one wavefront
__device__ inline void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// device copies of the rotation and normalisation tables (built once per context, order 30)
struct TablePtrs {
    const double *Hs, *zeta, *eta, *Mt, *Lt, *inv_fact, *fact, *Ad, *tfac;
};

// error method of an M2L launch
struct ErrParams {
    int kind;        // 0 fixed order, 1 rotated coefficients (gradient), 2 power potential, 3 power gradient
    int relative;
    double eps;
    int BE;
};

// processing order of the error-controlled M2L list within each target box (m2l_mx_order.hip)
const int32_t* m2l_batch_order(Context& C, Lists& Ls, const Tree& tt, const Tree& st);
// matrix-core M2L (m2l_mx_c*.hip): one launch over the M2L list at order (cap) p <= 20
#define FMMHIP_MX_DECL(name)                                                                                           \
    void name(Context& C, int p, bool LH, const Lists& Ls, const Tree& st, Tree& tt_dst, const ErrParams& E,         \
              const int32_t* perm, int32_t* flags, const TablePtrs& tb);
FMMHIP_MX_DECL(run_m2l_mx_c8)
FMMHIP_MX_DECL(run_m2l_mx_c12)
FMMHIP_MX_DECL(run_m2l_mx_c16)
FMMHIP_MX_DECL(run_m2l_mx_c20)
#undef FMMHIP_MX_DECL
void run_m2l_mx(Context& C, int p, bool LH, Lists& Ls, const Tree& st, Tree& tt_dst, const ErrParams& E, int32_t* flags,
                const TablePtrs& tb);

}  // namespace fmmhip
