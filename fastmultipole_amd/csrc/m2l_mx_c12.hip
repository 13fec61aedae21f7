// Matrix-core M2L kernels at order cap 12 (m2l_mx_gpu.hpp).
#include "m2l_mx_gpu.hpp"

namespace fmmhip {

void run_m2l_mx_c12(Context& C, int p, bool LH, const Lists& Ls, const Tree& st, Tree& tt_dst, const ErrParams& E,
                      const int32_t* perm, int32_t* flags, const TablePtrs& tb) {
    run_m2l_mx_cap<12>(C, p, LH, Ls, st, tt_dst, E, perm, flags, tb);
}

}  // namespace fmmhip
