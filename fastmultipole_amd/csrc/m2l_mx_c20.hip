// Matrix-core M2L kernels at order cap 20 (m2l_mx_gpu.hpp).
#include "m2l_mx_gpu.hpp"

namespace fmmhip {

void run_m2l_mx_c20(Context& C, int p, bool LH, const Lists& Ls, const Tree& st, Tree& tt_dst, const ErrParams& E,
                      const int32_t* perm, int32_t* flags, const TablePtrs& tb) {
    run_m2l_mx_cap<20>(C, p, LH, Ls, st, tt_dst, E, perm, flags, tb);
}

void run_m2l_mx(Context& C, int p, bool LH, Lists& Ls, const Tree& st, Tree& tt_dst, const ErrParams& E, int32_t* flags,
                const TablePtrs& tb) {
    // error-controlled orders vary per interaction: batches of similar order (fixed order: list order)
    const int32_t* perm = E.kind != 0 ? m2l_batch_order(C, Ls, tt_dst, st) : nullptr;
    if (p <= 8) run_m2l_mx_c8(C, p, LH, Ls, st, tt_dst, E, perm, flags, tb);
    else if (p <= 12) run_m2l_mx_c12(C, p, LH, Ls, st, tt_dst, E, perm, flags, tb);
    else if (p <= 16) run_m2l_mx_c16(C, p, LH, Ls, st, tt_dst, E, perm, flags, tb);
    else run_m2l_mx_c20(C, p, LH, Ls, st, tt_dst, E, perm, flags, tb);
}

}  // namespace fmmhip
