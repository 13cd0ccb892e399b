// FP64 parity-mode kernels.  Built with -ffp-contract=off so every expression
// keeps the reference's rounding sequence (no fused multiply-add).
#include "pe_kernels.inc"
#include "pe_plan.hpp"

namespace pe {

hipError_t launch_bs_f64(const BsArgs& a, bool risk, dim3 grid, hipStream_t s) {
    if (risk) hipLaunchKernelGGL((bs_kernel<double, true>), grid, dim3(kBlockThreads), 0, s, a);
    else hipLaunchKernelGGL((bs_kernel<double, false>), grid, dim3(kBlockThreads), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_lv_f64(const LvArgs& a, dim3 grid, size_t lds, int* err, hipStream_t s) {
    return launch_lv<double>(a, grid, lds, err, s);
}

}  // namespace pe
