// Mixed-precision kernels: FP32 transcendentals, inverse normal and variance
// lookups; FP64 path accumulators and statistics.
#include "pe_kernels.inc"
#include "pe_plan.hpp"

namespace pe {

hipError_t launch_bs_f32(const BsArgs& a, bool risk, dim3 grid, hipStream_t s) {
    if (risk) hipLaunchKernelGGL((bs_kernel<float, true>), grid, dim3(kBlockThreads), 0, s, a);
    else hipLaunchKernelGGL((bs_kernel<float, false>), grid, dim3(kBlockThreads), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_lv_f32(const LvArgs& a, dim3 grid, size_t lds, int* err, hipStream_t s) {
    return launch_lv<float>(a, grid, lds, err, s);
}

}  // namespace pe
