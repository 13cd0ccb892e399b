// Mixed-precision kernels: FP32 transcendentals, inverse normal and variance
// lookups; FP64 path accumulators and statistics.
#include <algorithm>

#define PE_MLP_VERIFY_KERNELS 1

#include "pe_kernels.inc"
#include "pe_bs.inc"
#include "pe_lv_risk.inc"
#include "pe_rough.inc"
#include "pe_rough_lsm.inc"
#include "pe_lsm.inc"
#include "pe_mlp.inc"
#define PE_ROUGH_NN_TRAIN_KERNELS
#include "pe_mlp_lsm.inc"
#include "pe_rough_nn.inc"
#include "pe_plan.hpp"

namespace pe {

// Accuracy of the mixed inverse normal against the FP64 AS241 over a range of
// words: per block the largest relative error of z and its word.
__global__ void __launch_bounds__(kBlockThreads) normal_mixed_error_kernel(uint64_t first, uint64_t count,
                                                                           double* block_err, uint32_t* block_word) {
    __shared__ double err[kBlockThreads];
    __shared__ uint32_t at[kBlockThreads];
    double worst = 0.0;
    uint32_t worst_word = 0;
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < count; i += stride) {
        const uint32_t w = static_cast<uint32_t>(first + i);
        const double z = static_cast<double>(normal_from_word_mixed(w));
        const double ref = normal_from_word_f64(w);
        const double e = fabs(z - ref) / fabs(ref);
        if (e > worst || !(e == e)) {
            worst = e == e ? e : 1e300;
            worst_word = w;
        }
    }
    err[threadIdx.x] = worst;
    at[threadIdx.x] = worst_word;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 1; k < kBlockThreads; ++k)
            if (err[k] > worst) {
                worst = err[k];
                worst_word = at[k];
            }
        block_err[blockIdx.x] = worst;
        block_word[blockIdx.x] = worst_word;
    }
}

hipError_t launch_normal_mixed_error(uint64_t first, uint64_t count, uint32_t blocks, double* block_err,
                                     uint32_t* block_word, hipStream_t s) {
    hipLaunchKernelGGL(normal_mixed_error_kernel, dim3(blocks), dim3(kBlockThreads), 0, s, first, count, block_err,
                       block_word);
    return hipGetLastError();
}

hipError_t add_bs_batch_node_f32(hipGraph_t graph, const BsBatchArgs& a, bool risk, bool cont, int kind,
                                 uint32_t graph_regs) {
    return add_bs_batch_node<float>(graph, a, risk, cont, kind, graph_regs);
}

hipError_t launch_bs_batch_f32(const BsBatchArgs& a, bool risk, bool cont, int kind, uint32_t graph_regs,
                               hipStream_t s) {
    return launch_bs_batch<float>(a, risk, cont, kind, graph_regs, s);
}

hipError_t launch_rough_lsm_f32(const RoughLsmArgs& a, dim3 train_grid, dim3 value_grid, hipStream_t s,
                               hipEvent_t paths_done, hipEvent_t between, void* cash_copy) {
    return launch_rough_lsm<float>(a, train_grid, value_grid, s, paths_done, between, cash_copy);
}

hipError_t launch_rough_nn_value_f32(const RoughNnArgs& A, const RoughArgs& a, dim3 grid, hipStream_t s) {
    return launch_rough_nn_value<float>(A, a, grid, s);
}

hipError_t launch_rough_f32(const RoughArgs& a, const RoughPayoffArgs& pa, dim3 grid, uint32_t payoff_items,
                            unsigned long long* error, hipStream_t s) {
    return launch_rough<float>(a, pa, grid, payoff_items, error, s);
}

hipError_t launch_lsm_value_f32(const LsmArgs& value, uint32_t value_items, uint32_t basis, bool risk, hipStream_t s) {
    return launch_lsm_value<float>(value, value_items, basis, risk, s);
}

hipError_t launch_lsm_f32(const LsmArgs& train, uint32_t train_items, const LsmArgs& value, uint32_t value_items,
                          uint32_t basis, bool risk, hipStream_t s, hipEvent_t paths_done, hipEvent_t between) {
    return launch_lsm<float>(train, train_items, value, value_items, basis, risk, s, paths_done, between);
}

hipError_t launch_mlp_gradient(const float* params, const float* x, const float* y, uint32_t rows, uint32_t rows_wg,
                               uint32_t wgs, float* partials, hipStream_t s) {
    const size_t lds = mlp_lds_bytes();
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&mlp_gradient_kernel),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds));
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(mlp_gradient_kernel, dim3(wgs), dim3(kMlpThreads), lds, s, params, x, y, rows, rows_wg,
                       partials);
    return hipGetLastError();
}

hipError_t launch_mlp_forward(const float* params, const float* x, uint32_t rows, float* out, hipStream_t s) {
    const size_t lds = mlp_lds_bytes();
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&mlp_forward_kernel),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds));
    if (e != hipSuccess) return e;
    const uint32_t tiles = (rows + 15) / 16;
    const uint32_t wgs = std::max<uint32_t>(1, std::min<uint32_t>(1024, (tiles + kMlpWaves - 1) / kMlpWaves));
    hipLaunchKernelGGL(mlp_forward_kernel, dim3(wgs), dim3(kMlpThreads), lds, s, params, x, rows, out);
    return hipGetLastError();
}

hipError_t launch_lsm_mlp_f32(const LsmArgs& train, uint32_t train_items, const MlpLsmArgs& m, const LsmArgs& judge,
                              const LsmArgs& value, const LsmArgs& judge_value, uint32_t value_items, bool risk,
                              hipStream_t s, const hipEvent_t* ev) {
    return launch_lsm_mlp<float>(train, train_items, m, judge, value, judge_value, value_items, risk, s, ev);
}

hipError_t launch_lv_graph_f32(const LvArgs& a, dim3 grid, size_t lds, int* err, hipStream_t s) {
    return launch_lv_graph<float>(a, grid, lds, err, s);
}

hipError_t launch_lv_f32(const LvArgs& a, dim3 grid, size_t lds, int* err, hipStream_t s) {
    return launch_lv<float>(a, grid, lds, err, s);
}
hipError_t launch_lv_risk_f32(const LvArgs& a, dim3 grid, size_t lds, int* err, hipStream_t s) {
    return launch_lv_risk<float>(a, grid, lds, err, s);
}

}  // namespace pe
