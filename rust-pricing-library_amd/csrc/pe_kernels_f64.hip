// FP64 parity-mode kernels.  Built with -ffp-contract=off so every expression
// keeps the reference's rounding sequence (no fused multiply-add).
#include "pe_kernels.inc"
#include "pe_bs.inc"
#include "pe_lv_risk.inc"
#include "pe_rough.inc"
#include "pe_rough_lsm.inc"
#include "pe_lsm.inc"
#include "pe_lsm_dual.inc"
#include "pe_mlp.inc"
#include "pe_mlp_lsm.inc"
#include "pe_rough_nn.inc"
#include "pe_plan.hpp"

namespace pe {

hipError_t add_bs_batch_node_f64(hipGraph_t graph, const BsBatchArgs& a, bool risk, bool cont, int kind,
                                 uint32_t graph_regs) {
    return add_bs_batch_node<double>(graph, a, risk, cont, kind, graph_regs);
}

hipError_t launch_bs_batch_f64(const BsBatchArgs& a, bool risk, bool cont, int kind, uint32_t graph_regs,
                               hipStream_t s) {
    return launch_bs_batch<double>(a, risk, cont, kind, graph_regs, s);
}

hipError_t launch_rough_lsm_f64(const RoughLsmArgs& a, dim3 train_grid, dim3 value_grid, hipStream_t s,
                               hipEvent_t paths_done, hipEvent_t between, void* cash_copy) {
    return launch_rough_lsm<double>(a, train_grid, value_grid, s, paths_done, between, cash_copy);
}

hipError_t launch_rough_nn_value_f64(const RoughNnArgs& A, const RoughArgs& a, dim3 grid, hipStream_t s) {
    return launch_rough_nn_value<double>(A, a, grid, s);
}

hipError_t launch_rough_f64(const RoughArgs& a, const RoughPayoffArgs& pa, dim3 grid, uint32_t payoff_items,
                            unsigned long long* error, hipStream_t s) {
    return launch_rough<double>(a, pa, grid, payoff_items, error, s);
}

hipError_t launch_lsm_f64(const LsmArgs& train, uint32_t train_items, const LsmArgs& value, uint32_t value_items,
                          uint32_t basis, bool risk, hipStream_t s, hipEvent_t paths_done, hipEvent_t between) {
    return launch_lsm<double>(train, train_items, value, value_items, basis, risk, s, paths_done, between);
}

hipError_t launch_lsm_mlp_f64(const LsmArgs& train, uint32_t train_items, const MlpLsmArgs& m, const LsmArgs& judge,
                              const LsmArgs& value, const LsmArgs& judge_value, uint32_t value_items, bool risk,
                              hipStream_t s, const hipEvent_t* ev) {
    return launch_lsm_mlp<double>(train, train_items, m, judge, value, judge_value, value_items, risk, s, ev);
}

hipError_t launch_lv_graph_f64(const LvArgs& a, dim3 grid, size_t lds, int* err, hipStream_t s) {
    return launch_lv_graph<double>(a, grid, lds, err, s);
}

hipError_t launch_lv_f64(const LvArgs& a, dim3 grid, size_t lds, int* err, hipStream_t s) {
    return launch_lv<double>(a, grid, lds, err, s);
}
hipError_t launch_lsm_dual_f64(const LsmDualArgs& da, uint64_t pairs, uint32_t outer_blocks, uint32_t basis,
                               hipStream_t s) {
    return launch_lsm_dual<double>(da, pairs, outer_blocks, basis, s);
}
hipError_t launch_lv_risk_f64(const LvArgs& a, dim3 grid, size_t lds, int* err, hipStream_t s) {
    return launch_lv_risk<double>(a, grid, lds, err, s);
}

}  // namespace pe
