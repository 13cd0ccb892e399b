// Dupire local variance from a StandardSsvi surface, one lane per grid node,
// in FP64 on the device.
//
// LocalVarianceGrid::from_surface (crates/pricing/src/market/local_variance.rs:149-206),
// StandardSsvi::total_variance_derivatives (crates/pricing/src/market/ssvi.rs:420-475),
// PhiSpec::PowerLaw (ssvi.rs:123-136), durrleman_density_factor (ssvi.rs:226-275),
// ThetaPchip::evaluate (crates/pricing/src/market/pchip.rs:111-170).
#include <hip/hip_runtime.h>

#include "pe_plan.hpp"

namespace pe {

namespace {

__device__ void theta_eval(const double* kt, const double* kv, const double* kd, uint32_t n, double slope, double t,
                           double* theta, double* deriv) {
    if (t < kt[0]) {
        const double d = kv[0] / kt[0];
        *theta = t * d;
        *deriv = d;
        return;
    }
    uint32_t right = 0;
    while (right < n && kt[right] < t) ++right;
    if (right < n && kt[right] == t) {
        *theta = kv[right];
        *deriv = kd[right];
        return;
    }
    if (right < n) {
        const uint32_t left = right - 1;
        const double w = kt[right] - kt[left];
        const double s = (t - kt[left]) / w;
        const double s2 = s * s, s3 = s2 * s;
        const double h00 = 2.0 * s3 - 3.0 * s2 + 1.0;
        const double h10 = s3 - 2.0 * s2 + s;
        const double h01 = -2.0 * s3 + 3.0 * s2;
        const double h11 = s3 - s2;
        *theta = h00 * kv[left] + h10 * w * kd[left] + h01 * kv[right] + h11 * w * kd[right];
        const double d00 = (6.0 * s2 - 6.0 * s) / w;
        const double d10 = 3.0 * s2 - 4.0 * s + 1.0;
        const double d01 = (-6.0 * s2 + 6.0 * s) / w;
        const double d11 = 3.0 * s2 - 2.0 * s;
        *deriv = d00 * kv[left] + d10 * kd[left] + d01 * kv[right] + d11 * kd[right];
        return;
    }
    *theta = kv[n - 1] + slope * (t - kt[n - 1]);
    *deriv = slope;
}

// status: 0 ok; 1 non-positive total variance; 2 non-positive density factor;
// 3 non-finite time derivative or phi.
__global__ void dupire_kernel(const double* kt, const double* kv, const double* kd, uint32_t nk, double slope,
                              double rho, double eta, double gamma, const double* t_nodes, uint32_t nt,
                              const double* x_nodes, uint32_t nx, double floor, double cap, double* out,
                              uint32_t* repairs, int* status) {
    const uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= nt * nx) return;
    const uint32_t i = idx / nx, j = idx % nx;
    const double t = t_nodes[i] == 0.0 ? t_nodes[1] : t_nodes[i];
    const double k = x_nodes[j];
    double theta, theta_d;
    theta_eval(kt, kv, kd, nk, slope, t, &theta, &theta_d);
    const double phi = eta / (pow(theta, gamma) * pow(1.0 + theta, 1.0 - gamma));
    const double phi_d = -phi * (gamma / theta + (1.0 - gamma) / (1.0 + theta));
    if (!(theta > 0.0) || !isfinite(phi) || !(phi > 0.0) || !isfinite(phi_d)) {
        atomicMax(status, 3);
        out[idx] = floor;
        return;
    }
    const double y = phi * k;
    const double a = y + rho;
    const double q = sqrt(a * a + 1.0 - rho * rho);
    const double shape = 1.0 + rho * y + q;
    const double w = theta * shape / 2.0;
    const double common = rho + a / q;
    const double wk = theta * phi * common / 2.0;
    const double wkk = theta * phi * phi * (1.0 - rho * rho) / (2.0 * q * q * q);
    const double wt = (shape / 2.0 + theta * k * phi_d * common / 2.0) * theta_d;
    if (!isfinite(wt)) { atomicMax(status, 3); out[idx] = floor; return; }
    if (!(w > 0.0) || !isfinite(w)) { atomicMax(status, 1); out[idx] = floor; return; }
    const double u = 1.0 - k * wk / (2.0 * w);
    const double g = u * u - (wk * wk / 4.0) * (1.0 / w + 1.0 / 4.0) + wkk / 2.0;
    if (!isfinite(g) || g <= 0.0) { atomicMax(status, 2); out[idx] = floor; return; }
    const double raw = wt / g;
    double v = raw;
    bool repaired = true;
    if (isnan(raw)) v = floor;
    else if (raw == -INFINITY) v = floor;
    else if (raw == INFINITY) v = cap;
    else if (raw <= 0.0) v = floor;
    else if (raw < floor) v = floor;
    else if (raw > cap) v = cap;
    else repaired = false;
    if (repaired) atomicAdd(repairs, 1u);
    out[idx] = v;
}

}  // namespace

hipError_t launch_dupire(const double* kt, const double* kv, const double* kd, uint32_t nk, double slope, double rho,
                         double eta, double gamma, const double* t, uint32_t nt, const double* x, uint32_t nx,
                         double floor, double cap, double* out, uint32_t* repairs, int* status, hipStream_t s) {
    const uint32_t n = nt * nx;
    const uint32_t blocks = (n + 255u) / 256u;
    hipLaunchKernelGGL(dupire_kernel, dim3(blocks), dim3(256), 0, s, kt, kv, kd, nk, slope, rho, eta, gamma, t, nt, x,
                       nx, floor, cap, out, repairs, status);
    return hipGetLastError();
}

}  // namespace pe
