// Host side of the device Dupire build: validation, the theta PCHIP knot
// derivatives (crates/pricing/src/market/pchip.rs:27-238) and the C ABI entry.
// Also unpacks the embedded Joe--Kuo direction asset.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "pe.h"
#include "pe_plan.hpp"

extern "C" const unsigned char pe_jk_directions_begin[];
extern "C" const unsigned char pe_jk_directions_end[];

namespace pe {

namespace {
double signum(double x) { return std::isnan(x) ? x : std::copysign(1.0, x); }

double endpoint_derivative(double aw, double nw, double as, double ns) {
    double d = ((2.0 * aw + nw) * as - aw * ns) / (aw + nw);
    if (signum(d) != signum(as)) d = 0.0;
    else if (signum(as) != signum(ns) && std::fabs(d) > 3.0 * std::fabs(as)) d = 3.0 * as;
    return d;
}
}  // namespace

std::vector<double> pchip_derivatives(const std::vector<double>& t, const std::vector<double>& v) {
    const size_t n = t.size();
    std::vector<double> w(n - 1), s(n - 1), d(n, 0.0);
    for (size_t i = 0; i + 1 < n; ++i) {
        w[i] = t[i + 1] - t[i];
        s[i] = (v[i + 1] - v[i]) / w[i];
    }
    if (n == 2) return {s[0], s[0]};
    d[0] = endpoint_derivative(w[0], w[1], s[0], s[1]);
    for (size_t i = 1; i + 1 < n; ++i) {
        const double l = s[i - 1], r = s[i];
        if (l == 0.0 || r == 0.0 || signum(l) != signum(r)) {
            d[i] = 0.0;
        } else {
            const double w1 = 2.0 * w[i] + w[i - 1];
            const double w2 = w[i] + 2.0 * w[i - 1];
            d[i] = (w1 + w2) / (w1 / l + w2 / r);
        }
    }
    d[n - 1] = endpoint_derivative(w[n - 2], w[n - 3], s[n - 2], s[n - 3]);
    return d;
}

const uint32_t* direction_words(uint32_t* dimensions) {
    static std::once_flag once;
    static std::vector<uint32_t> words;
    static uint32_t dims = 0;
    std::call_once(once, [] {
        const unsigned char* b = pe_jk_directions_begin;
        const size_t size = static_cast<size_t>(pe_jk_directions_end - pe_jk_directions_begin);
        auto be = [b](size_t off) {
            return (uint32_t(b[off]) << 24) | (uint32_t(b[off + 1]) << 16) | (uint32_t(b[off + 2]) << 8) | uint32_t(b[off + 3]);
        };
        if (size < 20 || std::memcmp(b, "JK621201", 8) != 0 || be(16) != 32) return;
        const uint32_t d = be(12);
        if (size != 20 + static_cast<size_t>(d) * 128) return;
        words.resize(static_cast<size_t>(d) * 32);
        for (size_t i = 0; i < words.size(); ++i) words[i] = be(20 + 4 * i);
        dims = d;
    });
    *dimensions = dims;
    return words.data();
}

}  // namespace pe

using namespace pe;

namespace {
thread_local std::string g_surface_error;
}

extern "C" pe_status pe_local_variance_from_ssvi(const pe_ssvi_surface* s, uint32_t nt, const double* t, uint32_t nx,
                                                 const double* x, double floor, double cap, double* values_out,
                                                 uint32_t* repair_count_out) {
    if (!s || !t || !x || !values_out) return PE_ERR_INVALID_ARGUMENT;
    // validate_nodes / validate_floor_cap (local_variance.rs:510-565); SSVI parameters (ssvi.rs:78-120).
    auto sorted = [](const double* v, uint32_t n, bool nonneg) {
        if (n < 2) return false;
        for (uint32_t i = 0; i < n; ++i) {
            if (!std::isfinite(v[i]) || (nonneg && v[i] < 0.0)) return false;
            if (i && !(v[i] > v[i - 1])) return false;
        }
        return true;
    };
    if (!sorted(t, nt, true) || !sorted(x, nx, false)) return PE_ERR_INVALID_ARGUMENT;
    if (!std::isfinite(floor) || floor <= 0.0 || !std::isfinite(cap) || cap < floor) return PE_ERR_INVALID_ARGUMENT;
    if (s->theta_count < 2 || !sorted(s->theta_times, s->theta_count, true) || s->theta_times[0] <= 0.0)
        return PE_ERR_INVALID_ARGUMENT;
    for (uint32_t i = 0; i < s->theta_count; ++i)
        if (!(s->theta_values[i] > 0.0) || (i && s->theta_values[i] < s->theta_values[i - 1]))
            return PE_ERR_INVALID_ARGUMENT;
    if (!std::isfinite(s->rho) || std::fabs(s->rho) >= 1.0 || !(s->eta > 0.0) || !(s->gamma > 0.0) ||
        !(s->gamma < 1.0) || !std::isfinite(s->theta_terminal_slope) || s->theta_terminal_slope < 0.0)
        return PE_ERR_INVALID_ARGUMENT;
    hipStream_t stream = nullptr;
    if (pe_init(0) != PE_OK) return PE_ERR_DEVICE;
    try {
        std::vector<double> kt(s->theta_times, s->theta_times + s->theta_count);
        std::vector<double> kv(s->theta_values, s->theta_values + s->theta_count);
        std::vector<double> kd = pchip_derivatives(kt, kv);
        DeviceBuffer dkt, dkv, dkd, dt, dx, dout, drep, dstatus;
        dkt.upload(kt);
        dkv.upload(kv);
        dkd.upload(kd);
        dt.upload(std::vector<double>(t, t + nt));
        dx.upload(std::vector<double>(x, x + nx));
        dout.reserve(sizeof(double) * nt * nx);
        const std::vector<uint32_t> zero{0u, 0u};
        drep.upload(zero);
        dstatus.upload(zero);
        if (launch_dupire(dkt.as<double>(), dkv.as<double>(), dkd.as<double>(), s->theta_count,
                          s->theta_terminal_slope, s->rho, s->eta, s->gamma, dt.as<double>(), nt, dx.as<double>(), nx,
                          floor, cap, dout.as<double>(), drep.as<uint32_t>(), dstatus.as<int>(), stream) != hipSuccess)
            return PE_ERR_DEVICE;
        if (hipDeviceSynchronize() != hipSuccess) return PE_ERR_DEVICE;
        int status = 0;
        uint32_t repairs = 0;
        if (hipMemcpy(values_out, dout.ptr, sizeof(double) * nt * nx, hipMemcpyDeviceToHost) != hipSuccess ||
            hipMemcpy(&status, dstatus.ptr, sizeof(int), hipMemcpyDeviceToHost) != hipSuccess ||
            hipMemcpy(&repairs, drep.ptr, sizeof(uint32_t), hipMemcpyDeviceToHost) != hipSuccess)
            return PE_ERR_DEVICE;
        if (repair_count_out) *repair_count_out = repairs;
        return status == 0 ? PE_OK : PE_ERR_INVALID_ARGUMENT;
    } catch (const PlanError& e) {
        return e.status;
    }
}
