// synth.hip -- the synthetic batched-closure workload on the device and on the
// host (same generator code, see synth.h), plus plain device-memory helpers of
// the C ABI for callers without their own allocator.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <thread>
#include <vector>

#include "synth.h"
#include "uqk.h"

namespace uqk {

__global__ void __launch_bounds__(256) synth_kernel(SynthParams p, const double* phi, const double* phiw,
                                                    uint64_t seed, long ncell, double* lam, double* u,
                                                    double* lam0, unsigned long long* nfail) {
    long cell = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (cell >= ncell) return;
    const long off = cell * 3 * p.M;
    int att = synth_cell(p, phi, phiw, seed, (uint64_t)cell, lam + off, u + off, lam0 + off);
    if (att > p.max_attempts) atomicAdd(nfail, 1ull);
}

static SynthParams synth_params(int M, int Q) {
    SynthParams p;
    p.M = M;
    p.Q = Q;
    p.gamma = 1.4;
    p.sigma = 0.05;
    p.temp_ratio = 1.0 / 3.0;
    p.rho_ratio = 100.0;
    p.max_attempts = 1 << 20;
    return p;
}

}  // namespace uqk

extern "C" {

int64_t uqk_synth_euler1d_closure_host(int64_t ncell, uint64_t seed, int32_t M, int32_t Q, const double* phi,
                                       const double* phiw, double* u, double* lam_star, double* lam0) {
    const uqk::SynthParams p = uqk::synth_params(M, Q);
    unsigned nth = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<std::thread> pool;
    std::vector<int64_t> fails(nth, 0);
    for (unsigned t = 0; t < nth; ++t) {
        pool.emplace_back([&, t]() {
            for (int64_t j = t; j < ncell; j += nth) {
                const int64_t off = j * 3 * M;
                int att = uqk::synth_cell(p, phi, phiw, seed, (uint64_t)j, lam_star + off, u + off, lam0 + off);
                if (att > p.max_attempts) ++fails[t];
            }
        });
    }
    for (auto& th : pool) th.join();
    int64_t f = 0;
    for (auto x : fails) f += x;
    return f;
}

int64_t uqk_synth_euler1d_closure(int64_t ncell, uint64_t seed, int32_t M, int32_t Q, const double* phi,
                                  const double* phiw, double* u, double* lam_star, double* lam0) {
    const uqk::SynthParams p = uqk::synth_params(M, Q);
    double *dphi = nullptr, *dphiw = nullptr;
    unsigned long long* dfail = nullptr;
    const size_t tb = sizeof(double) * (size_t)M * Q;
    if (hipMalloc(&dphi, tb) != hipSuccess || hipMalloc(&dphiw, tb) != hipSuccess ||
        hipMalloc(&dfail, sizeof(unsigned long long)) != hipSuccess) {
        hipFree(dphi);
        hipFree(dphiw);
        hipFree(dfail);
        return -2;
    }
    hipMemcpy(dphi, phi, tb, hipMemcpyHostToDevice);
    hipMemcpy(dphiw, phiw, tb, hipMemcpyHostToDevice);
    hipMemset(dfail, 0, sizeof(unsigned long long));
    int64_t result = 0;
    if (ncell > 0) {
        int blocks = (int)((ncell + 255) / 256);
        hipLaunchKernelGGL(uqk::synth_kernel, dim3(blocks), dim3(256), 0, 0, p, dphi, dphiw, seed, (long)ncell,
                           lam_star, u, lam0, dfail);
        if (hipGetLastError() != hipSuccess) result = -2;
    }
    unsigned long long nf = 0;
    if (hipMemcpy(&nf, dfail, sizeof(nf), hipMemcpyDeviceToHost) != hipSuccess) result = -2;
    hipFree(dphi);
    hipFree(dphiw);
    hipFree(dfail);
    return result < 0 ? result : (int64_t)nf;
}

void* uqk_device_alloc(int64_t bytes) {
    void* p = nullptr;
    if (bytes <= 0) bytes = 1;
    if (hipMalloc(&p, (size_t)bytes) != hipSuccess) return nullptr;
    return p;
}

int uqk_device_free(void* p) { return hipFree(p) == hipSuccess ? 0 : -2; }

int uqk_memcpy(void* dst, const void* src, int64_t bytes, int32_t kind) {
    if (bytes <= 0) return 0;
    hipMemcpyKind k = kind == 0 ? hipMemcpyHostToDevice : kind == 1 ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice;
    return hipMemcpy(dst, src, (size_t)bytes, k) == hipSuccess ? 0 : -2;
}

int uqk_device_synchronize(void) { return hipDeviceSynchronize() == hipSuccess ? 0 : -2; }

int32_t uqk_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

}  // extern "C"
