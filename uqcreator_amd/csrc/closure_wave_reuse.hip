// closure_wave_reuse.hip -- the time loop's mixed one-wavefront Newton
// closure with the reuse of the factorisation (closure_wave_kernel.h,
// PREC 2) and the kernel that inverts the factors it leaves, compiled in a
// file of their own with the default scheduler.
#include "closure_wave_kernel.h"

namespace uqk {

// Inverse of H = U^T U for the cells a closure launch factored: U packed
// upper (U(r, c) at c (c + 1) / 2 + r) then 1/U_rr, FP64; the launch listed
// those cells (list[0 .. *count)) and set their flags to 2.  One wave
// (workgroup) per listed cell, the workgroups strided over the list: lane c
// keeps column c of U in registers and solves U^T y = e_c, then U x = y, in
// FP64 registers; the entries of U reach the other lanes through SGPRs
// (v_readlane).  Column c (= row c) of H^{-1} goes out in FP32 (row stride
// N + 1) and the flag becomes 1.
template <int N>
__global__ void __launch_bounds__(64) factor_inverse_kernel(const unsigned long long* count, const int* list,
                                                            int* flags, const double* ufac, float* pinv,
                                                            const int* active) {
    constexpr int HP = N * (N + 1) / 2;
    if (active && *active == 0) return;
    const int lane = threadIdx.x;
    const int cl = lane < N ? lane : 0;
    const long cnt = (long)*count;
    for (long t = blockIdx.x; t < cnt; t += gridDim.x) {
        const long cell = list[t];
        const double* ug = ufac + cell * (long)(HP + N);
        double uc[N];  // U(r, cl), zero below the diagonal
#pragma unroll
        for (int r = 0; r < N; ++r) uc[r] = r <= cl ? ug[cl * (cl + 1) / 2 + r] : 0.0;
        const double rdl = ug[HP + cl];
        double y[N];
#pragma unroll
        for (int i = 0; i < N; ++i) y[i] = lane == i ? 1.0 : 0.0;
#pragma unroll
        for (int k = 0; k < N; ++k) {  // U^T y = e_lane, right-looking: U(k, i) from lane i
            y[k] *= readlane_f64(rdl, k);
#pragma unroll
            for (int i = k + 1; i < N; ++i) y[i] = fma(-readlane_f64(uc[k], i), y[k], y[i]);
        }
#pragma unroll
        for (int k = N - 1; k >= 0; --k) {  // U x = y, column k of U (lane k) at a time
            y[k] *= readlane_f64(rdl, k);
#pragma unroll
            for (int i = 0; i < k; ++i) y[i] = fma(-readlane_f64(uc[i], k), y[k], y[i]);
        }
        float* P = pinv + cell * (long)(N * (N + 1));
        if (lane < N) {
#pragma unroll
            for (int i = 0; i < N; ++i) P[i * (N + 1) + lane] = (float)y[i];
        }
        if (lane == 0) flags[cell] = 1;
    }
}

template <int KIND, int M, int Q>
static bool inverse_try(int kind, int m, int q, long ncell, const unsigned long long* count, const int* list,
                        int* flags, const double* ufac, float* pinv, const int* active, int num_cu, hipStream_t st,
                        hipError_t* err) {
    if (kind != KIND || m != M || q != Q) return false;
    constexpr int N = KindInfo<KIND>::S * M;
    // up to 8 workgroups (waves) per CU, strided over the list
    const long cap = 8L * (num_cu > 0 ? num_cu : 1);
    const unsigned blocks = (unsigned)(ncell < 1 ? 1 : (ncell < cap ? ncell : cap));
    hipLaunchKernelGGL(factor_inverse_kernel<N>, dim3(blocks), dim3(64), 0, st, count, list, flags, ufac, pinv,
                       active);
    *err = hipGetLastError();
    return true;
}

bool factor_inverse_launch(int kind, int M, int Q, long ncell, const unsigned long long* count, const int* list,
                           int* flags, const double* ufac, float* pinv, const int* active, int num_cu,
                           hipStream_t st, hipError_t* err) {
#define UQK_INV(K, MM, QQ) inverse_try<K, MM, QQ>(kind, M, Q, ncell, count, list, flags, ufac, pinv, active, num_cu, st, err)
    return UQK_INV(UQK_BOUNDED_BARRIER, 6, 12) || UQK_INV(UQK_EULER2D, 10, 25) ||
           UQK_INV(UQK_SHALLOW_WATER2D, 10, 25) || UQK_INV(UQK_SHALLOW_WATER2D_BARRIER, 10, 25);
#undef UQK_INV
}

template <int KIND, int M, int Q>
static bool reuse_try(int kind, int m, int q, const NewtonArgs& a, int max_lds, int num_cu, hipStream_t st,
                      hipError_t* err, const LaunchShape& shape) {
    if (!wave_fits<KIND, M, Q>(kind, m, q, max_lds)) return false;
    wave_go<KIND, M, Q, 2>(a, max_lds, num_cu, st, err, shape);
    return true;
}

bool newton_wave_reuse_fits(int kind, int M, int Q, int max_lds) {
    return wave_fits<UQK_BOUNDED_BARRIER, 6, 12>(kind, M, Q, max_lds) ||
           wave_fits<UQK_EULER2D, 10, 25>(kind, M, Q, max_lds) ||
           wave_fits<UQK_SHALLOW_WATER2D, 10, 25>(kind, M, Q, max_lds) ||
           wave_fits<UQK_SHALLOW_WATER2D_BARRIER, 10, 25>(kind, M, Q, max_lds);
}

bool newton_wave_reuse_launch(int kind, int M, int Q, const NewtonArgs& a, int max_lds, int num_cu, hipStream_t st,
                              hipError_t* err, const LaunchShape& shape) {
    return reuse_try<UQK_BOUNDED_BARRIER, 6, 12>(kind, M, Q, a, max_lds, num_cu, st, err, shape) ||
           reuse_try<UQK_EULER2D, 10, 25>(kind, M, Q, a, max_lds, num_cu, st, err, shape) ||
           reuse_try<UQK_SHALLOW_WATER2D, 10, 25>(kind, M, Q, a, max_lds, num_cu, st, err, shape) ||
           reuse_try<UQK_SHALLOW_WATER2D_BARRIER, 10, 25>(kind, M, Q, a, max_lds, num_cu, st, err, shape);
}

}  // namespace uqk
