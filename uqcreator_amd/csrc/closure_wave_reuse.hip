// closure_wave_reuse.hip -- the time loop's mixed one-wavefront Newton
// closure with the reuse of the factorisation (closure_wave_kernel.h,
// PREC 2) and the kernel that inverts the factors it leaves, compiled in a
// file of their own with the default scheduler.
#include "closure_wave_kernel.h"

namespace uqk {

// Inverse of H = U^T U for the cells a closure launch factored: U packed
// upper (U(r, c) at c (c + 1) / 2 + r) then 1/U_rr, FP64; the launch listed
// those cells (list[0 .. *count)) and set their flags to 2.  One wave
// (workgroup) per listed cell, the workgroups strided over the list, U in
// LDS, in FP64:
//   * W = U^-1 in place by 16 x 16 blocks: the diagonal blocks by 16-lane
//     groups, a column at a time (W U = I: W_jj = 1/U_jj, W_ij =
//     -(sum_{i<=k<j} W_ik U_kj) / U_jj); the blocks above the diagonal a
//     block column J at a time, W_IJ = -(sum_{I<=K<J} W_IK U_KJ) W_JJ, as
//     16 x 16 x 4 FP64 MFMA products (every T_I of the column summed before
//     any W_IJ replaces its U_IJ; T into the A layout through LDS);
//   * H^-1 = W W^T, tile (I, J), I <= J, as sum_{K >= J} W_IK W_JK^T on the
//     same MFMA, written in FP32 (row stride N + 1) with the mirror image of
//     the tiles off the diagonal.
// The flag becomes 1.
template <int N>
__global__ void __launch_bounds__(64) factor_inverse_kernel(const unsigned long long* count, const int* list,
                                                            int* flags, const double* ufac, float* pinv,
                                                            const int* active) {
    constexpr int HP = N * (N + 1) / 2, NB = (N + 15) / 16;
    static_assert(NB <= 4, "one wave inverts n <= 64");
    __shared__ double W[HP];
    __shared__ double rd[N];
    __shared__ double scr[256];
    if (active && *active == 0) return;
    const int lane = threadIdx.x;
    const int r16 = lane & 15, kq4 = lane >> 4;
    auto idx = [](int r, int c) { return c * (c + 1) / 2 + r; };  // r <= c
    const long cnt = (long)*count;
    for (long t = blockIdx.x; t < cnt; t += gridDim.x) {
        const long cell = list[t];
        const double* ug = ufac + cell * (long)(HP + N);
        for (int i = lane; i < HP; i += 64) W[i] = ug[i];
        for (int i = lane; i < N; i += 64) rd[i] = ug[HP + i];
        wave_sync();
        // the diagonal blocks: lanes 16 b .. 16 b + 15 on block b
        {
            const int b0 = 16 * kq4, ii = b0 + r16;
            for (int j = 0; j < 16; ++j) {
                const int jc = b0 + j;
                const bool col = jc < N;
                double s = 0.0;
                if (col && r16 < j)
                    for (int k = r16; k < j; ++k) s = fma(W[idx(ii, b0 + k)], W[idx(b0 + k, jc)], s);
                wave_sync();
                if (col && r16 < j)
                    W[idx(ii, jc)] = -s * rd[jc];
                else if (col && r16 == j)
                    W[idx(jc, jc)] = rd[jc];
                wave_sync();
            }
        }
        // the blocks above the diagonal
#pragma unroll
        for (int J = 1; J < NB; ++J) {
            const int gcol = 16 * J + r16;
            f64x4 tp[NB > 1 ? NB - 1 : 1];
#pragma unroll
            for (int I = 0; I < J; ++I) {
                f64x4 acc = {0.0, 0.0, 0.0, 0.0};
                const int gi = 16 * I + r16;
#pragma unroll
                for (int K = I; K < J; ++K)
#pragma unroll
                    for (int ks = 0; ks < 4; ++ks) {
                        const int kk = 16 * K + 4 * ks + kq4;
                        const double av = gi <= kk ? W[idx(gi, kk)] : 0.0;
                        const double bv = gcol < N ? W[idx(kk, gcol)] : 0.0;
                        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, acc, 0, 0, 0);
                    }
                tp[I] = acc;
            }
            wave_sync();
#pragma unroll
            for (int I = 0; I < J; ++I) {
#pragma unroll
                for (int e = 0; e < 4; ++e) scr[(4 * e + kq4) * 16 + r16] = tp[I][e];
                wave_sync();
                f64x4 w = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
                for (int ks = 0; ks < 4; ++ks) {
                    const int k = 4 * ks + kq4;
                    const double av = scr[r16 * 16 + k];
                    const double bv = (k <= r16 && gcol < N) ? W[idx(16 * J + k, gcol)] : 0.0;
                    w = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, w, 0, 0, 0);
                }
                wave_sync();
                if (gcol < N) {
#pragma unroll
                    for (int e = 0; e < 4; ++e) W[idx(16 * I + 4 * e + kq4, gcol)] = -w[e];
                }
            }
            wave_sync();
        }
        // H^-1 = W W^T by tiles (I, J), I <= J: sum over K >= J of W_IK W_JK^T
        float* P = pinv + cell * (long)(N * (N + 1));
#pragma unroll
        for (int J = 0; J < NB; ++J)
#pragma unroll
            for (int I = 0; I <= J; ++I) {
                f64x4 acc = {0.0, 0.0, 0.0, 0.0};
                const int gi = 16 * I + r16, gj = 16 * J + r16;
#pragma unroll
                for (int K = J; K < NB; ++K)
#pragma unroll
                    for (int ks = 0; ks < 4; ++ks) {
                        const int kk = 16 * K + 4 * ks + kq4;
                        const bool kin = kk < N;
                        const double av = (kin && gi <= kk) ? W[idx(gi, kk)] : 0.0;
                        const double bv = (kin && gj < N && gj <= kk) ? W[idx(gj, kk)] : 0.0;
                        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, acc, 0, 0, 0);
                    }
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const int r = 16 * I + 4 * e + kq4, c = gj;
                    if (r < N && c < N) {
                        P[r * (N + 1) + c] = (float)acc[e];
                        if (I < J) P[c * (N + 1) + r] = (float)acc[e];  // (a diagonal tile holds both)
                    }
                }
            }
        if (lane == 0) flags[cell] = 1;
        wave_sync();
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
