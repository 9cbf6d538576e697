// mlp.hip -- learned warm start for the Newton closure (no counterpart in
// UQCreator): a fully connected network maps normalised moments to a dual
// correction, trained on the GPU on (u, lambda) pairs that the Newton closure
// itself produced, and compared per cell with the plain start.
//
// Network: K0 inputs -> four hidden layers of width H with GELU (tanh form)
// -> linear head of NO outputs.  K0 = NO = S*M rounded up to 32 (zero
// padded).  Weights are FP32 masters with BF16 copies; every product runs on
// v_mfma_f32_16x16x32_bf16 with FP32 accumulation; Adam updates the masters.
//
// Features and targets (Euler closures, S = 3 or 4 states; shallow water,
// S = 3, with or without the barrier): from the cell mean u_0 take feature
// scales D and target scales ts (cell_scales below; Euler: D = (rho0,
// rho0 sqrt(T0) [per momentum], rho0 T0) with T0 = p0/rho0 and ts = D/rho0;
// shallow water: D = (h0, h0 a0, h0 a0) with a0 = sqrt(g h0) and ts = (1/a0^2,
// 1/a0, 1/a0)).  The input is u(i,s)/D_s; the target is (lambda -
// lambda_lin)(i,s) * ts_s with lambda_lin the linearised start (LinearStart
// below: the plain start (DS(u_0), 0, ...) followed by one Newton step, in
// closed form), so lambda = lambda_lin + y / ts_s inverts it at any scale.
// Both are standardised per feature with the training set's mean and std.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <type_traits>
#include <vector>

#include "synth.h"
#include "uqk.h"
#include "uqk_ctx.h"
#include "uqk_device.h"
#include "uqk_internal.h"

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

namespace {

constexpr int kLayers = 5;  // four hidden + head
constexpr int kTile = 64;   // GEMM workgroup tile (rows and columns)
constexpr int kKs = 32;     // GEMM K step (one MFMA 16x16x32)
constexpr int kLdsStride = 40;  // bf16 per LDS tile row (32 + 8 pad)
constexpr long kPredictChunk = 1L << 16;  // rows per inference pass of the warm start

enum { EPI_GELU = 0, EPI_LINEAR = 1, EPI_DGELU = 2, EPI_ATOMIC = 3 };

struct GemmArgs {
    int M, N, K;              // output M x N, reduction K (a multiple of 32)
    const __bf16* A;          // !TA: A[m*lda + k]   TA: A[k*lda + m]
    int lda;
    const __bf16* B;          // !TB: B[n*ldb + k]   TB: B[k*ldb + n]
    int ldb;
    const float* bias;        // [N]            (GELU, LINEAR)
    float* z;                 // GELU: z store (may be null); DGELU: z input
    int ldz;
    __bf16* h;                // GELU: gelu(z);  DGELU: dz
    int ldh;
    float* f;                 // LINEAR: y;      ATOMIC: f += acc
    int ldf;
    int k_split;              // K range per blockIdx.z (a multiple of 32)
};

__device__ __forceinline__ float gelu_f(float z) {
    const float u = 0.7978845608028654f * (z + 0.044715f * z * z * z);
    const float th = 1.0f - 2.0f / (expf(2.0f * u) + 1.0f);
    return 0.5f * z * (1.0f + th);
}

__device__ __forceinline__ float gelu_prime_f(float z) {
    const float u = 0.7978845608028654f * (z + 0.044715f * z * z * z);
    const float th = 1.0f - 2.0f / (expf(2.0f * u) + 1.0f);
    return 0.5f * (1.0f + th) + 0.5f * z * (1.0f - th * th) * 0.7978845608028654f * (1.0f + 0.134145f * z * z);
}

// Stage a 64 x 32 tile of an operand into LDS as [row][k] (row = m or n).
// Not transposed: rows r0.. of P[row*ld + k]; transposed: P[k*ld + row].
template <bool T>
__device__ __forceinline__ void stage(__bf16* S, const __bf16* P, int ld, int r0, int R, int k0) {
    const int t = threadIdx.x;
    if constexpr (!T) {
        const int row = t >> 2, kc = (t & 3) * 8;
        bf16x8 v = {};
        if (r0 + row < R) v = *reinterpret_cast<const bf16x8*>(P + (size_t)(r0 + row) * ld + k0 + kc);
        *reinterpret_cast<bf16x8*>(S + row * kLdsStride + kc) = v;
    } else {
        const int kk = t >> 3, rc = (t & 7) * 8;
        const __bf16* src = P + (size_t)(k0 + kk) * ld + r0 + rc;
        bf16x8 v = {};
        if (r0 + rc + 8 <= R) {
            v = *reinterpret_cast<const bf16x8*>(src);
        } else {
#pragma unroll
            for (int j = 0; j < 8; ++j)
                if (r0 + rc + j < R) v[j] = src[j];
        }
#pragma unroll
        for (int j = 0; j < 8; ++j) S[(rc + j) * kLdsStride + kk] = v[j];
    }
}

// C (M x N) = op(A) op(B)^T over K with a fused epilogue.  256 threads, a
// 64 x 64 tile per workgroup, each wave a 32 x 32 quarter (2 x 2 MFMA tiles).
// blockIdx.z splits K (ATOMIC epilogue only).
template <bool TA, bool TB, int EPI>
__global__ __launch_bounds__(256) void gemm_bf16_kernel(GemmArgs g) {
    __shared__ __attribute__((aligned(16))) __bf16 As[kTile * kLdsStride];
    __shared__ __attribute__((aligned(16))) __bf16 Bs[kTile * kLdsStride];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int m0 = blockIdx.x * kTile, n0 = blockIdx.y * kTile;
    const int kb = blockIdx.z * g.k_split;
    const int ke = min(g.K, kb + g.k_split);
    const int wr = (w >> 1) * 32, wc = (w & 1) * 32;
    f32x4 acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) acc[a][b] = f32x4{0.f, 0.f, 0.f, 0.f};
    for (int k0 = kb; k0 < ke; k0 += kKs) {
        stage<TA>(As, g.A, g.lda, m0, g.M, k0);
        stage<TB>(Bs, g.B, g.ldb, n0, g.N, k0);
        __syncthreads();
        bf16x8 a[2], b[2];
#pragma unroll
        for (int r = 0; r < 2; ++r)
            a[r] = *reinterpret_cast<const bf16x8*>(As + (wr + r * 16 + (lane & 15)) * kLdsStride + 8 * (lane >> 4));
#pragma unroll
        for (int c = 0; c < 2; ++c)
            b[c] = *reinterpret_cast<const bf16x8*>(Bs + (wc + c * 16 + (lane & 15)) * kLdsStride + 8 * (lane >> 4));
#pragma unroll
        for (int r = 0; r < 2; ++r)
#pragma unroll
            for (int c = 0; c < 2; ++c) acc[r][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[r], b[c], acc[r][c], 0, 0, 0);
        __syncthreads();
    }
#pragma unroll
    for (int r = 0; r < 2; ++r)
#pragma unroll
        for (int c = 0; c < 2; ++c)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int m = m0 + wr + r * 16 + 4 * (lane >> 4) + i;
                const int n = n0 + wc + c * 16 + (lane & 15);
                if (m >= g.M || n >= g.N) continue;
                const float v = acc[r][c][i];
                if constexpr (EPI == EPI_GELU) {
                    const float z = v + g.bias[n];
                    if (g.z) g.z[(size_t)m * g.ldz + n] = z;
                    g.h[(size_t)m * g.ldh + n] = (__bf16)gelu_f(z);
                } else if constexpr (EPI == EPI_LINEAR) {
                    g.f[(size_t)m * g.ldf + n] = v + g.bias[n];
                } else if constexpr (EPI == EPI_DGELU) {
                    g.h[(size_t)m * g.ldh + n] = (__bf16)(v * gelu_prime_f(g.z[(size_t)m * g.ldz + n]));
                } else {
                    atomicAdd(&g.f[(size_t)m * g.ldf + n], v);
                }
            }
}

// db[n] += sum over rows of x[row*ld + n] (x FP32 or BF16): a block of 256
// threads covers kColsumRows rows of 256 columns, eight loads in flight per
// thread, one atomic per column and block.
constexpr int kColsumRows = 32;
inline dim3 colsum_grid(long rows, int N) {
    return dim3((unsigned)((rows + kColsumRows - 1) / kColsumRows), (unsigned)((N + 255) / 256));
}
template <class T>
__global__ __launch_bounds__(256) void colsum_kernel(const T* x, int ld, int rows, int N, float* db) {
    const int n = blockIdx.y * 256 + threadIdx.x;
    if (n >= N) return;
    const int r0 = blockIdx.x * kColsumRows, r1 = min(rows, r0 + kColsumRows);
    float s = 0.f;
#pragma unroll 8
    for (int r = r0; r < r1; ++r) s += (float)x[(size_t)r * ld + n];
    atomicAdd(&db[n], s);
}

// dy = 2 (y - t) / (rows * n_real) for the first n_real outputs, 0 beyond;
// FP32 (bias gradient) and BF16 (products) copies; loss accumulated in FP64.
__global__ void loss_kernel(const float* y, const float* t, int rows, int NO, int n_real, float* dyf, __bf16* dyh,
                            double* loss) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    double e2 = 0.0;
    if (i < (size_t)rows * NO) {
        const int n = (int)(i % NO);
        float d = 0.f;
        if (n < n_real) {
            const float e = y[i] - t[i];
            e2 = (double)e * e;
            d = 2.0f * e / ((float)rows * (float)n_real);
        }
        dyf[i] = d;
        dyh[i] = (__bf16)d;
    }
    e2 = uqk::wave_sum(e2);
    if ((threadIdx.x & 63) == 0 && e2 != 0.0) atomicAdd(loss, e2 / ((double)rows * n_real));
}

// Adam (Kingma and Ba 2015) on the FP32 masters, BF16 copies refreshed.
// lr_t carries the bias corrections: lr sqrt(1 - b2^t) / (1 - b1^t).
__global__ void adam_kernel(float* p, const float* g, float* m, float* v, __bf16* ph, long n, float lr_t, float b1,
                            float b2, float eps) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float pi = p[i] - lr_t * mi / (sqrtf(vi) + eps);
    p[i] = pi;
    ph[i] = (__bf16)pi;
}

__global__ void fill_f32_kernel(float* x, long n, float v) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] = v;
}

__global__ void to_bf16_kernel(const float* x, __bf16* y, long n) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = (__bf16)x[i];
}

// Cell scales and plain start from the cell's mean state u_0: feature
// scales D (moments are divided by them), target scales ts (the dual
// corrections are multiplied by them, which makes them dimensionless) and
// ds = DS(u_0).  Returns false for a non-physical mean.
//   Euler: D = (rho, rho sqrt(T), ..., rho T), ts = D / rho; DS closed forms
//   of EulerClosure::DS (eulerclosure.cpp:70-80) and EulerClosure2D::DS
//   (eulerclosure2d.cpp:90-102), gamma negated there.
//   Shallow water: D = (h, h a, h a) with a = sqrt(g h), ts = (1/a^2, 1/a,
//   1/a); DS inverts U (shallowwaterclosure2d.cpp:10-14): Lambda_1,2 = the
//   velocity, Lambda_0 = g h - |v|^2 / 2 (with the barrier of the
//   log-barrier entropy, g h - eps / h - |v|^2 / 2).
template <int KIND>
__device__ __forceinline__ bool cell_scales(const double* u, const uqk::ClosureParams& cp, double* D, double* ts,
                                            double* ds) {
    constexpr int S = uqk::KindInfo<KIND>::S;
    if constexpr (KIND == UQK_EULER1D || KIND == UQK_EULER2D) {
        const double gm = cp.gamma;
        const double rho = u[0], E = u[S - 1];
        double q2 = 0.0;
        for (int s = 1; s < S - 1; ++s) q2 += u[s] * u[s];
        const double p = (gm - 1.0) * (E - 0.5 * q2 / rho);
        const double T = p / rho;
        if (!(rho > 0.0) || !(T > 0.0) || !isfinite(T)) return false;
        const double sq = sqrt(T);
        D[0] = rho;
        for (int s = 1; s < S - 1; ++s) D[s] = rho * sq;
        D[S - 1] = rho * T;
        for (int s = 0; s < S; ++s) ts[s] = D[s] / rho;
        const double g = -gm;
        const double e_int = E - q2 / (2.0 * rho);
        ds[0] = (q2 + g * (2.0 * rho * E - q2)) / (-2.0 * rho * E + q2) - log(pow(rho, g) * e_int);
        for (int s = 1; s < S - 1; ++s) ds[s] = -((2.0 * rho * u[s]) / (-2.0 * rho * E + q2));
        ds[S - 1] = -(rho / e_int);
        return true;
    } else {
        const double h = u[0];
        const double a2 = cp.g * h;
        if (!(h > 0.0) || !(a2 > 0.0) || !isfinite(a2)) return false;
        const double a = sqrt(a2);
        const double vx = u[1] / h, vy = u[2] / h;
        D[0] = h;
        D[1] = D[2] = h * a;
        ts[0] = 1.0 / a2;
        ts[1] = ts[2] = 1.0 / a;
        double c0 = a2;
        if constexpr (KIND == UQK_SHALLOW_WATER2D_BARRIER) c0 -= cp.barrier / h;
        ds[0] = c0 - 0.5 * (vx * vx + vy * vy);
        ds[1] = vx;
        ds[2] = vy;
        return isfinite(ds[0]) && isfinite(vx) && isfinite(vy);
    }
}

// The closures the warm start supports, dispatched at run time.
template <class F>
bool with_warm_kind(int kind, F&& f) {
    switch (kind) {
    case UQK_EULER1D: f(std::integral_constant<int, UQK_EULER1D>()); return true;
    case UQK_EULER2D: f(std::integral_constant<int, UQK_EULER2D>()); return true;
    case UQK_SHALLOW_WATER2D: f(std::integral_constant<int, UQK_SHALLOW_WATER2D>()); return true;
    case UQK_SHALLOW_WATER2D_BARRIER: f(std::integral_constant<int, UQK_SHALLOW_WATER2D_BARRIER>()); return true;
    default: return false;
    }
}

// Linearised start: at lambda = (DS(u_0), 0, ...) every quadrature point
// sees the same Lambda, so the gradient is (0, -u_1, ..., -u_{M-1}) and the
// Hessian is I_M (x) DU(DS(u_0)) (orthonormal basis).  One full Newton step
// from the plain start is therefore lambda_i = DU^{-1} u_i (i >= 1) with
// lambda_0 = DS(u_0) unchanged -- closed form, S x S per cell.  The network
// learns what remains.  LinearStart factors DU once; solve() gives z_i.
template <int KIND>
struct LinearStart {
    static constexpr int S = uqk::KindInfo<KIND>::S;
    double L[S][S];
    __device__ bool factor(const uqk::ClosureParams& cp, const double* ds) {
        double y[S * (S + 1) / 2];
        uqk::closure_DU<KIND>(cp, ds, y);
        for (int l = 0; l < S; ++l)
            for (int m = 0; m <= l; ++m) L[l][m] = y[l * (l + 1) / 2 + m];
        for (int j = 0; j < S; ++j) {
            double d = L[j][j];
            for (int k = 0; k < j; ++k) d -= L[j][k] * L[j][k];
            if (!(d > 0.0) || !isfinite(d)) return false;
            L[j][j] = sqrt(d);
            for (int i = j + 1; i < S; ++i) {
                double v = L[i][j];
                for (int k = 0; k < j; ++k) v -= L[i][k] * L[j][k];
                L[i][j] = v / L[j][j];
            }
        }
        return true;
    }
    __device__ void solve(const double* b, double* z) const {
        double w[S];
        for (int i = 0; i < S; ++i) {
            double v = b[i];
            for (int k = 0; k < i; ++k) v -= L[i][k] * w[k];
            w[i] = v / L[i][i];
        }
        for (int i = S - 1; i >= 0; --i) {
            double v = w[i];
            for (int k = i + 1; k < S; ++k) v -= L[k][i] * z[k];
            z[i] = v / L[i][i];
        }
    }
};

struct FeatArgs {
    long n;                   // rows
    const long* idx;          // row -> cell (null: identity)
    const double* u;          // cells x M x S
    const double* lam;        // cells x M x S (null: no targets)
    const int* nt;            // per-cell orders: moments per cell (null: M)
    int M;
    uqk::ClosureParams cp;
    int K0, NO;
    const float* mu_x;        // [K0] (null: raw features)
    const float* sd_x;
    const float* mu_t;        // [NO]
    const float* sd_t;
    __bf16* xh;               // rows x K0, standardised BF16 (may be null)
    float* xf;                // rows x K0, raw FP32 (may be null)
    float* tf;                // rows x NO, targets (raw if mu_t null)
};

// Per-row scales and linearised start of a block's rows, staged in LDS by
// the first kRowsPerBlock threads (one row each) for the block's other
// threads, which then spread over (row, feature) pairs.
constexpr int kRowsPerBlock = 64;
template <int S>
struct RowScales {
    double D[kRowsPerBlock][S];
    double ts[kRowsPerBlock][S];
    double ds[kRowsPerBlock][S];
    double L[kRowsPerBlock][S][S];
    int ok[kRowsPerBlock];
};

template <int KIND>
__device__ void stage_rows(RowScales<uqk::KindInfo<KIND>::S>& rs, long n, long r0, const long* idx,
                           const double* u, int M, const uqk::ClosureParams& cp) {
    constexpr int S = uqk::KindInfo<KIND>::S;
    const int t = threadIdx.x;
    if (t < kRowsPerBlock) {
        const long r = r0 + t;
        int ok = 0;
        if (r < n) {
            const long c = idx ? idx[r] : r;
            double D[S], ts[S], ds[S];
            LinearStart<KIND> lin;
            ok = cell_scales<KIND>(u + (size_t)c * M * S, cp, D, ts, ds) && lin.factor(cp, ds);
            for (int s = 0; s < S; ++s) {
                rs.D[t][s] = D[s];
                rs.ts[t][s] = ts[s];
                rs.ds[t][s] = ds[s];
                for (int q = 0; q < S; ++q) rs.L[t][s][q] = lin.L[s][q];
            }
        }
        rs.ok[t] = ok;
    }
    __syncthreads();
}

// z = the linearised start of moment i of block row `row` (i = 0: DS(u_0)).
template <int KIND>
__device__ void linear_start_z(const RowScales<uqk::KindInfo<KIND>::S>& rs, int row, int i, const double* ui,
                               double* z) {
    constexpr int S = uqk::KindInfo<KIND>::S;
    if (i == 0) {
        for (int s = 0; s < S; ++s) z[s] = rs.ds[row][s];
        return;
    }
    LinearStart<KIND> lin;
    for (int s = 0; s < S; ++s)
        for (int q = 0; q < S; ++q) lin.L[s][q] = rs.L[row][s][q];
    lin.solve(ui, z);
}

template <int KIND>
__global__ __launch_bounds__(256) void features_kernel(FeatArgs a) {
    constexpr int S = uqk::KindInfo<KIND>::S;
    __shared__ RowScales<S> rs;
    const long r0 = (long)blockIdx.x * kRowsPerBlock;
    stage_rows<KIND>(rs, a.n, r0, a.idx, a.u, a.M, a.cp);
    const int nf = a.M * S;
    // features x_f = u_f / D_s, standardised; consecutive threads write
    // consecutive features of a row
    for (int t = threadIdx.x; t < kRowsPerBlock * a.K0; t += blockDim.x) {
        const int row = t / a.K0, f = t - row * a.K0;
        const long r = r0 + row;
        if (r >= a.n) break;
        float x = 0.f;
        if (rs.ok[row] && f < nf) {
            const long c = a.idx ? a.idx[r] : r;
            // moments past the cell's order are absent: a zero moment
            if (!a.nt || f / S < a.nt[c]) x = (float)(a.u[(size_t)c * nf + f] / rs.D[row][f % S]);
            if (a.mu_x) x = (x - a.mu_x[f]) / a.sd_x[f];
        }
        if (a.xh) a.xh[(size_t)r * a.K0 + f] = (__bf16)x;
        if (a.xf) a.xf[(size_t)r * a.K0 + f] = x;
    }
    if (!a.lam) return;
    // targets: the correction to the linearised start, scaled by ts_s
    for (int t = threadIdx.x; t < kRowsPerBlock * a.M; t += blockDim.x) {
        const int row = t / a.M, i = t - row * a.M;
        const long r = r0 + row;
        if (r >= a.n) break;
        const long c = a.idx ? a.idx[r] : r;
        const double* u = a.u + (size_t)c * nf;
        const double* l = a.lam + (size_t)c * nf;
        float* tr = a.tf + (size_t)r * a.NO;
        double z[S];
        if (rs.ok[row]) linear_start_z<KIND>(rs, row, i, u + i * S, z);
        for (int s = 0; s < S; ++s) {
            const int f = i * S + s;
            float tv = 0.f;
            if (rs.ok[row]) {
                tv = (float)((l[f] - z[s]) * rs.ts[row][s]);
                if (a.mu_t) tv = (tv - a.mu_t[f]) / a.sd_t[f];
            }
            tr[f] = tv;
        }
        if (i == 0)
            for (int f = nf; f < a.NO; ++f) tr[f] = 0.f;
    }
}

// lambda = lambda_lin + (mu_t + sd_t y) / ts_s; NaN where the mean state is
// not physical (the selection then keeps the plain start); zero past a
// cell's order nt[c] (per-cell orders).  A block covers kRowsPerBlock cells;
// threads over (cell, moment) pairs.
template <int KIND>
__global__ __launch_bounds__(256) void apply_kernel(long n, const double* u, const int* nt, const float* y, int M,
                                                    uqk::ClosureParams cp, int NO, const float* mu_t,
                                                    const float* sd_t, double* lam) {
    constexpr int S = uqk::KindInfo<KIND>::S;
    __shared__ RowScales<S> rs;
    const long c0 = (long)blockIdx.x * kRowsPerBlock;
    stage_rows<KIND>(rs, n, c0, nullptr, u, M, cp);
    for (int t = threadIdx.x; t < kRowsPerBlock * M; t += blockDim.x) {
        const int row = t / M, i = t - row * M;
        const long c = c0 + row;
        if (c >= n) break;
        const double* uc = u + (size_t)c * M * S;
        double z[S];
        if (rs.ok[row]) linear_start_z<KIND>(rs, row, i, uc + i * S, z);
        for (int s = 0; s < S; ++s) {
            const int f = i * S + s;
            double v = NAN;
            if (nt && i >= nt[c]) {
                v = 0.0;  // duals past the cell's order stay zero
            } else if (rs.ok[row]) {
                const double tv = (double)mu_t[f] + (double)sd_t[f] * (double)y[(size_t)c * NO + f];
                v = z[s] + tv / rs.ts[row][s];
            }
            lam[(size_t)c * M * S + f] = v;
        }
    }
}

// Selection: per cell (one wave), ||grad L||_2 at lambda and at lambda_alt
// (closure.cpp:234-253 gradient, (:195-203) norm); the smaller one stays in
// lambda.  Non-finite norms never win.  counter += cells that took lambda_alt.
// Per-cell orders (nt not null): the sums run over the cell's nt[c] moments.
// Persistent workgroups: phi and phiW are staged in LDS once per workgroup,
// each wave walks cells with a grid stride and stages the candidate duals of
// its cell in LDS (the sums run in the same order as a plain loop over k, i).
template <int KIND>
__global__ __launch_bounds__(256) void select_kernel(uqk::ClosureParams cp, long ncell, int M, int Q,
                                                     const double* phi, const double* phiw, const double* u,
                                                     double* lam, const double* lam_alt, const int* nt,
                                                     unsigned long long* counter, const int* active) {
    constexpr int S = uqk::KindInfo<KIND>::S;
    if (active && *active == 0) return;
    extern __shared__ double sel_lds[];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int nf = M * S;
    double* tphi = sel_lds;          // Q x M
    double* tphiw = tphi + Q * M;    // Q x M
    double* uq = tphiw + Q * M + (size_t)w * (Q * S + nf);
    double* lw = uq + Q * S;
    for (int t = threadIdx.x; t < Q * M; t += blockDim.x) {
        tphi[t] = phi[t];
        tphiw[t] = phiw[t];
    }
    __syncthreads();
    unsigned long long took = 0;
    for (long c = (long)blockIdx.x * 4 + w; c < ncell; c += (long)gridDim.x * 4) {
        const double* uc = u + (size_t)c * nf;
        const int ntc = nt ? nt[c] : M;
        double norm[2];
        for (int cand = 0; cand < 2; ++cand) {
            const double* l = (cand ? lam_alt : lam) + (size_t)c * nf;
            for (int f = lane; f < nf; f += 64) lw[f] = l[f];
            uqk::wave_sync();
            for (int k = lane; k < Q; k += 64) {
                double L[S], U[S];
                for (int s = 0; s < S; ++s) L[s] = 0.0;
                for (int i = 0; i < ntc; ++i) {
                    const double p = tphi[k * M + i];
                    for (int s = 0; s < S; ++s) L[s] += lw[i * S + s] * p;
                }
                uqk::closure_U<KIND>(cp, L, U);
                for (int s = 0; s < S; ++s) uq[k * S + s] = U[s];
            }
            uqk::wave_sync();
            double acc = 0.0;
            for (int f = lane; f < ntc * S; f += 64) {
                const int i = f / S, s = f % S;
                double gsum = 0.0;
                for (int k = 0; k < Q; ++k) gsum += uq[k * S + s] * tphiw[k * M + i];
                const double gf = gsum - uc[f];
                acc += gf * gf;
            }
            norm[cand] = sqrt(uqk::wave_sum(acc));
            uqk::wave_sync();
        }
        if (norm[1] < norm[0]) {
            for (int f = lane; f < nf; f += 64) lam[(size_t)c * nf + f] = lam_alt[(size_t)c * nf + f];
            ++took;
        }
    }
    if (lane == 0 && took) atomicAdd(counter, took);
}

__global__ void rand_index_kernel(long* idx, long rows, long n, uint64_t seed, uint64_t step) {
    const long r = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= rows) return;
    const uqk::Philox4x32 x =
        uqk::philox4x32_10((uint32_t)r, (uint32_t)step, 0x6d6c70u, 0u, (uint32_t)seed, (uint32_t)(seed >> 32));
    const uint64_t v = ((uint64_t)x.v[0] << 32) | x.v[1];
    idx[r] = (long)(v % (uint64_t)n);
}

__global__ void zero_u64_kernel(unsigned long long* x) { *x = 0ull; }

long pad32(long v) { return (v + 31) / 32 * 32; }

int fail_ctx(uqk_ctx* ctx, int code, const std::string& msg) {
    if (ctx) ctx->err = msg;
    return code;
}

#define UQK_HIP_CTX(ctx, call)                                                                          \
    do {                                                                                                \
        hipError_t e_ = (call);                                                                         \
        if (e_ != hipSuccess) return fail_ctx(ctx, -2, std::string(#call ": ") + hipGetErrorString(e_)); \
    } while (0)

}  // namespace

struct uqk_mlp {
    uqk_ctx* ctx = nullptr;
    int S = 0, M = 0, K0 = 0, NO = 0, H = 0;
    int dims[kLayers + 1];
    long off[kLayers + 1];     // parameter offsets of W_l (b_l follows W_l)
    long np = 0;
    float *p = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;
    __bf16* ph = nullptr;
    float* norm = nullptr;     // mu_x[K0], sd_x[K0], mu_t[NO], sd_t[NO]
    bool normalised = false;
    long adam_t = 0;
    // work buffers for up to `rows_cap` rows
    long rows_cap = 0;
    __bf16* act[kLayers];      // inputs of each layer (BF16)
    float* z[kLayers - 1];     // pre-activations of the hidden layers
    __bf16* dz = nullptr;      // gradient of a hidden pre-activation (BF16)
    __bf16* dz2 = nullptr;
    float* y = nullptr;        // head output
    float* t = nullptr;        // targets
    float* dyf = nullptr;
    __bf16* dyh = nullptr;
    long* idx = nullptr;
    double* d_loss = nullptr;
    unsigned long long* d_count = nullptr;
    double* lam_alt = nullptr;
    long lam_alt_cells = 0;
    std::vector<void*> owned;
};

namespace {

int mfail(uqk_mlp* n, int code, const std::string& msg) {
    if (n && n->ctx) n->ctx->err = msg;
    return code;
}

#define MLP_HIP(net, call)                                                                              \
    do {                                                                                                \
        hipError_t e_ = (call);                                                                         \
        if (e_ != hipSuccess) return mfail(net, -2, std::string(#call ": ") + hipGetErrorString(e_));   \
    } while (0)

template <class T>
hipError_t alloc(uqk_mlp* n, T** p, size_t count) {
    void* q = nullptr;
    hipError_t e = hipMalloc(&q, std::max<size_t>(count, 1) * sizeof(T));
    if (e == hipSuccess) {
        n->owned.push_back(q);
        *p = static_cast<T*>(q);
    }
    return e;
}

unsigned blocks(long n, int b = 256) { return (unsigned)((n + b - 1) / b); }

// the GEMMs of one layer; rows = batch rows (M of the forward GEMM)
int gemm(uqk_mlp* n, int epi, bool ta, bool tb, const GemmArgs& a, int splits) {
    dim3 grid(blocks(a.M, kTile), blocks(a.N, kTile), splits);
    hipStream_t st = n->ctx->stream;
    if (a.K % kKs || a.k_split % kKs) return mfail(n, -1, "mlp gemm: K not a multiple of 32");
    if (!ta && !tb && epi == EPI_GELU)
        hipLaunchKernelGGL((gemm_bf16_kernel<false, false, EPI_GELU>), grid, dim3(256), 0, st, a);
    else if (!ta && !tb && epi == EPI_LINEAR)
        hipLaunchKernelGGL((gemm_bf16_kernel<false, false, EPI_LINEAR>), grid, dim3(256), 0, st, a);
    else if (!ta && tb && epi == EPI_DGELU)
        hipLaunchKernelGGL((gemm_bf16_kernel<false, true, EPI_DGELU>), grid, dim3(256), 0, st, a);
    else if (ta && tb && epi == EPI_ATOMIC)
        hipLaunchKernelGGL((gemm_bf16_kernel<true, true, EPI_ATOMIC>), grid, dim3(256), 0, st, a);
    else
        return mfail(n, -1, "mlp gemm: unsupported operand form");
    MLP_HIP(n, hipGetLastError());
    return 0;
}

int ensure_rows(uqk_mlp* n, long rows) {
    if (rows <= n->rows_cap) return 0;
    if (n->ctx->capturing) return mfail(n, -1, "warm start: buffers must exist before graph capture");
    // buffers are only ever grown; old ones stay owned until destroy
    const long R = (rows + kTile - 1) / kTile * kTile;
    for (int l = 0; l < kLayers; ++l) MLP_HIP(n, alloc(n, &n->act[l], (size_t)R * n->dims[l]));
    for (int l = 0; l < kLayers - 1; ++l) MLP_HIP(n, alloc(n, &n->z[l], (size_t)R * n->H));
    MLP_HIP(n, alloc(n, &n->dz, (size_t)R * n->H));
    MLP_HIP(n, alloc(n, &n->dz2, (size_t)R * n->H));
    MLP_HIP(n, alloc(n, &n->y, (size_t)R * n->NO));
    MLP_HIP(n, alloc(n, &n->t, (size_t)R * n->NO));
    MLP_HIP(n, alloc(n, &n->dyf, (size_t)R * n->NO));
    MLP_HIP(n, alloc(n, &n->dyh, (size_t)R * n->NO));
    MLP_HIP(n, alloc(n, &n->idx, (size_t)R));
    n->rows_cap = R;
    return 0;
}

// forward over rows already staged in act[0]; store_z keeps what backward needs
int forward(uqk_mlp* n, long rows, bool store_z) {
    for (int l = 0; l < kLayers; ++l) {
        GemmArgs a{};
        a.M = (int)rows;
        a.N = n->dims[l + 1];
        a.K = n->dims[l];
        a.A = n->act[l];
        a.lda = n->dims[l];
        a.B = n->ph + n->off[l];
        a.ldb = n->dims[l];
        a.bias = n->p + n->off[l] + (long)n->dims[l + 1] * n->dims[l];
        a.k_split = a.K;
        int rc;
        if (l + 1 < kLayers) {
            a.z = store_z ? n->z[l] : nullptr;
            a.ldz = n->H;
            a.h = n->act[l + 1];
            a.ldh = n->dims[l + 1];
            rc = gemm(n, EPI_GELU, false, false, a, 1);
        } else {
            a.f = n->y;
            a.ldf = n->NO;
            rc = gemm(n, EPI_LINEAR, false, false, a, 1);
        }
        if (rc) return rc;
    }
    return 0;
}

// loss and gradient (into n->g) of the rows staged in act[0] / t
int backward(uqk_mlp* n, long rows, int n_real) {
    hipStream_t st = n->ctx->stream;
    hipLaunchKernelGGL(fill_f32_kernel, dim3(blocks(n->np)), dim3(256), 0, st, n->g, n->np, 0.f);
    hipLaunchKernelGGL(loss_kernel, dim3(blocks(rows * n->NO)), dim3(256), 0, st, n->y, n->t, (int)rows, n->NO,
                       n_real, n->dyf, n->dyh, n->d_loss);
    const __bf16* dcur = n->dyh;
    int ldcur = n->NO;
    const int splits = (int)std::max(1L, std::min(64L, rows / 512));
    const int ksplit = (int)((rows / splits + kKs - 1) / kKs * kKs);
    for (int l = kLayers - 1; l >= 0; --l) {
        const int N = n->dims[l + 1], K = n->dims[l];
        float* gW = n->g + n->off[l];
        float* gb = gW + (long)N * K;
        // dW_l (N x K) = dz_l^T act_l, reduction over rows (split with atomics)
        GemmArgs a{};
        a.M = N;
        a.N = K;
        a.K = (int)rows;
        a.A = dcur;
        a.lda = ldcur;
        a.B = n->act[l];
        a.ldb = K;
        a.f = gW;
        a.ldf = K;
        a.k_split = ksplit;
        if (int rc = gemm(n, EPI_ATOMIC, true, true, a, splits)) return rc;
        if (l == kLayers - 1)
            hipLaunchKernelGGL(colsum_kernel<float>, colsum_grid(rows, N), dim3(256), 0, st, n->dyf, n->NO, (int)rows, N,
                               gb);
        else
            hipLaunchKernelGGL(colsum_kernel<__bf16>, colsum_grid(rows, N), dim3(256), 0, st, dcur, ldcur, (int)rows, N,
                               gb);
        if (l == 0) break;
        // dz_{l-1} = (dz_l W_l) * gelu'(z_{l-1})
        GemmArgs b{};
        b.M = (int)rows;
        b.N = K;
        b.K = N;
        b.A = dcur;
        b.lda = ldcur;
        b.B = n->ph + n->off[l];
        b.ldb = K;
        b.z = n->z[l - 1];
        b.ldz = n->H;
        __bf16* dnext = dcur == n->dz ? n->dz2 : n->dz;
        b.h = dnext;
        b.ldh = K;
        b.k_split = N;
        if (int rc = gemm(n, EPI_DGELU, false, true, b, 1)) return rc;
        dcur = dnext;
        ldcur = K;
    }
    MLP_HIP(n, hipGetLastError());
    return 0;
}

int launch_features(uqk_mlp* n, long rows, const long* idx, const double* u, const double* lam, bool standardise,
                    __bf16* xh, float* xf, float* tf, const int* nt = nullptr) {
    FeatArgs a{};
    a.n = rows;
    a.idx = idx;
    a.u = u;
    a.lam = lam;
    a.nt = nt;
    a.M = n->M;
    a.cp = n->ctx->c;
    a.K0 = n->K0;
    a.NO = n->NO;
    if (standardise) {
        a.mu_x = n->norm;
        a.sd_x = n->norm + n->K0;
        a.mu_t = n->norm + 2 * n->K0;
        a.sd_t = n->norm + 2 * n->K0 + n->NO;
    }
    a.xh = xh;
    a.xf = xf;
    a.tf = tf;
    if (!with_warm_kind(n->ctx->kind, [&](auto k) {
            hipLaunchKernelGGL(features_kernel<decltype(k)::value>, dim3(blocks(rows, kRowsPerBlock)), dim3(256), 0,
                               n->ctx->stream, a);
        }))
        return mfail(n, -1, "warm start: unsupported closure");
    MLP_HIP(n, hipGetLastError());
    return 0;
}

}  // namespace

namespace uqk {

// features -> network -> duals for ncell cells, in chunks of 64K rows; enqueue only
int enqueue_predict(uqk_mlp* n, long ncell, const double* u, double* lambda_out, const int* nt) {
    if (ncell == 0) return 0;
    const long chunk = kPredictChunk;
    if (int rc = ensure_rows(n, std::min<long>(ncell, chunk))) return rc;
    hipStream_t st = n->ctx->stream;
    const float* mu_t = n->norm + 2 * n->K0;
    const float* sd_t = mu_t + n->NO;
    for (long c0 = 0; c0 < ncell; c0 += chunk) {
        const long rows = std::min<long>(chunk, ncell - c0);
        const double* uc = u + (size_t)c0 * n->M * n->S;
        const int* ntc = nt ? nt + c0 : nullptr;
        if (int rc = launch_features(n, rows, nullptr, uc, nullptr, true, n->act[0], nullptr, nullptr, ntc)) return rc;
        if (int rc = forward(n, rows, false)) return rc;
        double* lc = lambda_out + (size_t)c0 * n->M * n->S;
        with_warm_kind(n->ctx->kind, [&](auto k) {
            hipLaunchKernelGGL(apply_kernel<decltype(k)::value>, dim3(blocks(rows, kRowsPerBlock)), dim3(256), 0, st,
                               rows, uc, ntc, n->y, n->M, n->ctx->c, n->NO, mu_t, sd_t, lc);
        });
    }
    MLP_HIP(n, hipGetLastError());
    return 0;
}

// Selection kernel only (no readback): capturable.  counters[4] = cells that took
// lambda_alt.
int enqueue_warm_select(uqk_ctx* ctx, long ncell, const double* u, double* lambda, const double* lambda_alt,
                        const int* active, const int* nt) {
    if (!with_warm_kind(ctx->kind, [](auto) {}))
        return fail_ctx(ctx, -1, "warm start: supported for the Euler and shallow-water closures only");
    unsigned long long* cnt = ctx->d_counters + 4;
    hipLaunchKernelGGL(zero_u64_kernel, dim3(1), dim3(1), 0, ctx->stream, cnt);
    // tables once per workgroup, then per wave the points' U values and the duals
    const size_t lds = sizeof(double) * (2 * (size_t)ctx->Q * ctx->M + 4 * ((size_t)ctx->Q * ctx->S + (size_t)ctx->M * ctx->S));
    if (lds > ctx->max_lds) return fail_ctx(ctx, -1, "warm start: selection tables exceed LDS");
    const long want = (ncell + 3) / 4, cap = 8L * (ctx->num_cu > 0 ? ctx->num_cu : 1);
    const dim3 grid((unsigned)(want < cap ? (want > 0 ? want : 1) : cap));
    with_warm_kind(ctx->kind, [&](auto k) {
        hipLaunchKernelGGL(select_kernel<decltype(k)::value>, grid, dim3(256), lds, ctx->stream, ctx->c, ncell,
                           ctx->M, ctx->Q, ctx->d_phi, ctx->d_phiw, u, lambda, lambda_alt, nt, cnt, active);
    });
    UQK_HIP_CTX(ctx, hipGetLastError());
    return 0;
}

// Warm start of the attached network before a closure: predict into the
// network's lambda_alt buffer, then select against lambda.  Capturable.
int prepare_warm_start(uqk_ctx* ctx, long ncell) {
    uqk_mlp* n = ctx->warm;
    if (!n || ncell <= 0) return 0;
    if (n->lam_alt_cells < ncell) {
        if (ctx->capturing) return fail_ctx(ctx, -1, "warm start: buffers must exist before graph capture");
        double* p = nullptr;
        UQK_HIP_CTX(ctx, alloc(n, &p, (size_t)ncell * n->M * n->S));
        n->lam_alt = p;
        n->lam_alt_cells = ncell;
    }
    return ensure_rows(n, std::min<long>(ncell, kPredictChunk));
}

int enqueue_warm_predict(uqk_ctx* ctx, long ncell, const double* u, const double** lam_alt, const int* nt) {
    uqk_mlp* n = ctx->warm;
    *lam_alt = nullptr;
    if (!n) return 0;
    if (int rc = prepare_warm_start(ctx, ncell)) return rc;
    if (int rc = enqueue_predict(n, ncell, u, n->lam_alt, nt)) return rc;
    *lam_alt = n->lam_alt;
    return 0;
}

}  // namespace uqk

extern "C" {

int uqk_set_warm_start(uqk_ctx* ctx, uqk_mlp* net) {
    if (!ctx) return -1;
    if (net && net->ctx != ctx) return fail_ctx(ctx, -1, "uqk_set_warm_start: network belongs to another context");
    if (ctx->warm != net) uqk::invalidate_graph(ctx);
    ctx->warm = net;
    return 0;
}

int uqk_mlp_create(uqk_ctx* ctx, int32_t hidden, uint64_t seed, uqk_mlp** out) {
    if (!ctx || !out) return -1;
    *out = nullptr;
    if (!with_warm_kind(ctx->kind, [](auto) {}))
        return fail_ctx(ctx, -1, "uqk_mlp_create: the warm start supports the Euler and shallow-water closures");
    if (hidden <= 0 || hidden % 32) return fail_ctx(ctx, -1, "uqk_mlp_create: hidden width must be a multiple of 32");
    auto* n = new uqk_mlp();
    n->ctx = ctx;
    n->S = ctx->S;
    n->M = ctx->M;
    n->H = hidden;
    n->K0 = (int)pad32(ctx->S * ctx->M);
    n->NO = n->K0;
    n->dims[0] = n->K0;
    for (int l = 1; l < kLayers; ++l) n->dims[l] = hidden;
    n->dims[kLayers] = n->NO;
    n->np = 0;
    for (int l = 0; l < kLayers; ++l) {
        n->off[l] = n->np;
        n->np += (long)n->dims[l + 1] * n->dims[l] + n->dims[l + 1];
    }
    n->off[kLayers] = n->np;
    // He-style uniform init on the host (Philox stream keyed by seed); padded
    // input columns and output rows start at zero and receive zero gradient
    std::vector<float> h(n->np, 0.f);
    const int nf = n->S * n->M;
    for (int l = 0; l < kLayers; ++l) {
        const int N = n->dims[l + 1], K = n->dims[l];
        const float bound = std::sqrt(6.0f / (float)K) * (l + 1 < kLayers ? 1.0f : 0.1f);
        for (int r = 0; r < N; ++r)
            for (int k = 0; k < K; ++k) {
                if ((l == 0 && k >= nf) || (l + 1 == kLayers && r >= nf)) continue;
                const uqk::Philox4x32 x = uqk::philox4x32_10((uint32_t)l, (uint32_t)r, (uint32_t)k, 0x77u,
                                                             (uint32_t)seed, (uint32_t)(seed >> 32));
                h[n->off[l] + (long)r * K + k] = bound * (2.0f * (float)uqk::u01(x.v[0], x.v[1]) - 1.0f);
            }
    }
    hipError_t e = hipSuccess;
    if ((e = alloc(n, &n->p, n->np)) != hipSuccess || (e = alloc(n, &n->g, n->np)) != hipSuccess ||
        (e = alloc(n, &n->m, n->np)) != hipSuccess || (e = alloc(n, &n->v, n->np)) != hipSuccess ||
        (e = alloc(n, &n->ph, n->np)) != hipSuccess || (e = alloc(n, &n->norm, 2 * n->K0 + 2 * n->NO)) != hipSuccess ||
        (e = alloc(n, &n->d_loss, 1)) != hipSuccess || (e = alloc(n, &n->d_count, 1)) != hipSuccess) {
        for (void* q : n->owned) (void)hipFree(q);
        delete n;
        return fail_ctx(ctx, -2, std::string("uqk_mlp_create: ") + hipGetErrorString(e));
    }
    std::vector<float> nrm(2 * n->K0 + 2 * n->NO, 0.f);
    for (int i = 0; i < n->K0; ++i) nrm[n->K0 + i] = 1.f;
    for (int i = 0; i < n->NO; ++i) nrm[2 * n->K0 + n->NO + i] = 1.f;
    if (hipMemcpy(n->p, h.data(), sizeof(float) * n->np, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(n->norm, nrm.data(), sizeof(float) * nrm.size(), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemset(n->m, 0, sizeof(float) * n->np) != hipSuccess ||
        hipMemset(n->v, 0, sizeof(float) * n->np) != hipSuccess) {
        uqk_mlp_destroy(n);
        return fail_ctx(ctx, -2, "uqk_mlp_create: initialisation copy failed");
    }
    hipLaunchKernelGGL(to_bf16_kernel, dim3(blocks(n->np)), dim3(256), 0, ctx->stream, n->p, n->ph, n->np);
    if (hipStreamSynchronize(ctx->stream) != hipSuccess) {
        uqk_mlp_destroy(n);
        return fail_ctx(ctx, -2, "uqk_mlp_create: initialisation failed");
    }
    *out = n;
    return 0;
}

int uqk_mlp_destroy(uqk_mlp* n) {
    if (!n) return 0;
    if (n->ctx && n->ctx->warm == n) {
        // a captured step would still launch on this network's buffers
        uqk::invalidate_graph(n->ctx);
        n->ctx->warm = nullptr;
    }
    if (n->ctx) (void)hipStreamSynchronize(n->ctx->stream);
    for (void* q : n->owned) (void)hipFree(q);
    delete n;
    return 0;
}

int uqk_mlp_shape(const uqk_mlp* n, int32_t* dims, int64_t* n_params) {
    if (!n) return -1;
    if (dims)
        for (int l = 0; l <= kLayers; ++l) dims[l] = n->dims[l];
    if (n_params) *n_params = n->np;
    return kLayers;
}

int uqk_mlp_get_params(uqk_mlp* n, float* host_params, float* host_norm) {
    if (!n) return -1;
    MLP_HIP(n, hipStreamSynchronize(n->ctx->stream));
    if (host_params) MLP_HIP(n, hipMemcpy(host_params, n->p, sizeof(float) * n->np, hipMemcpyDeviceToHost));
    if (host_norm)
        MLP_HIP(n, hipMemcpy(host_norm, n->norm, sizeof(float) * (2 * n->K0 + 2 * n->NO), hipMemcpyDeviceToHost));
    return 0;
}

int uqk_mlp_set_params(uqk_mlp* n, const float* host_params, const float* host_norm) {
    if (!n) return -1;
    MLP_HIP(n, hipStreamSynchronize(n->ctx->stream));
    if (host_params) {
        MLP_HIP(n, hipMemcpy(n->p, host_params, sizeof(float) * n->np, hipMemcpyHostToDevice));
        hipLaunchKernelGGL(to_bf16_kernel, dim3(blocks(n->np)), dim3(256), 0, n->ctx->stream, n->p, n->ph, n->np);
    }
    if (host_norm) {
        MLP_HIP(n, hipMemcpy(n->norm, host_norm, sizeof(float) * (2 * n->K0 + 2 * n->NO), hipMemcpyHostToDevice));
        n->normalised = true;
    }
    MLP_HIP(n, hipStreamSynchronize(n->ctx->stream));
    return 0;
}

int uqk_mlp_features(uqk_mlp* n, int64_t ncell, const double* u, const double* lambda, int32_t standardise,
                     float* x, float* t) {
    if (!n || ncell < 0) return -1;
    if (ncell == 0) return 0;
    if (lambda && !t) return mfail(n, -1, "uqk_mlp_features: targets need an output buffer");
    if (int rc = launch_features(n, ncell, nullptr, u, lambda, standardise != 0, nullptr, x, t)) return rc;
    MLP_HIP(n, hipStreamSynchronize(n->ctx->stream));
    return 0;
}

int uqk_mlp_fit_normalization(uqk_mlp* n, int64_t ncell, const double* u, const double* lambda) {
    if (!n || ncell <= 1 || !u || !lambda) return mfail(n, -1, "uqk_mlp_fit_normalization: need at least two pairs");
    const long rows = std::min<long>(ncell, 1L << 18);
    float *xf = nullptr, *tf = nullptr;
    MLP_HIP(n, hipMalloc(&xf, sizeof(float) * rows * n->K0));
    if (hipMalloc(&tf, sizeof(float) * rows * n->NO) != hipSuccess) {
        (void)hipFree(xf);
        return mfail(n, -2, "uqk_mlp_fit_normalization: allocation failed");
    }
    // an evenly strided sample of the pairs
    std::vector<long> hidx(rows);
    for (long r = 0; r < rows; ++r) hidx[r] = (long)((double)r * ncell / rows);
    long* didx = nullptr;
    int rc = 0;
    std::vector<float> hx((size_t)rows * n->K0), ht((size_t)rows * n->NO);
    if (hipMalloc(&didx, sizeof(long) * rows) != hipSuccess ||
        hipMemcpy(didx, hidx.data(), sizeof(long) * rows, hipMemcpyHostToDevice) != hipSuccess) {
        rc = mfail(n, -2, "uqk_mlp_fit_normalization: index upload failed");
    } else if ((rc = launch_features(n, rows, didx, u, lambda, false, nullptr, xf, tf)) == 0) {
        if (hipStreamSynchronize(n->ctx->stream) != hipSuccess ||
            hipMemcpy(hx.data(), xf, sizeof(float) * hx.size(), hipMemcpyDeviceToHost) != hipSuccess ||
            hipMemcpy(ht.data(), tf, sizeof(float) * ht.size(), hipMemcpyDeviceToHost) != hipSuccess)
            rc = mfail(n, -2, "uqk_mlp_fit_normalization: download failed");
    }
    (void)hipFree(xf);
    (void)hipFree(tf);
    (void)hipFree(didx);
    if (rc) return rc;
    std::vector<float> nrm(2 * n->K0 + 2 * n->NO);
    auto stats = [&](const std::vector<float>& a, int W, float* mu, float* sd) {
        for (int f = 0; f < W; ++f) {
            double s = 0.0, s2 = 0.0;
            for (long r = 0; r < rows; ++r) s += a[(size_t)r * W + f];
            const double m = s / rows;
            for (long r = 0; r < rows; ++r) {
                const double d = a[(size_t)r * W + f] - m;
                s2 += d * d;
            }
            const double sdv = std::sqrt(s2 / (rows - 1));
            mu[f] = (float)m;
            sd[f] = sdv > 1e-6 * (std::fabs(m) + 1e-30) && sdv > 1e-30 ? (float)sdv : 1.0f;
        }
    };
    stats(hx, n->K0, nrm.data(), nrm.data() + n->K0);
    stats(ht, n->NO, nrm.data() + 2 * n->K0, nrm.data() + 2 * n->K0 + n->NO);
    return uqk_mlp_set_params(n, nullptr, nrm.data());
}

int uqk_mlp_loss_grad(uqk_mlp* n, int64_t rows, const float* x, const float* t, int32_t n_real, double* loss,
                      float* grad) {
    if (!n || rows <= 0) return -1;
    if (rows % kKs) return mfail(n, -1, "uqk_mlp_loss_grad: rows must be a multiple of 32");
    if (int rc = ensure_rows(n, rows)) return rc;
    hipStream_t st = n->ctx->stream;
    hipLaunchKernelGGL(to_bf16_kernel, dim3(blocks(rows * n->K0)), dim3(256), 0, st, x, n->act[0], rows * n->K0);
    MLP_HIP(n, hipMemcpyAsync(n->t, t, sizeof(float) * rows * n->NO, hipMemcpyDeviceToDevice, st));
    MLP_HIP(n, hipMemsetAsync(n->d_loss, 0, sizeof(double), st));
    if (int rc = forward(n, rows, true)) return rc;
    if (int rc = backward(n, rows, n_real)) return rc;
    MLP_HIP(n, hipStreamSynchronize(st));
    if (loss) MLP_HIP(n, hipMemcpy(loss, n->d_loss, sizeof(double), hipMemcpyDeviceToHost));
    if (grad) MLP_HIP(n, hipMemcpy(grad, n->g, sizeof(float) * n->np, hipMemcpyDeviceToDevice));
    return 0;
}

int uqk_mlp_forward(uqk_mlp* n, int64_t rows, const float* x, float* y) {
    if (!n || rows < 0) return -1;
    if (rows == 0) return 0;
    if (int rc = ensure_rows(n, rows)) return rc;
    hipStream_t st = n->ctx->stream;
    hipLaunchKernelGGL(to_bf16_kernel, dim3(blocks(rows * n->K0)), dim3(256), 0, st, x, n->act[0], rows * n->K0);
    if (int rc = forward(n, rows, false)) return rc;
    MLP_HIP(n, hipMemcpyAsync(y, n->y, sizeof(float) * rows * n->NO, hipMemcpyDeviceToDevice, st));
    MLP_HIP(n, hipStreamSynchronize(st));
    return 0;
}

int uqk_mlp_train(uqk_mlp* n, int64_t ncell, const double* u, const double* lambda, int32_t batch, int32_t steps,
                  double lr, uint64_t seed, double* mean_loss) {
    if (!n || ncell <= 0 || !u || !lambda || steps < 0) return -1;
    if (batch <= 0 || batch % kTile) return mfail(n, -1, "uqk_mlp_train: batch must be a positive multiple of 64");
    if (!n->normalised) {
        if (int rc = uqk_mlp_fit_normalization(n, ncell, u, lambda)) return rc;
    }
    if (int rc = ensure_rows(n, batch)) return rc;
    hipStream_t st = n->ctx->stream;
    MLP_HIP(n, hipMemsetAsync(n->d_loss, 0, sizeof(double), st));
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    for (int s = 0; s < steps; ++s) {
        ++n->adam_t;
        hipLaunchKernelGGL(rand_index_kernel, dim3(blocks(batch)), dim3(256), 0, st, n->idx, (long)batch, (long)ncell,
                           seed, (uint64_t)n->adam_t);
        if (int rc = launch_features(n, batch, n->idx, u, lambda, true, n->act[0], nullptr, n->t)) return rc;
        if (int rc = forward(n, batch, true)) return rc;
        if (int rc = backward(n, batch, n->S * n->M)) return rc;
        const double bc1 = 1.0 - std::pow((double)b1, (double)n->adam_t);
        const double bc2 = 1.0 - std::pow((double)b2, (double)n->adam_t);
        const float lr_t = (float)(lr * std::sqrt(bc2) / bc1);
        hipLaunchKernelGGL(adam_kernel, dim3(blocks(n->np)), dim3(256), 0, st, n->p, n->g, n->m, n->v, n->ph, n->np,
                           lr_t, b1, b2, eps);
    }
    MLP_HIP(n, hipGetLastError());
    MLP_HIP(n, hipStreamSynchronize(st));
    if (mean_loss) {
        double l = 0.0;
        MLP_HIP(n, hipMemcpy(&l, n->d_loss, sizeof(double), hipMemcpyDeviceToHost));
        *mean_loss = steps ? l / steps : 0.0;
    }
    return 0;
}

int uqk_mlp_predict(uqk_mlp* n, int64_t ncell, const double* u, double* lambda_out) {
    if (!n || ncell < 0) return -1;
    if (int rc = uqk::enqueue_predict(n, ncell, u, lambda_out, nullptr)) return rc;
    MLP_HIP(n, hipStreamSynchronize(n->ctx->stream));
    return 0;
}

int64_t uqk_warm_start_select(uqk_ctx* ctx, int64_t ncell, const double* u, double* lambda, const double* lambda_alt) {
    if (!ctx || ncell < 0) return -1;
    if (ncell == 0) return 0;
    if (int rc = uqk::enqueue_warm_select(ctx, ncell, u, lambda, lambda_alt, nullptr, nullptr)) return rc;
    unsigned long long h = 0;
    UQK_HIP_CTX(ctx, hipMemcpyAsync(&h, ctx->d_counters + 4, sizeof(h), hipMemcpyDeviceToHost, ctx->stream));
    UQK_HIP_CTX(ctx, hipStreamSynchronize(ctx->stream));
    return (int64_t)h;
}

}  // extern "C"
