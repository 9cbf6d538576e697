// closure_check.hip -- the residual check that precedes the Newton kernel
// in the time loop.  Closure::SolveClosure (closure.cpp:133-193) first
// evaluates ||grad L||_2 at the cell's lambda and stops when it is below
// epsilon; in a time loop about half of the cells end there, and their
// outputs need only that gradient.  This kernel evaluates it for every cell
// and finishes the cells that pass -- status CONVERGED, no Newton step,
// uQ = U(lambda phi), the moments <U phiW> and their ComputeDt -- and lists
// the others for the wave kernel.
//
// One lane per cell (64 cells per wave): the work of a cell is a chain of
// independent point evaluations, so lanes need no exchange and the tables
// are LDS broadcasts.  Per point k: Lambda_k = sum_i lambda_i phi_i(k) (the
// wave kernel's order), U(Lambda_k) -> uQ (written for every cell; the
// Newton kernel rewrites the cells it takes), ComputeDt there, and the
// projection sum_k U(k) phiW(k, i) accumulated in k order.  The list grows
// by one atomic per wave (ballot and prefix count).
#include <hip/hip_runtime.h>

#include "uqk_device.h"
#include "uqk_internal.h"

namespace uqk {

constexpr int kCheckThreads = 256;

template <int KIND, int M, int Q>
__global__ void __launch_bounds__(kCheckThreads) closure_check_kernel(NewtonArgs a) {
    constexpr int S = KindInfo<KIND>::S, N = S * M;
    __shared__ double tphi[Q * M];   // phi(k, i) at k * M + i
    __shared__ double tphiw[Q * M];
    if (a.active && *a.active == 0) return;
    for (int t = threadIdx.x; t < Q * M; t += blockDim.x) {
        tphi[t] = a.phi[t];
        tphiw[t] = a.phiw[t];
    }
    __syncthreads();
    const long cell = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = cell < a.ncell;
    const long cs = valid ? cell : 0;  // (invalid lanes compute on cell 0, write nothing)
    const ClosureParams& c = a.c;
    double lam[N];  // lambda(i, s) at i * S + s
    const double* lg = a.lam + cs * N;
#pragma unroll
    for (int t = 0; t < N; ++t) lam[t] = lg[t];
    double acc[N];  // <U phiW>(i, s) at i * S + s
#pragma unroll
    for (int t = 0; t < N; ++t) acc[t] = 0.0;
    const double dx = a.dt_min ? a.dx_param[cs] : 0.0;
    double dtm = 1e10;
    double* uqg = a.uq_out + cs * (long)Q * S;
#pragma unroll 1
    for (int k = 0; k < Q; ++k) {
        double L[S], U[S];
#pragma unroll
        for (int s = 0; s < S; ++s) L[s] = 0.0;
#pragma unroll
        for (int i = 0; i < M; ++i) {
            const double f = tphi[k * M + i];
#pragma unroll
            for (int s = 0; s < S; ++s) L[s] += lam[i * S + s] * f;
        }
        closure_U<KIND>(c, L, U);
        if (valid) {
#pragma unroll
            for (int s = 0; s < S; ++s) uqg[k * S + s] = U[s];
        }
        if (a.dt_min) dtm = fmin(dtm, point_dt<S>(a.dtp, U, dx));
#pragma unroll
        for (int i = 0; i < M; ++i) {
            const double w = tphiw[k * M + i];
#pragma unroll
            for (int s = 0; s < S; ++s) acc[i * S + s] = fma(U[s], w, acc[i * S + s]);
        }
    }
    const double* ug = a.u + cs * N;
    double sq = 0.0;
#pragma unroll
    for (int t = 0; t < N; ++t) {
        const double g = acc[t] - ug[t];
        sq += g * g;
    }
    const bool conv = valid && sqrt(sq) < c.eps;
    if (conv) {
        double* uo = a.u_out + cs * N;
#pragma unroll
        for (int t = 0; t < N; ++t) uo[t] = acc[t];
        if (a.status) a.status[cs] = UQK_STATUS_CONVERGED;
        if (a.iters) a.iters[cs] = 0;
    }
    // the others to the list: one atomic per wave
    const unsigned long long left = __ballot(valid && !conv);
    const int lane = threadIdx.x & 63;
    unsigned long long base = 0;
    if (lane == 0 && left) base = atomicAdd(a.list_count, (unsigned long long)__popcll(left));
    base = ((unsigned long long)__builtin_amdgcn_readfirstlane((unsigned)(base >> 32)) << 32) |
           (unsigned)__builtin_amdgcn_readfirstlane((unsigned)base);
    if (valid && !conv) {
        const unsigned long long below = lane ? (left & (~0ull >> (64 - lane))) : 0ull;
        a.list_out[base + __popcll(below)] = (int)cell;
    }
    // counters: one gradient per finished cell; ComputeDt of finished cells
    const unsigned long long done = __ballot(conv);
    double m = conv ? dtm : 1e10;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) m = fmin(m, __shfl_xor(m, off, 64));
    if (lane == 0) {
        if (done) atomicAdd(a.nfail + 2, (unsigned long long)__popcll(done));
        if (a.dt_min && m < 1e10) atomicMin(a.dt_min, (unsigned long long)__double_as_longlong(m));
    }
}

template <int KIND, int M, int Q>
static bool check_try(int kind, int m, int q, const NewtonArgs& a, int num_cu, hipStream_t st, hipError_t* err) {
    if (kind != KIND || m != M || q != Q) return false;
    (void)num_cu;
    const long blocks = (a.ncell + kCheckThreads - 1) / kCheckThreads;
    hipLaunchKernelGGL((closure_check_kernel<KIND, M, Q>), dim3((unsigned)(blocks > 0 ? blocks : 1)),
                       dim3(kCheckThreads), 0, st, a);
    *err = hipGetLastError();
    return true;
}

bool closure_check_launch(int kind, int M, int Q, const NewtonArgs& a, int num_cu, hipStream_t st, hipError_t* err) {
    return check_try<UQK_BOUNDED_BARRIER, 6, 12>(kind, M, Q, a, num_cu, st, err) ||
           check_try<UQK_EULER1D, 15, 49>(kind, M, Q, a, num_cu, st, err) ||
           check_try<UQK_EULER2D, 10, 25>(kind, M, Q, a, num_cu, st, err) ||
           check_try<UQK_SHALLOW_WATER2D, 10, 25>(kind, M, Q, a, num_cu, st, err) ||
           check_try<UQK_SHALLOW_WATER2D_BARRIER, 10, 25>(kind, M, Q, a, num_cu, st, err);
}

}  // namespace uqk
