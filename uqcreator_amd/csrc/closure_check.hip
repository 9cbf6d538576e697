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
// by one atomic per wave (ballot and prefix count); the gradients of the
// listed cells are stored for the Newton kernel, which starts from them.
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
        if (a.dt_cell) a.dt_cell[cs] = dtm;
    } else if (valid && a.grad_at_lam) {
        // the gradient the Newton kernel starts from
        double* gg = a.grad_at_lam + cs * N;
#pragma unroll
        for (int t = 0; t < N; ++t) gg[t] = acc[t] - ug[t];
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
    // counters: one gradient per cell (the listed cells' is their Newton
    // start); ComputeDt of finished cells
    const unsigned long long evaluated = __ballot(valid);
    double m = conv ? dtm : 1e10;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) m = fmin(m, __shfl_xor(m, off, 64));
    if (lane == 0) {
        if (evaluated) atomicAdd(a.nfail + 2, (unsigned long long)__popcll(evaluated));
        if (a.dt_min && m < 1e10) atomicMin(a.dt_min, (unsigned long long)__double_as_longlong(m));
    }
}

// The residual check of a run's later steps: the last step's closure left,
// for every cell, the moments <U(lambda phi) phiW> of the current lambda in
// u_out, uQ in uq_out and the cell's ComputeDt in dt_cell, so the gradient
// at lambda is u_out - u without any evaluation of U, and a cell that
// passes keeps those outputs.  One lane per cell, as closure_check_kernel,
// whose gradient it equals (the same differences, the same sum of squares).
// The host launches it for a run's steps after the first; the run's flag
// (set when a step has left those outputs) guards it.
template <int KIND, int M, int Q>
__global__ void __launch_bounds__(kCheckThreads) closure_recheck_kernel(NewtonArgs a) {
    constexpr int S = KindInfo<KIND>::S, N = S * M;
    if (a.active && *a.active == 0) return;
    if (!*a.recon_ok) return;
    const long cell = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = cell < a.ncell;
    const long cs = valid ? cell : 0;
    const ClosureParams& c = a.c;
    const double* ur = a.u_out + cs * N;
    const double* ug = a.u + cs * N;
    double g[N];
    double sq = 0.0;
#pragma unroll
    for (int t = 0; t < N; ++t) {
        g[t] = ur[t] - ug[t];
        sq += g[t] * g[t];
    }
    const bool conv = valid && sqrt(sq) < c.eps;
    if (conv) {
        if (a.status) a.status[cs] = UQK_STATUS_CONVERGED;
        if (a.iters) a.iters[cs] = 0;
    } else if (valid && a.grad_at_lam) {
        double* gg = a.grad_at_lam + cs * N;
#pragma unroll
        for (int t = 0; t < N; ++t) gg[t] = g[t];
    }
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
    // ComputeDt of the finished cells, as the last step found it
    double m = (conv && a.dt_min) ? a.dt_cell[cs] : 1e10;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) m = fmin(m, __shfl_xor(m, off, 64));
    if (lane == 0 && a.dt_min && m < 1e10) atomicMin(a.dt_min, (unsigned long long)__double_as_longlong(m));
}

// The same check for the workgroup kernel's sizes (n = S*M up to 128, any
// Q, per-cell orders): one wave per cell, the workgroup's waves strided over
// the cells.  Lanes over points evaluate Lambda (over the cell's nTotal), U
// and uQ; lanes over duals (i, s), i < nTotal, sum the projection over the
// points in order; a wave keeps the cells it leaves in LDS and appends them
// 64 at a time.  Tables (Q x M, k-major) in LDS.  With a learned start
// (a.lam_alt), a cell that fails at its lambda is also evaluated at the
// network's candidate, which replaces lambda when its ||grad L|| is smaller
// (a NaN norm never wins), as the wave kernel chooses; a candidate that
// passes finishes the cell.  The gradients of the listed cells at the duals
// they keep are stored for the workgroup kernel, which starts from them.  In
// a run's later steps (*recon_ok) the gradient at lambda comes from the last
// step's moments, as in closure_recheck_kernel.
template <int KIND>
__global__ void __launch_bounds__(512) closure_check_wave_kernel(NewtonArgs a) {
    constexpr int S = KindInfo<KIND>::S;
    extern __shared__ double smem[];
    if (a.active && *a.active == 0) return;
    const int M = a.M, Q = a.Q, SM = S * M;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
    double* tphi = smem;
    double* tphiw = tphi + Q * M;
    for (int t = threadIdx.x; t < Q * M; t += blockDim.x) {
        tphi[t] = a.phi[t];
        tphiw[t] = a.phiw[t];
    }
    double* lamw = tphiw + Q * M + (size_t)wave * (SM + S * Q);
    double* Ub = lamw + SM;
    int* pend = reinterpret_cast<int*>(tphiw + Q * M + (size_t)nw * (SM + S * Q)) + 64 * wave;
    __syncthreads();
    const ClosureParams& c = a.c;
    int np = 0;  // cells pending in this wave's buffer (wave-uniform)
    auto flush = [&]() {
        wave_sync();
        unsigned long long base = 0;
        if (lane == 0) base = atomicAdd(a.list_count, (unsigned long long)np);
        base = ((unsigned long long)__builtin_amdgcn_readfirstlane((unsigned)(base >> 32)) << 32) |
               (unsigned)__builtin_amdgcn_readfirstlane((unsigned)base);
        if (lane < np) a.list_out[base + lane] = pend[lane];
        np = 0;
        wave_sync();
    };
    unsigned long long w_done = 0, w_took = 0, w_alt = 0, w_eval = 0;
    double w_dt = 1e10;
    // a run's later steps: the last step left every cell's outputs at its
    // lambda (u_out = <U(lambda phi) phiW>, uQ, dt_cell), so the gradient at
    // lambda is u_out - u, and a cell that passes there keeps them
    const bool recon = a.recon_ok && *a.recon_ok;
    const long stride = (long)gridDim.x * nw;
    for (long cell = (long)blockIdx.x * nw + wave; cell < a.ncell; cell += stride) {
        const int nt = a.nt_cell ? a.nt_cell[cell] : M;
        double* lg = a.lam + cell * SM;
        const double* ug = a.u + cell * SM;
        double m = 1e10;
        const double dx = a.dt_min ? a.dx_param[cell] : 0.0;
        double acc[2] = {0.0, 0.0};  // this lane's duals e = lane, lane + 64 (n <= 128)
        // ||grad L|| at the duals X (the cell's nTotal of them); leaves U at
        // the points in Ub, this lane's projections in acc, min ComputeDt in m
        auto eval = [&](const double* X) -> double {
            for (int t = lane; t < S * nt; t += 64) lamw[t] = X[t];
            wave_sync();
            m = 1e10;
            for (int k = lane; k < Q; k += 64) {
                double L[S], U[S];
#pragma unroll
                for (int s = 0; s < S; ++s) L[s] = 0.0;
                for (int i = 0; i < nt; ++i) {
                    const double f = tphi[k * M + i];
#pragma unroll
                    for (int s = 0; s < S; ++s) L[s] += lamw[i * S + s] * f;
                }
                closure_U<KIND>(c, L, U);
#pragma unroll
                for (int s = 0; s < S; ++s) Ub[s * Q + k] = U[s];
                if (a.dt_min) m = fmin(m, point_dt<S>(a.dtp, U, dx));
            }
            wave_sync();
            double sq = 0.0;
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int e = lane + 64 * h;
                acc[h] = 0.0;
                if (e < S * nt) {
                    const int i = e / S, s = e - i * S;
                    double v = 0.0;
                    for (int k = 0; k < Q; ++k) v = fma(Ub[s * Q + k], tphiw[k * M + i], v);
                    acc[h] = v;
                    const double g = v - ug[e];
                    sq += g * g;
                }
            }
            const double nrm = sqrt(wave_sum(sq));
            wave_sync();
            return nrm;
        };
        // ||grad L|| at lambda from the last step's moments (recon)
        auto from_outputs = [&]() -> double {
            const double* uo = a.u_out + cell * SM;
            double sq = 0.0;
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int e = lane + 64 * h;
                acc[h] = 0.0;
                if (e < S * nt) {
                    const double v = uo[e];
                    acc[h] = v;
                    const double g = v - ug[e];
                    sq += g * g;
                }
            }
            return sqrt(wave_sum(sq));
        };
        double ng = recon ? from_outputs() : eval(lg);
        if (!recon) ++w_eval;
        bool kept = recon;  // the cell's outputs stay the last step's
        double gk[2] = {acc[0], acc[1]};  // <U phiW> at the duals the cell keeps
        if (!(ng < c.eps) && a.lam_alt) {
            const double* la = a.lam_alt + cell * SM;
            const double nalt = eval(la);
            ++w_alt;
            if (nalt < ng) {
                for (int t = lane; t < SM; t += 64) lg[t] = la[t];
                ng = nalt;
                ++w_took;
                kept = false;
                gk[0] = acc[0];
                gk[1] = acc[1];
            }
        }
        if (!(ng < c.eps)) {
            if (a.grad_at_lam) {
                // the gradient the Newton kernel starts from
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    const int e = lane + 64 * h;
                    if (e < S * nt) a.grad_at_lam[cell * SM + e] = gk[h] - ug[e];
                }
            }
            if (lane == 0) pend[np] = (int)cell;
            ++np;
            if (np == 64) flush();
            continue;
        }
        ++w_done;
        if (kept) {
            if (a.dt_min) w_dt = fmin(w_dt, a.dt_cell[cell]);
        } else {
            double* uqg = a.uq_out + cell * (long)Q * S;
            for (int t = lane; t < Q * S; t += 64) {
                const int k = t / S, s = t - k * S;
                uqg[t] = Ub[s * Q + k];
            }
            double* uo = a.u_out + cell * SM;
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int e = lane + 64 * h;
                if (e < SM) uo[e] = e < S * nt ? acc[h] : 0.0;
            }
            if (a.dt_min) {
                const double mc = wave_min(m);
                w_dt = fmin(w_dt, mc);
                if (a.dt_cell && lane == 0) a.dt_cell[cell] = mc;
            }
        }
        if (lane == 0) {
            if (a.status) a.status[cell] = UQK_STATUS_CONVERGED;
            if (a.iters) a.iters[cell] = 0;
        }
        wave_sync();
    }
    if (np > 0) flush();
    if (lane == 0) {
        // gradients evaluated: at lambda (not from the last step's moments)
        // and at every candidate
        if (w_eval + w_alt) atomicAdd(a.nfail + 2, w_eval + w_alt);
        if (w_took && a.took) atomicAdd(a.took, w_took);
        if (a.dt_min && w_dt < 1e10) atomicMin(a.dt_min, (unsigned long long)__double_as_longlong(w_dt));
    }
}

template <int KIND, int M, int Q>
static bool check_try(int kind, int m, int q, const NewtonArgs& a, int num_cu, hipStream_t st, hipError_t* err) {
    if (kind != KIND || m != M || q != Q) return false;
    (void)num_cu;
    const long blocks = (a.ncell + kCheckThreads - 1) / kCheckThreads;
    const dim3 grid((unsigned)(blocks > 0 ? blocks : 1));
    // a run's later steps (recon_ok): the check from the last step's outputs
    if (a.recon_ok)
        hipLaunchKernelGGL((closure_recheck_kernel<KIND, M, Q>), grid, dim3(kCheckThreads), 0, st, a);
    else
        hipLaunchKernelGGL((closure_check_kernel<KIND, M, Q>), grid, dim3(kCheckThreads), 0, st, a);
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

size_t closure_check_wave_lds(int S, int M, int Q, int waves) {
    return sizeof(double) * (2 * (size_t)Q * M + (size_t)waves * ((size_t)S * M + (size_t)S * Q)) +
           sizeof(int) * 64 * (size_t)waves;
}

template <int KIND>
static hipError_t check_wave_go(const NewtonArgs& a, int max_lds, int num_cu, hipStream_t st) {
    constexpr int S = KindInfo<KIND>::S;
    if (S * a.M > 128) return hipErrorInvalidValue;
    int waves = 8;
    while (waves > 1 && closure_check_wave_lds(S, a.M, a.Q, waves) > (size_t)max_lds) waves /= 2;
    const size_t shmem = closure_check_wave_lds(S, a.M, a.Q, waves);
    if (shmem > (size_t)max_lds) return hipErrorInvalidValue;
    int occ = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, closure_check_wave_kernel<KIND>, 64 * waves, shmem) !=
            hipSuccess ||
        occ <= 0)
        occ = 1;
    const long want = (a.ncell + waves - 1) / waves;
    const long cap = (long)occ * (num_cu > 0 ? num_cu : 1);
    const unsigned blocks = (unsigned)(want < cap ? (want > 0 ? want : 1) : cap);
    hipLaunchKernelGGL(closure_check_wave_kernel<KIND>, dim3(blocks), dim3(64 * waves), shmem, st, a);
    return hipGetLastError();
}

bool closure_check_wave_fits(int S, int M, int Q, int max_lds) {
    return S * M <= 128 && closure_check_wave_lds(S, M, Q, 1) <= (size_t)max_lds;
}

hipError_t closure_check_wave_launch(int kind, const NewtonArgs& a, int max_lds, int num_cu, hipStream_t st) {
    switch (kind) {
        case UQK_BOUNDED_BARRIER: return check_wave_go<UQK_BOUNDED_BARRIER>(a, max_lds, num_cu, st);
        case UQK_EULER1D: return check_wave_go<UQK_EULER1D>(a, max_lds, num_cu, st);
        case UQK_EULER2D: return check_wave_go<UQK_EULER2D>(a, max_lds, num_cu, st);
        case UQK_SHALLOW_WATER2D: return check_wave_go<UQK_SHALLOW_WATER2D>(a, max_lds, num_cu, st);
        case UQK_SHALLOW_WATER2D_BARRIER: return check_wave_go<UQK_SHALLOW_WATER2D_BARRIER>(a, max_lds, num_cu, st);
    }
    return hipErrorInvalidValue;
}

}  // namespace uqk
