// sweeps.hip -- HBM-bound sweeps of the moment solver's step body:
//   u(lambda) at the quadrature points   (momentsolver.cpp:99-101)
//   moment reconstruction u = uQ phiw    (momentsolver.cpp:120-124)
//   CFL time step, global min            (momentsolver.cpp:127-133)
//   kinetic flux + explicit Euler update (expliciteuler.cpp:8-75) with the
//     residual of the state-0 mean fused in (momentsolver.cpp:149-152)
//   mean / variance of the solution      (momentsolver.cpp:302-328)
#include <hip/hip_runtime.h>

#include "uqk_device.h"
#include "uqk_internal.h"

namespace uqk {

typedef double f64x4 __attribute__((ext_vector_type(4)));

constexpr int kSweepThreads = 256;
constexpr int kPartialBlocks = 8192;  // grid of the grid-stride sweeps and size of the partials

int sweep_partial_blocks(long) { return kPartialBlocks; }

// One thread per (cell, quadrature point): uq(s,k) = U_mat(sum_i lambda(s,i) phi(k,i)).
#define UQK_SKIP_IF_INACTIVE(a) \
    if ((a).active && *(a).active == 0) return;

template <int KIND>
__global__ void __launch_bounds__(kSweepThreads) eval_uq_kernel(SweepArgs a) {
    constexpr int S = KindInfo<KIND>::S;
    UQK_SKIP_IF_INACTIVE(a)
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.ncell * a.Q) return;
    long cell = t / a.Q;
    int k = (int)(t - cell * a.Q);
    int nt = a.nt_cell ? a.nt_cell[cell] : a.M;
    const double* lam = a.lam + cell * S * a.M;
    double L[S], Uv[S];
#pragma unroll
    for (int s = 0; s < S; ++s) L[s] = 0.0;
    for (int i = 0; i < nt; ++i) {
        double p = a.phi[k * a.M + i];
#pragma unroll
        for (int s = 0; s < S; ++s) L[s] += lam[i * S + s] * p;
    }
    closure_U_mat<KIND>(a.c, L, Uv);
    double* out = a.uq_out + t * S;
#pragma unroll
    for (int s = 0; s < S; ++s) out[s] = Uv[s];
}

// One thread per output moment (cell, i, s): u(s,i) = sum_k uq(s,k) phiw(k,i) for i < nt.
__global__ void __launch_bounds__(kSweepThreads) reconstruct_kernel(SweepArgs a) {
    UQK_SKIP_IF_INACTIVE(a)
    const int S = a.S, M = a.M, Q = a.Q;
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.ncell * S * M) return;
    long cell = t / (S * M);
    int e = (int)(t - cell * S * M);
    int i = e / S, s = e - i * S;
    int nt = a.nt_cell ? a.nt_cell[cell] : M;
    double acc = 0.0;
    if (i < nt) {
        const double* q = a.uq + cell * S * Q;
        for (int k = 0; k < Q; ++k) acc += q[k * S + s] * a.phiw[k * M + i];
    }
    a.u_out[t] = acc;
}

// MomentSolver refinement (momentsolver.cpp:104-117) with ComputeRefIndicator
// (:271-289), one thread per cell.  The indicator of a cell at level l is
// |u(0, nt-1)| + |u(0, nt-2)| for one uncertain dimension, else the sum of
// |u(0, i)| over the moments the level adds (i from nTotal(l-1), or from 1 at
// level 0).  Above refine_above the cell moves up a level, below
// coarsen_below down, and never above max_level (the cap of refinement
// retardation).  A cell whose nt_cell is no level's nTotal is left alone.
__global__ void __launch_bounds__(kSweepThreads) refine_kernel(RefineArgs a) {
    UQK_SKIP_IF_INACTIVE(a)
    const long j = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= a.ncell) return;
    const int nt = a.nt_cell[j];
    int lev = -1;
    for (int l = 0; l < a.nlev; ++l)
        if (a.nt_level[l] == nt) lev = l;
    if (lev < 0) return;
    const double* u = a.u + j * a.S * a.M;
    double ind = 0.0;
    if (a.dim_xi == 1) {
        ind = fabs(u[(nt - 1) * a.S]);
        if (nt >= 2) ind += fabs(u[(nt - 2) * a.S]);
    } else {
        for (int i = lev == 0 ? 1 : a.nt_level[lev - 1]; i < nt; ++i) ind += fabs(u[i * a.S]);
    }
    if (ind > a.refine_above && lev < a.nlev - 1)
        ++lev;
    else if (ind < a.coarsen_below && lev > 0)
        --lev;
    if (lev > a.max_level) lev = a.max_level;
    a.nt_cell[j] = a.nt_level[lev];
}

__device__ __forceinline__ double block_min(double v) {
    __shared__ double red[kSweepThreads / 64];
    v = wave_min(v);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    double r = red[0];
    for (int w = 1; w < (int)(blockDim.x >> 6); ++w) r = fmin(r, red[w]);
    __syncthreads();
    return r;
}

__device__ __forceinline__ double block_sum(double v) {
    __shared__ double red[kSweepThreads / 64];
    v = wave_sum(v);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    double r = red[0];
    for (int w = 1; w < (int)(blockDim.x >> 6); ++w) r += red[w];
    __syncthreads();
    return r;
}

// Per-block minimum of ComputeDt over cells (grid-stride), then a one-block finish.
template <int KIND>
__global__ void __launch_bounds__(kSweepThreads) dt_partial_kernel(SweepArgs a) {
    constexpr int S = KindInfo<KIND>::S;
    UQK_SKIP_IF_INACTIVE(a)
    double m = 1e10;
    for (long cell = (long)blockIdx.x * blockDim.x + threadIdx.x; cell < a.ncell; cell += (long)gridDim.x * blockDim.x) {
        const double* q = a.uq + cell * S * a.Q;
        double dx = a.dx_param[cell];
        for (int k = 0; k < a.Q; ++k) m = fmin(m, point_dt<S>(a.p, q + k * S, dx));
    }
    m = block_min(m);
    if (threadIdx.x == 0) a.partial[blockIdx.x] = m;
}

__global__ void __launch_bounds__(kSweepThreads) min_finish_kernel(SweepArgs a) {
    UQK_SKIP_IF_INACTIVE(a)
    double m = 1e10;
    for (int t = threadIdx.x; t < a.npartial; t += blockDim.x) m = fmin(m, a.partial[t]);
    m = block_min(m);
    if (threadIdx.x == 0) *a.scalar_out = m;
}

__global__ void __launch_bounds__(kSweepThreads) sum_finish_kernel(SweepArgs a) {
    UQK_SKIP_IF_INACTIVE(a)
    double m = 0.0;
    for (int t = threadIdx.x; t < a.npartial; t += blockDim.x) m += a.partial[t];
    m = block_sum(m);
    if (threadIdx.x == 0) *a.scalar_out = m;
}

// Cells per wavefront of the advance: lanes take (cell slot, quadrature
// point) for Q <= 32, as many cells as 64 lanes and the 16 rows of one MFMA
// tile (cells x states) allow; one cell with ceil(Q/64) points per lane else.
__host__ __device__ inline int advance_cells_per_wave(int S, int Q) {
    if (Q > 32) return 1;
    const int by_lanes = 64 / Q, by_rows = 16 / S;
    return by_lanes < by_rows ? by_lanes : by_rows;
}

// LDS doubles per wavefront of the advance: the edge-summed point fluxes of
// its cells, (cell slot, point, state).
__host__ __device__ inline int advance_wave_doubles(int S, int Q) {
    return advance_cells_per_wave(S, Q) * Q * S;
}

// ExplicitEuler::Advance (expliciteuler.cpp:8-75) with numFlux
// (momentsolver.cpp:413-416) for a wavefront's cells: each lane owns a
// quadrature point of one cell and sums the point fluxes of all the cell's
// edges in registers, in edge order (the projection onto phiw is linear, so
// the edges share one projection); the shallow-water bed friction enters as
// -area * S(u) at the point before the projection; the projection
// rhs(s, i) = sum_k G(s, k) phiW(k, i) of all the wave's cells is one FP64
// MFMA tile row block (rows: cell slot x state), and the lanes holding the
// tile's elements update the moments.  The residual |u_new(0,0) - u_old(0,0)|
// * area is accumulated per block.
// (four waves per SIMD: the kernel waits on memory and square roots, and the
// register budget of four waves holds it without spilling)
template <int KIND>
__global__ void __launch_bounds__(kSweepThreads) __attribute__((amdgpu_waves_per_eu(4))) advance_kernel(SweepArgs a) {
    constexpr int S = KindInfo<KIND>::S;
    UQK_SKIP_IF_INACTIVE(a)
    extern __shared__ double smem[];
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int M = a.M, Q = a.Q, SM = S * M, E = a.nedge, D = a.dim;
    const int CPW = advance_cells_per_wave(S, Q);
    double* Gs = smem + wave * advance_wave_doubles(S, Q);
    const bool wlds = a.table_lds != 0;
    double* wt = smem + (blockDim.x >> 6) * advance_wave_doubles(S, Q);
    if (wlds) {
        for (int t = threadIdx.x; t < Q * M; t += blockDim.x) wt[t] = a.phiw[t];
        __syncthreads();
    }
    const double* phiw = wlds ? wt : a.phiw;
    const double dt = a.dt_dev ? *a.dt_dev : a.dt;
    const bool fric = (S == 3 && a.p.problem == UQK_SHALLOW_WATER2D_PROBLEM && a.p.friction > 0.0);
    // this lane's cell slot and first point
    const int slot = CPW > 1 ? lane / Q : 0;
    const int k0 = CPW > 1 ? lane - slot * Q : lane;
    const bool lane_on = slot < CPW && k0 < Q;
    double res = 0.0;
    const long nwaves = (long)gridDim.x * (blockDim.x >> 6);
    for (long cell0 = ((long)blockIdx.x * (blockDim.x >> 6) + wave) * CPW; cell0 < a.ncell; cell0 += nwaves * CPW) {
        const long cell = cell0 + slot;
        const bool on = lane_on && cell < a.ncell;
        const int bc = on ? a.bc[cell] : UQK_BC_DIRICHLET;
        if (on && bc != UQK_BC_DIRICHLET) {
            const double* uqj = a.uq + cell * S * Q;
            const bool wall = (bc == UQK_BC_NOSLIP || bc == UQK_BC_SWWALL);
            const double area = a.area[cell];
            for (int k = k0; k < Q; k += 64) {
                double uj[S], gj[S], Gsum[S];
#pragma unroll
                for (int s = 0; s < S; ++s) {
                    uj[s] = uqj[k * S + s];
                    Gsum[s] = 0.0;
                }
                if constexpr (S >= 3) {
                    if (wall) {
                        // ghost state: velocity reflected to the tangential
                        // part (expliciteuler.cpp:25-58)
                        const double bx = a.bnormal[cell * D + 0], by = a.bnormal[cell * D + 1];
                        const double rho = uj[0], vx = uj[1] / rho, vy = uj[2] / rho;
                        const double vn = bx * vx + by * vy;
                        const double tx = vx - vn * bx, ty = vy - vn * by;
#pragma unroll
                        for (int s = 0; s < S; ++s) gj[s] = uj[s];
                        if constexpr (S == 4) {
                            if (bc == UQK_BC_NOSLIP) gj[3] += rho * 0.5 * ((tx * tx + ty * ty) - (vx * vx + vy * vy));
                        }
                        gj[1] = rho * tx;
                        gj[2] = rho * ty;
                    }
                }
                for (int e = 0; e < E; ++e) {
                    const int nb = a.nbr[cell * E + e];
                    const double* nu = a.unit_normal + (cell * E + e) * D;
                    const double* nv = a.normal + (cell * E + e) * D;
                    const double nux = nu[0], nuy = D > 1 ? nu[1] : 0.0, nx = nv[0], ny = D > 1 ? nv[1] : 0.0;
                    double l[S], r[S], G[S];
                    if (wall && nb == a.ncell) {
                        // the flux is G(ghost, ghost)
#pragma unroll
                        for (int s = 0; s < S; ++s) l[s] = r[s] = S >= 3 ? gj[s] : uj[s];
                    } else if (nb == a.ncell) {
                        // the reference reads the never-written uQ[nCells] (zero)
#pragma unroll
                        for (int s = 0; s < S; ++s) {
                            l[s] = uj[s];
                            r[s] = 0.0;
                        }
                    } else {
#pragma unroll
                        for (int s = 0; s < S; ++s) {
                            l[s] = uj[s];
                            r[s] = a.uq[((long)nb * Q + k) * S + s];
                        }
                    }
                    point_flux<S>(a.p, l, r, nux, nuy, nx, ny, G);
#pragma unroll
                    for (int s = 0; s < S; ++s) Gsum[s] += G[s];
                }
                if (fric) {
                    // bed friction S_s = -cf |v| q_s / h (s = 1, 2) at this point
                    const double cf =
                        a.p.friction * (1.0 + (a.p.friction_dim >= 0 ? a.xi[k * a.d + a.p.friction_dim] : 0.0));
                    const double sp = sqrt(uj[1] * uj[1] + uj[2] * uj[2]) / uj[0];
#pragma unroll
                    for (int s = 1; s < S; ++s) Gsum[s] += area * (cf * sp * uj[s] / uj[0]);
                }
#pragma unroll
                for (int s = 0; s < S; ++s) Gs[(slot * Q + k) * S + s] = Gsum[s];
            }
        }
        wave_sync();
        // projection and update: D = A B with A(r, k) = G(cell slot r / S,
        // state r % S, point k) and B(k, i) = phiW(k, i), 16 x 16 x 4 FP64
        // MFMA steps (lane l: row/column l % 16, k = 4 ks + l / 16); element e
        // of lane l is D(4 e + l / 16, l % 16)
        const int r16 = lane & 15, kq4 = lane >> 4;
        const int ks_n = (Q + 3) >> 2;
        const int ar_slot = r16 / S, ar_s = r16 - ar_slot * S;
        const long ar_cell = cell0 + ar_slot;
        const bool ar_on = ar_slot < CPW && ar_cell < a.ncell && a.bc[ar_cell] != UQK_BC_DIRICHLET;
        for (int ti = 0; 16 * ti < M; ++ti) {
            const int ic = 16 * ti + r16;
            f64x4 pacc = {0.0, 0.0, 0.0, 0.0};
            for (int ks = 0; ks < ks_n; ++ks) {
                const int k = 4 * ks + kq4;
                const double av = (ar_on && k < Q) ? Gs[(ar_slot * Q + k) * S + ar_s] : 0.0;
                const double bv = (ic < M && k < Q) ? phiw[k * M + ic] : 0.0;
                pacc = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, pacc, 0, 0, 0);
            }
            if (ic < M) {
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const int row = 4 * e + kq4;
                    const int cs = row / S, st = row - cs * S;
                    const long c = cell0 + cs;
                    if (cs < CPW && c < a.ncell) {
                        const int nt = a.nt_cell ? a.nt_cell[c] : M;
                        const int bcc = a.bc[c];
                        const int f = ic * S + st;
                        const double us = a.u[c * SM + f];  // reconstructed moments
                        double* un = a.u_out + c * SM;       // in: previous moments, out: new moments
                        const double old = un[f];
                        double v = us;
                        if (bcc != UQK_BC_DIRICHLET && ic < nt) v = us - (dt / a.area[c]) * pacc[e];
                        un[f] = v;
                        if (f == 0) res += fabs(v - old) * a.area[c];
                    }
                }
            }
        }
        wave_sync();
    }
    __syncthreads();
    res = block_sum(res);
    if (threadIdx.x == 0) a.partial[blockIdx.x] = res;
}

// E and Var of each state from the closure at lambda (WriteMeanAndVar,
// momentsolver.cpp:302-328): one wavefront per cell, lanes over quadrature
// points (up to 128 per cell), the two weighted sums as wave reductions.  The
// cell's duals are read by all lanes at once (one cache line, broadcast).
template <int KIND>
__global__ void __launch_bounds__(kSweepThreads) mean_var_kernel(SweepArgs a) {
    constexpr int S = KindInfo<KIND>::S;
    constexpr int KP = 2;  // quadrature points per lane
    const int lane = threadIdx.x & 63;
    const long nwaves = (long)gridDim.x * (blockDim.x >> 6);
    for (long cell = (long)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); cell < a.ncell; cell += nwaves) {
        const int nt = a.nt_cell ? a.nt_cell[cell] : a.M;
        const double* lam = a.lam + cell * S * a.M;
        double Uv[KP][S], w0[KP];
#pragma unroll
        for (int h = 0; h < KP; ++h) {
            const int k = lane + 64 * h;
            double L[S];
#pragma unroll
            for (int s = 0; s < S; ++s) L[s] = 0.0;
            w0[h] = 0.0;
            if (k < a.Q) {
                for (int i = 0; i < nt; ++i) {
                    const double p = a.phi[k * a.M + i];
#pragma unroll
                    for (int s = 0; s < S; ++s) L[s] += lam[i * S + s] * p;
                }
                w0[h] = a.phiw[k * a.M + 0];
            }
            closure_U<KIND>(a.c, L, Uv[h]);
            if (k >= a.Q) {
#pragma unroll
                for (int s = 0; s < S; ++s) Uv[h][s] = 0.0;  // U of Lambda = 0 may be NaN
            }
        }
        double E[S];
#pragma unroll
        for (int s = 0; s < S; ++s) {
            double e = 0.0;
#pragma unroll
            for (int h = 0; h < KP; ++h) e += Uv[h][s] * w0[h];
            E[s] = wave_sum(e);
        }
        double* out = a.mv_out + cell * 2 * S;
#pragma unroll
        for (int s = 0; s < S; ++s) {
            double v = 0.0;
#pragma unroll
            for (int h = 0; h < KP; ++h) v += (Uv[h][s] - E[s]) * (Uv[h][s] - E[s]) * w0[h];
            v = wave_sum(v);
            if (lane == 0) {
                out[s] = E[s];
                out[S + s] = v;
            }
        }
    }
}

// t < t_end and steps < max_steps and no failed cell: the step runs
// The end of a replayed step: residual and t (momentsolver.cpp:134-152), run
// totals, and what the next step's launches need (zero closure counters, the
// start value 1e10 of ComputeDt's MIN, the active flag; set_recon: the
// step's closure left its outputs at the current lambda for every cell).
__global__ void __launch_bounds__(kSweepThreads) step_end_kernel(RunState* rs, double* scalars,
                                                                 unsigned long long* counters,
                                                                 const double* partial, int npartial, int set_recon) {
    if (!rs->active) return;
    double m = 0.0;
    for (int t = threadIdx.x; t < npartial; t += blockDim.x) m += partial[t];
    m = block_sum(m);
    if (threadIdx.x == 0) {
        const double t = rs->t + scalars[0];
        const long long steps = rs->steps + 1;
        const long long nfail = rs->nfail + (long long)counters[1];
        rs->t = t;
        rs->residual = m;
        rs->steps = steps;
        rs->nfail = nfail;
        rs->newton += (long long)counters[2];
        rs->gradients += (long long)counters[3];
        rs->factorizations += (long long)counters[5];
        rs->reuse_steps += (long long)counters[6];
#pragma unroll
        for (int i = 0; i < kClosureCounters; ++i) counters[i] = 0ull;
        scalars[0] = 1e10;
        rs->active = (t < rs->t_end && (rs->max_steps < 0 || steps < rs->max_steps) && nfail == 0) ? 1 : 0;
        if (set_recon) rs->recon = 1;
    }
}

hipError_t launch_step_end(RunState* rs, double* scalars, unsigned long long* counters, const double* partial,
                           int set_recon, int npartial, hipStream_t st) {
    hipLaunchKernelGGL(step_end_kernel, dim3(1), dim3(kSweepThreads), 0, st, rs, scalars, counters, partial,
                       npartial, set_recon);
    return hipGetLastError();
}

__global__ void fill_u64_kernel(unsigned long long* p, int n) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) p[t] = 0ull;
}

__global__ void fill_f64_kernel(double* p, int n, double v) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) p[t] = v;
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
#define UQK_DISPATCH(kind, FN, ...)                                                                     \
    switch (kind) {                                                                                     \
        case UQK_BOUNDED_BARRIER: FN<UQK_BOUNDED_BARRIER>(__VA_ARGS__); break;                         \
        case UQK_EULER1D: FN<UQK_EULER1D>(__VA_ARGS__); break;                                         \
        case UQK_EULER2D: FN<UQK_EULER2D>(__VA_ARGS__); break;                                         \
        case UQK_SHALLOW_WATER2D: FN<UQK_SHALLOW_WATER2D>(__VA_ARGS__); break;                         \
        case UQK_SHALLOW_WATER2D_BARRIER: FN<UQK_SHALLOW_WATER2D_BARRIER>(__VA_ARGS__); break;         \
        default: return hipErrorInvalidValue;                                                           \
    }

template <int KIND> static void go_eval_uq(const SweepArgs& a, hipStream_t st) {
    long total = a.ncell * a.Q;
    int blocks = (int)((total + kSweepThreads - 1) / kSweepThreads);
    hipLaunchKernelGGL(eval_uq_kernel<KIND>, dim3(blocks), dim3(kSweepThreads), 0, st, a);
}
hipError_t launch_eval_uq(int kind, const SweepArgs& a, hipStream_t st) {
    if (a.ncell <= 0) return hipSuccess;
    UQK_DISPATCH(kind, go_eval_uq, a, st);
    return hipGetLastError();
}

hipError_t launch_refine(const RefineArgs& a, hipStream_t st) {
    if (a.ncell <= 0) return hipSuccess;
    const int blocks = (int)((a.ncell + kSweepThreads - 1) / kSweepThreads);
    hipLaunchKernelGGL(refine_kernel, dim3(blocks), dim3(kSweepThreads), 0, st, a);
    return hipGetLastError();
}

hipError_t launch_reconstruct(const SweepArgs& a, hipStream_t st) {
    if (a.ncell <= 0) return hipSuccess;
    long total = a.ncell * a.S * a.M;
    int blocks = (int)((total + kSweepThreads - 1) / kSweepThreads);
    hipLaunchKernelGGL(reconstruct_kernel, dim3(blocks), dim3(kSweepThreads), 0, st, a);
    return hipGetLastError();
}

template <int KIND> static void go_dt(const SweepArgs& a, hipStream_t st) {
    hipLaunchKernelGGL(dt_partial_kernel<KIND>, dim3(kPartialBlocks), dim3(kSweepThreads), 0, st, a);
}
hipError_t launch_dt(int kind, const SweepArgs& a, hipStream_t st) {
    UQK_DISPATCH(kind, go_dt, a, st);
    SweepArgs b = a;
    b.npartial = kPartialBlocks;
    hipLaunchKernelGGL(min_finish_kernel, dim3(1), dim3(kSweepThreads), 0, st, b);
    return hipGetLastError();
}

template <int KIND> static void go_advance(const SweepArgs& a, hipStream_t st) {
    constexpr int S = KindInfo<KIND>::S;
    size_t shmem = sizeof(double) * (kSweepThreads / 64) * advance_wave_doubles(S, a.Q);
    SweepArgs b = a;
    // PhiW goes to LDS while a workgroup stays within 64 KB
    b.table_lds = shmem + sizeof(double) * a.Q * a.M <= 65536;
    if (b.table_lds) shmem += sizeof(double) * a.Q * a.M;
    hipLaunchKernelGGL(advance_kernel<KIND>, dim3(kPartialBlocks), dim3(kSweepThreads), shmem, st, b);
}
hipError_t launch_advance(int kind, const SweepArgs& a, hipStream_t st) {
    UQK_DISPATCH(kind, go_advance, a, st);
    return hipGetLastError();
}

hipError_t launch_residual(const SweepArgs& a, hipStream_t st) {
    SweepArgs b = a;
    b.npartial = kPartialBlocks;
    hipLaunchKernelGGL(sum_finish_kernel, dim3(1), dim3(kSweepThreads), 0, st, b);
    return hipGetLastError();
}

template <int KIND> static void go_mean_var(const SweepArgs& a, hipStream_t st) {
    const long waves = kSweepThreads / 64;
    const long want = (a.ncell + waves - 1) / waves;
    const int blocks = (int)(want < kPartialBlocks ? want : kPartialBlocks);
    hipLaunchKernelGGL(mean_var_kernel<KIND>, dim3(blocks), dim3(kSweepThreads), 0, st, a);
}
hipError_t launch_mean_var(int kind, const SweepArgs& a, hipStream_t st) {
    if (a.ncell <= 0) return hipSuccess;
    if (a.Q > 128) return hipErrorInvalidValue;  // two quadrature points per lane
    UQK_DISPATCH(kind, go_mean_var, a, st);
    return hipGetLastError();
}

hipError_t launch_fill_f64(double* p, int n, double v, hipStream_t st) {
    hipLaunchKernelGGL(fill_f64_kernel, dim3((n + 63) / 64), dim3(64), 0, st, p, n, v);
    return hipGetLastError();
}

hipError_t launch_fill_u64(unsigned long long* p, int n, hipStream_t st) {
    hipLaunchKernelGGL(fill_u64_kernel, dim3((n + 63) / 64), dim3(64), 0, st, p, n);
    return hipGetLastError();
}

}  // namespace uqk
