// closure_fp64.hip -- batched FP64 Newton dual closure, one wavefront per cell.
//
// Restates Closure::SolveClosure / SolveClosureSafe (closure.cpp:133-193,
// :297-339) for a whole batch of cells:
//   * persistent workgroups pull cells from a device work queue, so cells
//     with many Newton iterations do not stall the others;
//   * each wave64 owns one cell: lambda, u, gradient, the trial gradient and
//     the packed upper triangle of the Hessian live in the wave's LDS slice;
//   * the basis tables phi / phiw are staged once per workgroup in LDS;
//   * Hessian (closure.cpp:255-271): every lane owns pairs (j <= i) of basis
//     indices and accumulates all S(S+1)/2 state blocks over the quadrature
//     points -- each block phi^T diag(DU_lm w) phi is symmetric in (j, i);
//   * Cholesky (dposv 'U', vectorspace.cpp:152-171): right-looking on the
//     packed upper triangle, lane c owns column c of the trailing update;
//   * gradient norms (closure.cpp:195-203) are butterfly wave reductions, so
//     every lane takes the same branch.
#include <hip/hip_runtime.h>

#include "uqk_device.h"
#include "uqk_internal.h"

namespace uqk {

__device__ __forceinline__ int pidx(int r, int c) { return c * (c + 1) / 2 + r; }  // r <= c

template <int KIND, bool TLDS>
__global__ void __launch_bounds__(256) newton_fp64_kernel(NewtonArgs a) {
    constexpr int S = KindInfo<KIND>::S;
    constexpr int P = S * (S + 1) / 2;
    extern __shared__ double smem[];
    if (a.active && *a.active == 0) return;
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int M = a.M, Q = a.Q;

    const double* phi;
    const double* phiw;
    double* wbase;
    if constexpr (TLDS) {
        double* sphi = smem;
        double* sphiw = smem + Q * M;
        for (int t = threadIdx.x; t < Q * M; t += blockDim.x) {
            sphi[t] = a.phi[t];
            sphiw[t] = a.phiw[t];
        }
        __syncthreads();
        phi = sphi;
        phiw = sphiw;
        wbase = smem + 2 * Q * M + wave * a.wave_doubles;
    } else {
        phi = a.phi;
        phiw = a.phiw;
        wbase = smem + wave * a.wave_doubles;
    }
    const int SM = S * M;
    const int nmax = S * M;
    double* lam = wbase;
    double* lamnew = lam + SM;
    double* uu = lamnew + SM;
    double* gA = uu + SM;
    double* gB = gA + nmax;
    double* d = gB + nmax;
    double* Lk = d + nmax;
    double* Uk = Lk + Q * S;
    double* DUk = Uk + Q * S;
    double* H = DUk + Q * P;

    const ClosureParams& c = a.c;

    // Work queue, guided: a wave takes (remaining / 2 waves) queue positions
    // with one atomicAdd, so chunks start large (few atomics) and shrink to
    // single positions at the end (balance).  Position t of the 2^qbits queue
    // is cell bitreverse(t) (skipped when >= ncell): neighbouring costly cells
    // go to different chunks.  Newton and failure counts are summed per wave
    // (wave-uniform, scalar) and added once at the end.
    unsigned long long w_fail = 0, w_newton = 0, w_grad = 0;
    const long qsize = 1L << a.qbits;
    long qpos = 0, qend = 0;
    for (;;) {
        if (qpos == qend) {
            long want = (qsize - qend) / (2 * a.waves);
            if (want < 1) want = 1;
            unsigned long long got = 0;
            if (lane == 0) got = atomicAdd(a.counter, (unsigned long long)want);
            const long base = (long)(((unsigned long long)__builtin_amdgcn_readfirstlane((unsigned)(got >> 32)) << 32) |
                                     (unsigned)__builtin_amdgcn_readfirstlane((unsigned)got));
            if (base >= qsize) break;
            qpos = base;
            qend = base + want < qsize ? base + want : qsize;
        }
        const unsigned pos = (unsigned)qpos;
        ++qpos;
        const long cell = a.qbits ? (long)(__builtin_bitreverse32(pos) >> (32 - a.qbits)) : 0;
        if (cell >= a.ncell) continue;
        const int nt = a.nt_cell ? a.nt_cell[cell] : M;
        const int n = S * nt;
        const double* ug = a.u + cell * SM;
        double* lg = a.lam + cell * SM;
        for (int t = lane; t < SM; t += 64) {
            uu[t] = ug[t];
            lam[t] = lg[t];
            lamnew[t] = lg[t];
        }
        wave_sync();

        // gradient of the dual objective at X, into gout; returns ||gout||_2
        // (closure.cpp:234-253).  Leaves Lambda(X) at the points in Lk.
        auto gradient = [&](const double* X, double* gout) -> double {
            for (int k = lane; k < Q; k += 64) {
                double L[S], Uv[S];
#pragma unroll
                for (int s = 0; s < S; ++s) L[s] = 0.0;
                for (int i = 0; i < nt; ++i) {
                    double p = phi[k * M + i];
#pragma unroll
                    for (int s = 0; s < S; ++s) L[s] += X[i * S + s] * p;
                }
                closure_U<KIND>(c, L, Uv);
#pragma unroll
                for (int s = 0; s < S; ++s) {
                    Lk[k * S + s] = L[s];
                    Uk[k * S + s] = Uv[s];
                }
            }
            wave_sync();
            double sq = 0.0;
            for (int e = lane; e < n; e += 64) {
                int l = e / nt, i = e - l * nt;
                double acc = 0.0;
                for (int k = 0; k < Q; ++k) acc += Uk[k * S + l] * phiw[k * M + i];
                acc -= uu[i * S + l];
                gout[e] = acc;
                sq += acc * acc;
            }
            double nrm = sqrt(wave_sum(sq));
            wave_sync();
            return nrm;
        };

        // Hessian at the iterate whose Lambda(x_k) is in Lk (closure.cpp:255-271).
        auto hessian = [&]() {
            for (int k = lane; k < Q; k += 64) closure_DU<KIND>(c, Lk + k * S, DUk + k * P);
            wave_sync();
            const int npair = nt * (nt + 1) / 2;
            for (int e = lane; e < npair; e += 64) {
                int i = (int)((sqrt(8.0 * e + 1.0) - 1.0) * 0.5);
                while ((i + 1) * (i + 2) / 2 <= e) ++i;
                while (i * (i + 1) / 2 > e) --i;
                int j = e - i * (i + 1) / 2;  // j <= i
                double acc[P];
#pragma unroll
                for (int p = 0; p < P; ++p) acc[p] = 0.0;
                for (int k = 0; k < Q; ++k) {
                    double pk = phi[k * M + j] * phiw[k * M + i];
#pragma unroll
                    for (int p = 0; p < P; ++p) acc[p] += pk * DUk[k * P + p];
                }
#pragma unroll
                for (int l = 0; l < S; ++l)
#pragma unroll
                    for (int m = 0; m <= l; ++m) {
                        double v = acc[l * (l + 1) / 2 + m];
                        if (m < l) {
                            H[pidx(m * nt + j, l * nt + i)] = v;
                            H[pidx(m * nt + i, l * nt + j)] = v;
                        } else {
                            H[pidx(m * nt + j, m * nt + i)] = v;
                        }
                    }
            }
            wave_sync();
        };

        // d <- H^{-1} d via U^T U = H (upper, packed).  false if not SPD.
        auto cholesky_solve = [&]() -> bool {
            for (int j = 0; j < n; ++j) {
                double ajj = H[pidx(j, j)];
                if (!(ajj > 0.0)) return false;
                double ujj = sqrt(ajj);
                double inv = 1.0 / ujj;
                for (int col = j + 1 + lane; col < n; col += 64) H[pidx(j, col)] *= inv;
                wave_sync();
                if (lane == 0) H[pidx(j, j)] = ujj;
                for (int col = j + 1 + lane; col < n; col += 64) {
                    double ujc = H[pidx(j, col)];
                    int base = col * (col + 1) / 2;
                    for (int r = j + 1; r <= col; ++r) H[base + r] -= H[pidx(j, r)] * ujc;
                }
                wave_sync();
            }
            for (int i = 0; i < n; ++i) {
                double yi = d[i] / H[pidx(i, i)];
                for (int col = i + 1 + lane; col < n; col += 64) d[col] -= H[pidx(i, col)] * yi;
                if (lane == 0) d[i] = yi;
                wave_sync();
            }
            for (int i = n - 1; i >= 0; --i) {
                double xi = d[i] / H[pidx(i, i)];
                int base = i * (i + 1) / 2;
                for (int r = lane; r < i; r += 64) d[r] -= H[base + r] * xi;
                if (lane == 0) d[i] = xi;
                wave_sync();
            }
            return true;
        };

        double* g = gA;
        double* gnew = gB;
        int ngrad = 1;
        double ng = gradient(lam, g);
        int status = -1;
        int iters = 0;
        if (!c.safe && ng < c.eps) status = UQK_STATUS_CONVERGED;
        for (int it = 0; status < 0 && it < c.max_iter; ++it) {
            if (it > 0) {
                double* t = g;
                g = gnew;
                gnew = t;
            }
            hessian();
            for (int e = lane; e < n; e += 64) d[e] = g[e];
            wave_sync();
            ++iters;
            if (!cholesky_solve()) {
                status = UQK_STATUS_NOT_SPD;
                break;
            }
            if (c.max_iter == 1) {  // one-shot IPM (closure.cpp:151-154)
                for (int e = lane; e < n; e += 64) {
                    int l = e / nt, i = e - l * nt;
                    lam[i * S + l] -= c.alpha * d[e];
                }
                wave_sync();
                status = UQK_STATUS_CONVERGED;
                break;
            }
            auto trial = [&](double step) -> double {
                for (int e = lane; e < n; e += 64) {
                    int l = e / nt, i = e - l * nt;
                    lamnew[i * S + l] = lam[i * S + l] - step * c.alpha * d[e];
                }
                wave_sync();
                ++ngrad;
                return gradient(lamnew, gnew);
            };
            double step = 1.0;
            double nnew = trial(step);
            int halvings = 0;
            bool done = false;
            while (ng < nnew || (c.safe && !isfinite(nnew))) {
                step *= 0.5;
                nnew = trial(step);
                if (nnew < c.eps) {
                    status = UQK_STATUS_CONVERGED;
                    done = true;
                    break;
                }
                if (++halvings > c.max_halvings) {
                    status = UQK_STATUS_LINE_SEARCH;
                    done = true;
                    break;
                }
            }
            if (done) {
                if (status == UQK_STATUS_CONVERGED) {
                    for (int t = lane; t < SM; t += 64) lam[t] = lamnew[t];
                    wave_sync();
                }
                break;
            }
            for (int t = lane; t < SM; t += 64) lam[t] = lamnew[t];
            wave_sync();
            ng = nnew;
            if (!isfinite(nnew)) {
                status = UQK_STATUS_NONFINITE;
                break;
            }
            if (nnew < c.eps) status = UQK_STATUS_CONVERGED;
        }
        if (status < 0) status = UQK_STATUS_MAX_ITERATIONS;
        for (int t = lane; t < SM; t += 64) lg[t] = lam[t];
        if (lane == 0) {
            if (a.status) a.status[cell] = status;
            if (a.iters) a.iters[cell] = iters;
        }
        // status, iters and ngrad are wave-uniform: these sums stay scalar
        w_fail += status != UQK_STATUS_CONVERGED ? 1ull : 0ull;
        w_newton += (unsigned long long)iters;
        w_grad += (unsigned long long)ngrad;
        wave_sync();
    }
    if (lane == 0) {
        if (w_fail) atomicAdd(a.nfail, w_fail);
        if (w_newton) atomicAdd(a.nfail + 1, w_newton);
        if (w_grad) atomicAdd(a.nfail + 2, w_grad);
    }
}

// LDS doubles needed by one wave for the given sizes.
int newton_fp64_wave_doubles(int S, int M, int Q) {
    const int P = S * (S + 1) / 2;
    const int n = S * M;
    return 3 * S * M + 3 * n + 2 * Q * S + Q * P + n * (n + 1) / 2;
}

template <int KIND>
static hipError_t launch_kind(const NewtonArgs& a, bool tlds, int blocks, int threads, size_t shmem,
                              hipStream_t st) {
    if (tlds) {
        hipLaunchKernelGGL((newton_fp64_kernel<KIND, true>), dim3(blocks), dim3(threads), shmem, st, a);
    } else {
        hipLaunchKernelGGL((newton_fp64_kernel<KIND, false>), dim3(blocks), dim3(threads), shmem, st, a);
    }
    return hipGetLastError();
}

template <int KIND>
static int occupancy_kind(bool tlds, int threads, size_t shmem) {
    int nb = 0;
    if (tlds)
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, newton_fp64_kernel<KIND, true>, threads, shmem);
    else
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, newton_fp64_kernel<KIND, false>, threads, shmem);
    return nb;
}

int newton_fp64_occupancy(int kind, bool tlds, int threads, size_t shmem) {
    switch (kind) {
        case UQK_BOUNDED_BARRIER: return occupancy_kind<UQK_BOUNDED_BARRIER>(tlds, threads, shmem);
        case UQK_EULER1D: return occupancy_kind<UQK_EULER1D>(tlds, threads, shmem);
        case UQK_EULER2D: return occupancy_kind<UQK_EULER2D>(tlds, threads, shmem);
        case UQK_SHALLOW_WATER2D: return occupancy_kind<UQK_SHALLOW_WATER2D>(tlds, threads, shmem);
        case UQK_SHALLOW_WATER2D_BARRIER: return occupancy_kind<UQK_SHALLOW_WATER2D_BARRIER>(tlds, threads, shmem);
    }
    return 0;
}

hipError_t newton_fp64_launch(int kind, const NewtonArgs& args, bool tlds, int blocks, int threads, size_t shmem,
                              hipStream_t st) {
    NewtonArgs a = args;
    set_queue(a, (long)blocks * (threads / 64));
    switch (kind) {
        case UQK_BOUNDED_BARRIER: return launch_kind<UQK_BOUNDED_BARRIER>(a, tlds, blocks, threads, shmem, st);
        case UQK_EULER1D: return launch_kind<UQK_EULER1D>(a, tlds, blocks, threads, shmem, st);
        case UQK_EULER2D: return launch_kind<UQK_EULER2D>(a, tlds, blocks, threads, shmem, st);
        case UQK_SHALLOW_WATER2D: return launch_kind<UQK_SHALLOW_WATER2D>(a, tlds, blocks, threads, shmem, st);
        case UQK_SHALLOW_WATER2D_BARRIER:
            return launch_kind<UQK_SHALLOW_WATER2D_BARRIER>(a, tlds, blocks, threads, shmem, st);
    }
    return hipErrorInvalidValue;
}

}  // namespace uqk
