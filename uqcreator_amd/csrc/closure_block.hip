// closure_block.hip -- batched FP64 Newton dual closure for large systems
// (n = S * nTotal from 64 up to 128, e.g. 105 duals with 125 quadrature
// points), one 256-thread workgroup per cell.
//
// Restates Closure::SolveClosure / SolveClosureSafe (closure.cpp:133-193,
// :297-339), like closure_fp64.hip, with the work of a cell spread over four
// wavefronts instead of one:
//   * gradient (closure.cpp:234-253): threads over quadrature points evaluate
//     Lambda and U (Lambda stays in the thread's registers for DU), threads
//     over duals project; the norm is a workgroup reduction;
//   * Hessian (closure.cpp:255-271): each symmetric state block on MFMA tiles
//     (FP64, or FP32 with FP32 accumulation for the mixed path), operands
//     straight from the L2-resident tables;
//   * Cholesky (dposv 'U', vectorspace.cpp:152-171) on register tiles spread
//     cyclically over the workgroup, two pivots per barrier; the triangular
//     solves in blocks of 16 unknowns (a 16-lane sequential solve, then a
//     workgroup-wide update of the rest);
//   * every branch of the Newton iteration depends on workgroup-uniform
//     values (reduced norms, pivots read from LDS).
// The tables phi / phiw are read from global memory (L2-resident): at these
// sizes they would take the LDS the Hessian needs.
#include <hip/hip_runtime.h>

#include <type_traits>

#include "uqk_device.h"
#include "uqk_internal.h"

namespace uqk {

namespace {

constexpr int kBlockThreads = 256;
constexpr int BRA = 4;  // row-pair slots: 128 rows
constexpr int BNB = 8;  // column slots: 128 columns
constexpr int kBlockMaxN = 128;

__device__ __forceinline__ int bpidx(int r, int c) { return c * (c + 1) / 2 + r; }  // r <= c

// sum over the workgroup; every thread gets the result
__device__ __forceinline__ double block_allsum(double v, double* red) {
    v = wave_sum(v);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) red[w] = v;
    __syncthreads();
    const double s = red[0] + red[1] + red[2] + red[3];
    __syncthreads();
    return s;
}

__device__ __forceinline__ double lane_bcast(double v, int src) {
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), src);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), src);
    return __hiloint2double(hi, lo);
}

}  // namespace

typedef double f64x4 __attribute__((ext_vector_type(4)));
// 1/sqrt(h) from v_rsq_f64 (about 24 bits) refined by two Newton steps
__device__ __forceinline__ double block_rsqrt(double h) {
    double r = __builtin_amdgcn_rsq(h);
    r = r * (1.5 - 0.5 * h * r * r);
    r = r * (1.5 - 0.5 * h * r * r);
    return r;
}
typedef float f32x4 __attribute__((ext_vector_type(4)));

template <int KIND, int PREC>
__global__ void __launch_bounds__(kBlockThreads) __attribute__((amdgpu_waves_per_eu(2))) newton_block_kernel(NewtonArgs a) {
    constexpr int S = KindInfo<KIND>::S;
    constexpr int P = S * (S + 1) / 2;
    extern __shared__ double smem[];
    if (a.active && *a.active == 0) return;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int M = a.M, Q = a.Q, SM = S * M;
    const double* phi = a.phi;
    const double* phiw = a.phiw;
    // LDS: H (packed, SM(SM+1)/2) | lam | lamnew | u | g | gnew | d | Ub (S x Q) |
    //      DU (Q x P) | rdiag (SM) | reduction (8) | queue (2)
    double* H = smem;
    double* lam = H + SM * (SM + 1) / 2;
    double* lamnew = lam + SM;
    double* uu = lamnew + SM;
    double* gA = uu + SM;
    double* gB = gA + SM;
    double* dv = gB + SM;
    double* Ub = dv + SM;
    double* DUk = Ub + S * Q;
    double* rdiag = DUk + Q * P;
    double* red = rdiag + SM;  // [0..3] reduction, [4] factorisation verdict
    double* rowbuf = red + 8;  // 2 x 256: rows j, j+1 of U, double buffered
    long* qshare = reinterpret_cast<long*>(rowbuf + 512);
    const ClosureParams& c = a.c;

    // guided work queue over workgroups (as in the wave kernels, per block)
    unsigned long long w_fail = 0, w_newton = 0, w_grad = 0;
    double w_dt = 1e10;  // min ComputeDt over this workgroup's cells (fused path)
    const long qsize = 1L << a.qbits;
    long qpos = 0, qend = 0;
    for (;;) {
        if (qpos == qend) {
            long want = (qsize - qend) / (2 * a.waves);
            if (want < 1) want = 1;
            if (tid == 0) qshare[0] = (long)atomicAdd(a.counter, (unsigned long long)want);
            __syncthreads();
            const long base = qshare[0];
            __syncthreads();
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
        for (int t = tid; t < SM; t += kBlockThreads) {
            uu[t] = ug[t];
            lam[t] = lg[t];
            lamnew[t] = lg[t];
        }
        __syncthreads();

        double Lk[S];  // Lambda at this thread's point (tid < Q) of the last gradient
        double uproj = 0.0;  // <U phiW> of this thread's dual (tid < n) at the last gradient
        // gradient at X into gout (state-major e = l*nt + i); returns ||g||_2
        auto gradient = [&](const double* X, double* gout) -> double {
            // phase: gradient
            if (tid < Q) {
                const int k = tid;
                double Uv[S];
#pragma unroll
                for (int s = 0; s < S; ++s) Lk[s] = 0.0;
                for (int i = 0; i < nt; ++i) {
                    const double p = phi[k * M + i];
#pragma unroll
                    for (int s = 0; s < S; ++s) Lk[s] += X[i * S + s] * p;
                }
                closure_U<KIND>(c, Lk, Uv);
#pragma unroll
                for (int s = 0; s < S; ++s) Ub[s * Q + k] = Uv[s];
            }
            __syncthreads();
            // <U phiW>(s, i) = sum_k U_s(k) phiW(k, i) as FP64 MFMA tiles of the
            // product (S x Q)(Q x nt): one 16-column tile per wavefront, the
            // table loads four k-steps ahead; results land in gout, state-major
            {
                const int r16 = lane & 15, kq4 = lane >> 4;
                const int tnt = (nt + 15) >> 4, ks_n = (Q + 3) >> 2;
                for (int ti = wave; ti < tnt; ti += kBlockThreads / 64) {
                    const int ic = 16 * ti + r16;
                    constexpr int PF = 4;
                    auto load = [&](int ks, double& av, double& bv) {
                        const int k = 4 * ks + kq4;
                        av = (r16 < S && k < Q) ? Ub[r16 * Q + k] : 0.0;
                        bv = (ic < nt && k < Q) ? phiw[k * M + ic] : 0.0;
                    };
                    double av[PF], bv[PF];
#pragma unroll
                    for (int u = 0; u < PF; ++u) load(u, av[u], bv[u]);
                    f64x4 pacc = {0.0, 0.0, 0.0, 0.0};
                    for (int ks0 = 0; ks0 < ks_n; ks0 += PF) {
                        double an[PF], bn[PF];
#pragma unroll
                        for (int u = 0; u < PF; ++u) load(ks0 + PF + u, an[u], bn[u]);
#pragma unroll
                        for (int u = 0; u < PF; ++u) pacc = __builtin_amdgcn_mfma_f64_16x16x4f64(av[u], bv[u], pacc, 0, 0, 0);
#pragma unroll
                        for (int u = 0; u < PF; ++u) {
                            av[u] = an[u];
                            bv[u] = bn[u];
                        }
                    }
#pragma unroll
                    for (int e = 0; e < 4; ++e) {
                        const int sr = 4 * e + kq4;  // element e at (4e + l/16, l%16)
                        if (sr < S && ic < nt) gout[sr * nt + ic] = pacc[e];
                    }
                }
            }
            __syncthreads();
            double sq = 0.0;
            if (tid < n) {
                const int l = tid / nt, i = tid - l * nt;
                const double proj = gout[tid];
                uproj = proj;
                const double acc = proj - uu[i * S + l];
                gout[tid] = acc;
                sq = acc * acc;
            }
            const double nrm = sqrt(block_allsum(sq, red));
            // phase end: gradient
            return nrm;
        };

        auto hessian = [&]() {
            // phase: hessian
            if (tid < Q) closure_DU<KIND>(c, Lk, DUk + tid * P);
            __syncthreads();
            // every state block H_lm = Phi^T diag(DU_lm) PhiW (m <= l) on MFMA
            // tiles of 16 x 16, k in steps of 4, FP64 (PREC 0) or FP32 with
            // FP32 accumulation (PREC 1).  A wavefront takes a tile position
            // (tj, ti) and computes it for all S(S+1)/2 pairs at once, so the
            // two table loads per k step feed P products; positions below the
            // diagonal (tj > ti) only need the pairs l != m.  Positions are
            // dealt to the four wavefronts in snake order, full ones first.
            const int tmt = (nt + 15) >> 4, ks_n = (Q + 3) >> 2;
            const int r16 = lane & 15, kq = lane >> 4;
            const int nfull = tmt * (tmt + 1) / 2, npos = tmt * tmt;
            for (int base = 0; base < npos; base += 4) {
                const int idx = base + (((base >> 2) & 1) ? 3 - wave : wave);
                if (idx >= npos) continue;
                int tj, ti;
                bool lower;
                if (idx < nfull) {  // upper positions, row-major over tj <= ti
                    int r = idx;
                    tj = 0;
                    while (r >= tmt - tj) {
                        r -= tmt - tj;
                        ++tj;
                    }
                    ti = tj + r;
                    lower = false;
                } else {  // lower positions tj > ti
                    int r = idx - nfull;
                    tj = 1;
                    while (r >= tj) {
                        r -= tj;
                        ++tj;
                    }
                    ti = r;
                    lower = true;
                }
                const int jr = 16 * tj + r16, ic = 16 * ti + r16;
                const bool jin = jr < nt, iin = ic < nt;
                using V = typename std::conditional<PREC == 0, f64x4, f32x4>::type;
                V acc[P];
#pragma unroll
                for (int pp = 0; pp < P; ++pp) acc[pp] = V{0, 0, 0, 0};
                // table loads run PF k steps ahead of the products (L2 latency)
                constexpr int PF = 4;
                auto load = [&](int ks, double& av, double& bv) {
                    const int k = 4 * ks + kq;
                    const bool kin = k < Q;
                    av = (kin && jin) ? phi[k * M + jr] : 0.0;
                    bv = (kin && iin) ? phiw[k * M + ic] : 0.0;
                };
                double av[PF], bv[PF];
#pragma unroll
                for (int u = 0; u < PF; ++u) load(u, av[u], bv[u]);
                for (int ks0 = 0; ks0 < ks_n; ks0 += PF) {
                    double an[PF], bn[PF];
#pragma unroll
                    for (int u = 0; u < PF; ++u) load(ks0 + PF + u, an[u], bn[u]);
#pragma unroll
                    for (int u = 0; u < PF; ++u) {
                        // beyond Q the loaded values are zero, any DU row will do
                        const int k = 4 * (ks0 + u) + kq;
                        const double* du = DUk + (k < Q ? k : 0) * P;
#pragma unroll
                        for (int l = 0; l < S; ++l)
#pragma unroll
                            for (int m = 0; m <= l; ++m) {
                                const int pp = l * (l + 1) / 2 + m;
                                if (lower && l == m) continue;
                                if constexpr (PREC == 0)
                                    acc[pp] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[u], du[pp] * bv[u], acc[pp], 0, 0, 0);
                                else
                                    acc[pp] = __builtin_amdgcn_mfma_f32_16x16x4f32((float)av[u], (float)(du[pp] * bv[u]),
                                                                                    acc[pp], 0, 0, 0);
                            }
                    }
#pragma unroll
                    for (int u = 0; u < PF; ++u) {
                        av[u] = an[u];
                        bv[u] = bn[u];
                    }
                }
                if (!iin) continue;
#pragma unroll
                for (int l = 0; l < S; ++l)
#pragma unroll
                    for (int m = 0; m <= l; ++m) {
                        const int pp = l * (l + 1) / 2 + m;
                        if (lower && l == m) continue;
#pragma unroll
                        for (int e = 0; e < 4; ++e) {
                            // D layouts: FP64 (4e + l/16, l%16), FP32 (4 (l/16) + e, l%16)
                            const int j = 16 * tj + (PREC == 0 ? 4 * e + kq : 4 * kq + e);
                            if (j < nt && (l != m || j <= ic)) H[bpidx(m * nt + j, l * nt + ic)] = (double)acc[pp][e];
                        }
                    }
            }
            __syncthreads();
            // phase end: hessian
        };

        // d <- H^{-1} g (U^T U = H, upper, packed); false if not SPD
        auto cholesky_solve = [&](const double* g) -> bool {
            // phase: factor
            // H (upper) is spread over the workgroup in registers: thread
            // 16p + q holds rows 32a + 2p + i (i = 0, 1) of columns 16b + q in
            // T[a][i][b] -- row pairs cyclic over the 16 row groups, columns
            // cyclic over the 16 column groups.  A step eliminates two pivots:
            // the 16 owner threads of the two rows (one wavefront) factor the
            // 2x2 diagonal block from SGPR broadcasts, scale the rows and
            // publish them in LDS; after one barrier every thread applies the
            // rank-2 update to its tiles.
            const int p16 = tid >> 4, q16 = tid & 15;
            // with a free column the right-hand side rides along as column n:
            // its elimination is the forward substitution U^T y = g
            const bool aug = n < 16 * BNB;
            double T[BRA][2][BNB];
#pragma unroll
            for (int ta = 0; ta < BRA; ++ta)
#pragma unroll
                for (int i = 0; i < 2; ++i)
#pragma unroll
                    for (int tb = 0; tb < BNB; ++tb) {
                        const int r = 32 * ta + 2 * p16 + i, cc = 16 * tb + q16;
                        T[ta][i][tb] = (r <= cc && cc < n) ? H[bpidx(r, cc)] : (aug && cc == n && r < n) ? g[r] : 0.0;
                    }
            for (int j = 0; j < n; j += 2) {
                const bool two = j + 1 < n;
                const int pj = (j & 31) >> 1, aj = j >> 5;
                double* urow = rowbuf + 256 * ((j >> 1) & 1);  // rows j, j+1 at [c], [128 + c]
                if (p16 == pj) {
                    // the owner wavefront: lanes 16 (pj % 4) + q hold rows j, j+1
                    const int lbase = 16 * (pj & 3);
                    double h00 = 0.0, h01 = 0.0, h11 = 0.0;
#pragma unroll
                    for (int ta = 0; ta < BRA; ++ta) {
                        if (ta != aj) continue;
#pragma unroll
                        for (int tb = 0; tb < BNB; ++tb) {
                            if (tb == (j >> 4)) h00 = lane_bcast(T[ta][0][tb], lbase + (j & 15));
                            if (two && tb == ((j + 1) >> 4)) {
                                h01 = lane_bcast(T[ta][0][tb], lbase + ((j + 1) & 15));
                                h11 = lane_bcast(T[ta][1][tb], lbase + ((j + 1) & 15));
                            }
                        }
                    }
                    // U_j+1,j+1^2 = det / h00, so 1/U_j+1,j+1 = sqrt(h00) / sqrt(det)
                    // and both square roots start at once
                    const double det = fma(h00, h11, -h01 * h01);
                    bool ok = h00 > 0.0 && (!two || det > 0.0);
                    const double r0 = ok ? block_rsqrt(h00) : 0.0;
                    const double u01 = h01 * r0;
                    const double r1 = (two && ok) ? block_rsqrt(det) * (h00 * r0) : 0.0;
                    if (q16 == 0) {
                        rdiag[j] = r0;
                        if (two) rdiag[j + 1] = r1;
                        red[4] = ok ? 0.0 : 1.0;
                    }
#pragma unroll
                    for (int ta = 0; ta < BRA; ++ta) {
                        if (ta != aj) continue;
#pragma unroll
                        for (int tb = 0; tb < BNB; ++tb) {
                            const int cc = 16 * tb + q16;
                            // the same formulas give the diagonal; columns left of
                            // the pivots or past n get values nothing reads
                            const double v0 = T[ta][0][tb] * r0;
                            T[ta][0][tb] = v0;
                            urow[cc] = v0;
                            if (two) {
                                const double v1 = (T[ta][1][tb] - u01 * v0) * r1;
                                T[ta][1][tb] = v1;
                                urow[128 + cc] = v1;
                            }
                        }
                    }
                }
                __syncthreads();
                if (red[4] != 0.0) return false;  // uniform: every thread reads the owners' verdict
                const int jl = two ? j + 1 : j;
                double uc0[BNB], uc1[BNB];
#pragma unroll
                for (int tb = 0; tb < BNB; ++tb) {
                    const int cc = 16 * tb + q16;
                    const bool live = cc > jl && cc <= n;  // column n: the right-hand side
                    uc0[tb] = live ? urow[cc] : 0.0;
                    uc1[tb] = (live && two) ? urow[128 + cc] : 0.0;
                }
#pragma unroll
                for (int ta = 0; ta < BRA; ++ta) {
                    if (32 * ta + 31 <= jl) continue;  // uniform: every row of the slot is done
#pragma unroll
                    for (int i = 0; i < 2; ++i) {
                        const int r = 32 * ta + 2 * p16 + i;
                        // only rows past the pivots must keep their values;
                        // entries below the diagonal or past n are never read
                        if (r <= jl) continue;
                        const double ur0 = urow[r], ur1 = two ? urow[128 + r] : 0.0;
#pragma unroll
                        for (int tb = 0; tb < BNB; ++tb) {
                            if (16 * tb + 15 < 32 * ta) continue;  // slot wholly below the diagonal
                            T[ta][i][tb] -= ur0 * uc0[tb] + ur1 * uc1[tb];
                        }
                    }
                }
            }
            // U back to LDS (packed) for the triangular solves
#pragma unroll
            for (int ta = 0; ta < BRA; ++ta)
#pragma unroll
                for (int i = 0; i < 2; ++i)
#pragma unroll
                    for (int tb = 0; tb < BNB; ++tb) {
                        const int r = 32 * ta + 2 * p16 + i, cc = 16 * tb + q16;
                        if (r <= cc && cc < n)
                            H[bpidx(r, cc)] = T[ta][i][tb];
                        else if (aug && cc == n && r < n)
                            dv[r] = T[ta][i][tb];
                    }
            if (!aug)
                for (int e = tid; e < n; e += kBlockThreads) dv[e] = g[e];
            __syncthreads();
            // phase end: factor
            // phase: backsub
            // U^T y = g (unless eliminated with the factor) then U x = y, in
            // blocks of 16 unknowns: the 16 x 16 diagonal block is solved by
            // lanes 0..15 of wavefront 0 (SGPR broadcasts), the rest of the
            // vector is updated by all threads.
            const int nb = (n + 15) / 16;
            for (int blk = 0; blk < (aug ? 0 : nb); ++blk) {
                const int i0 = 16 * blk, i1 = i0 + 16 < n ? i0 + 16 : n;
                if (wave == 0) {
                    const int col = i0 + (lane & 15);
                    double y = (lane < 16 && col < i1) ? dv[col] : 0.0;
                    for (int i = i0; i < i1; ++i) {
                        const double yi = lane_bcast(y, i - i0) * rdiag[i];
                        if (lane < 16 && col > i && col < i1) y -= H[bpidx(i, col)] * yi;
                        if (lane == i - i0) y = yi;
                    }
                    if (lane < 16 && col < i1) dv[col] = y;
                }
                __syncthreads();
                for (int cc = i1 + tid; cc < n; cc += kBlockThreads) {
                    double acc = dv[cc];
                    for (int r = i0; r < i1; ++r) acc -= H[bpidx(r, cc)] * dv[r];
                    dv[cc] = acc;
                }
                __syncthreads();
            }
            for (int blk = nb - 1; blk >= 0; --blk) {
                const int i0 = 16 * blk, i1 = i0 + 16 < n ? i0 + 16 : n;
                if (wave == 0) {
                    // lane t owns x_i0+t; its row of the block and 1/U_tt sit in
                    // registers, so a step is one multiply, broadcast and FMA
                    const int row = i0 + (lane & 15);
                    const bool own = lane < 16 && row < i1;
                    double x = own ? dv[row] : 0.0;
                    const double rd = own ? rdiag[row] : 0.0;
                    double uh[16];
#pragma unroll
                    for (int t = 0; t < 16; ++t)
                        uh[t] = (own && i0 + t < i1 && row < i0 + t) ? H[bpidx(row, i0 + t)] : 0.0;
#pragma unroll
                    for (int t = 15; t >= 0; --t) {
                        if (i0 + t >= i1) continue;  // uniform
                        const double xt = lane_bcast(x * rd, t);
                        x = lane == t ? xt : fma(-uh[t], xt, x);
                    }
                    if (own) dv[row] = x;
                }
                __syncthreads();
                for (int r = tid; r < i0; r += kBlockThreads) {
                    double acc = dv[r];
                    for (int cc = i0; cc < i1; ++cc) acc -= H[bpidx(r, cc)] * dv[cc];
                    dv[r] = acc;
                }
                __syncthreads();
            }
            // phase end: backsub
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
            ++iters;
            if (!cholesky_solve(g)) {
                status = UQK_STATUS_NOT_SPD;
                break;
            }
            if (c.max_iter == 1) {  // one-shot IPM (closure.cpp:151-154)
                if (tid < n) {
                    const int l = tid / nt, i = tid - l * nt;
                    lam[i * S + l] -= c.alpha * dv[tid];
                }
                __syncthreads();
                status = UQK_STATUS_CONVERGED;
                break;
            }
            auto trial = [&](double step) -> double {
                if (tid < n) {
                    const int l = tid / nt, i = tid - l * nt;
                    lamnew[i * S + l] = lam[i * S + l] - step * c.alpha * dv[tid];
                }
                __syncthreads();
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
                    double* t = lam;
                    lam = lamnew;
                    lamnew = t;
                }
                break;
            }
            double* t = lam;
            lam = lamnew;
            lamnew = t;
            ng = nnew;
            if (!isfinite(nnew)) {
                status = UQK_STATUS_NONFINITE;
                break;
            }
            if (nnew < c.eps) status = UQK_STATUS_CONVERGED;
        }
        if (status < 0) status = UQK_STATUS_MAX_ITERATIONS;
        for (int t = tid; t < SM; t += kBlockThreads) lg[t] = lam[t];
        if (a.uq_out) {
            // the step's next two sweeps from the last gradient, which is at
            // the returned lambda: uQ = U(lambda phi) (momentsolver.cpp:100)
            // and u = <uQ phiW> over the cell's nTotal, zero above (:120-124)
            double* uqg = a.uq_out + cell * (long)Q * S;
            for (int t = tid; t < Q * S; t += kBlockThreads) {
                const int k = t / S, sx = t - k * S;
                uqg[t] = Ub[sx * Q + k];
            }
            double* uo = a.u_out + cell * SM;
            if (tid < n) uo[(tid - (tid / nt) * nt) * S + tid / nt] = uproj;
            for (int t = nt * S + tid; t < SM; t += kBlockThreads) uo[t] = 0.0;
            if (a.dt_min) {  // momentsolver.cpp:127-133 for this cell
                double m = 1e10;
                if (tid < Q) {
                    double uk[S];
#pragma unroll
                    for (int sx = 0; sx < S; ++sx) uk[sx] = Ub[sx * Q + tid];
                    m = point_dt<S>(a.dtp, uk, a.dx_param[cell]);
                }
                w_dt = fmin(w_dt, wave_min(m));  // per wavefront; combined below
            }
        }
        if (tid == 0) {
            if (a.status) a.status[cell] = status;
            if (a.iters) a.iters[cell] = iters;
        }
        w_fail += status != UQK_STATUS_CONVERGED ? 1ull : 0ull;
        w_newton += (unsigned long long)iters;
        w_grad += (unsigned long long)ngrad;
        __syncthreads();
    }
    if (tid == 0) {
        if (w_fail) atomicAdd(a.nfail, w_fail);
        if (w_newton) atomicAdd(a.nfail + 1, w_newton);
        if (w_grad) atomicAdd(a.nfail + 2, w_grad);
    }
    // positive doubles order as their bit patterns; one atomic per wavefront
    if (a.dt_min && lane == 0 && w_dt < 1e10) atomicMin(a.dt_min, (unsigned long long)__double_as_longlong(w_dt));
}

size_t newton_block_lds(int S, int M, int Q) {
    const int SM = S * M, P = S * (S + 1) / 2;
    return sizeof(double) * ((size_t)SM * (SM + 1) / 2 + 6 * (size_t)SM + (size_t)S * Q + (size_t)Q * P + SM + 8 + 512 + 2);
}

template <int KIND>
static hipError_t block_go(const NewtonArgs& args, int max_lds, int num_cu, hipStream_t st) {
    constexpr int S = KindInfo<KIND>::S;
    const size_t shmem = newton_block_lds(S, args.M, args.Q);
    if (shmem > (size_t)max_lds || S * args.M > kBlockMaxN || args.Q > kBlockThreads) return hipErrorInvalidValue;
    int occ = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, newton_block_kernel<KIND, 0>, kBlockThreads, shmem) !=
            hipSuccess ||
        occ <= 0)
        occ = 1;
    long blocks = args.ncell < (long)occ * num_cu ? args.ncell : (long)occ * num_cu;
    if (blocks < 1) blocks = 1;
    NewtonArgs a = args;
    // the queue counts workgroups here: one cell per workgroup at a time
    set_queue(a, blocks);
    if (a.prec == UQK_FP64)
        hipLaunchKernelGGL((newton_block_kernel<KIND, 0>), dim3((unsigned)blocks), dim3(kBlockThreads), shmem, st, a);
    else
        hipLaunchKernelGGL((newton_block_kernel<KIND, 1>), dim3((unsigned)blocks), dim3(kBlockThreads), shmem, st, a);
    return hipGetLastError();
}

bool newton_block_fits(int S, int M, int Q, int max_lds) {
    return S * M <= kBlockMaxN && Q <= kBlockThreads && newton_block_lds(S, M, Q) <= (size_t)max_lds;
}

hipError_t newton_block_launch(int kind, const NewtonArgs& a, int max_lds, int num_cu, hipStream_t st) {
    switch (kind) {
        case UQK_BOUNDED_BARRIER: return block_go<UQK_BOUNDED_BARRIER>(a, max_lds, num_cu, st);
        case UQK_EULER1D: return block_go<UQK_EULER1D>(a, max_lds, num_cu, st);
        case UQK_EULER2D: return block_go<UQK_EULER2D>(a, max_lds, num_cu, st);
        case UQK_SHALLOW_WATER2D: return block_go<UQK_SHALLOW_WATER2D>(a, max_lds, num_cu, st);
        case UQK_SHALLOW_WATER2D_BARRIER: return block_go<UQK_SHALLOW_WATER2D_BARRIER>(a, max_lds, num_cu, st);
    }
    return hipErrorInvalidValue;
}

}  // namespace uqk
