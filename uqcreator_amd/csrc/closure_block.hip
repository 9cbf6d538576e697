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
//   * Cholesky (dposv 'U', vectorspace.cpp:152-171) on 16 x 16 register tiles
//     dealt over the four wavefronts, four pivots per step with the panel and
//     the trailing update as FP64 MFMAs, one barrier per step; the triangular
//     solves in blocks of 16 unknowns (a 16-lane sequential solve, then a
//     workgroup-wide update of the rest);
//   * every branch of the Newton iteration depends on workgroup-uniform
//     values (reduced norms, pivots read from LDS);
//   * the time loop's mixed closure (PREC 2) keeps W = U^-1 of each cell's
//     last factor (inverted in place by 16 x 16 blocks, the blocks above the
//     diagonal on FP64 MFMA), rounded to FP32, in a per-cell slot: while it
//     stays good, the next steps' Newton steps take d = P g + P (g - H P g),
//     P = W W^T, against the FP64 Hessian-vector product -- four triangular
//     matrix-vector products, with no Hessian, factorisation or triangular
//     solve.
// The tables phi / phiw are read from global memory (L2-resident): at these
// sizes they would take the LDS the Hessian needs.
#include <hip/hip_runtime.h>

#include <type_traits>

#include "uqk_device.h"
#include "uqk_internal.h"

namespace uqk {

namespace {

constexpr int kBlockThreads = 256;
constexpr int kBlockMaxN = 128;
// register tile slots per wavefront: 8 x 8 tiles of 16 x 16 cover n <= 128,
// 36 of them on or above the diagonal, dealt to 4 wavefronts
constexpr int kSlots = 9;
// pinv_ok of a kept factor (workgroup kernel): the cell's nTotal, plus this
// bit once the slot holds the factor's inverse
constexpr int kKeptInverted = 0x100;

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
    //      DU (Q x P) | rdiag (SM) | reduction (8) | diagonal tile (256) |
    //      rows of U (2 x 512) | queue (2)
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
    double* red = rdiag + SM;  // [0..3] reduction
    double* dbuf = red + 8;      // 256: the current diagonal tile of the factorisation
    double* rowbuf = dbuf + 256;  // 2 x 512: four rows of U, 64 per tile column, two interleaved buffers
    long* qshare = reinterpret_cast<long*>(rowbuf + 1024);
    const ClosureParams& c = a.c;

    // guided work queue over workgroups (as in the wave kernels, per block;
    // after a residual check over its list of cells, bit-reversed)
    unsigned long long w_fail = 0, w_newton = 0, w_grad = 0, w_factor = 0, w_reuse = 0;
    double w_dt = 1e10;  // min ComputeDt over this workgroup's cells (fused path)
    // kept factors (PREC 2): floats per cell slot, the packed U of the
    // largest order
    const long fslot = (long)SM * (SM + 1) / 2;
    const long nlist = a.cell_list ? (long)*a.cell_count : 0;
    int lbits = 0;
    while ((1L << lbits) < nlist) ++lbits;
    const int qb = a.cell_list ? lbits : a.qbits;
    const long qsize = 1L << qb;
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
        const long qcell = qb ? (long)(__builtin_bitreverse32(pos) >> (32 - qb)) : 0;
        long cell = qcell;
        if (a.cell_list) {
            if (qcell >= nlist) continue;
            cell = a.cell_list[qcell];
        }
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
        // the kept factor (PREC 2) is usable when pinv_ok[cell] is this
        // cell's nTotal (the order it was factored at), 0 when there is none
        bool reuse = false, refactor = true, factored = false, flds = false;
        int kept_tag = 0;
        if constexpr (PREC == 2) {
            if (a.pinv_mode == 1 && c.max_iter != 1 && cell < a.pinv_slots) {
                reuse = true;
                kept_tag = a.pinv_ok[cell];
                refactor = (kept_tag & 0xff) != nt;
            }
        }
        __syncthreads();

        double Lk[S];  // Lambda at this thread's point (tid < Q) of the last gradient
        double uproj = 0.0;  // <U phiW> of this thread's dual (tid < n) at the last gradient
        // <Ub phiW>: out(s, i) = sum_k Ub(s, k) phiW(k, i), state-major
        // (e = s nt + i), as FP64 MFMA tiles of the product (S x Q)(Q x nt):
        // one 16-column tile per wavefront, the table loads four k-steps ahead
        auto project = [&](double* out) {
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
                        if (sr < S && ic < nt) out[sr * nt + ic] = pacc[e];
                    }
                }
            }
        };
        // Lambda at this thread's point (tid < Q) for the duals X, into Lk
        auto point_lambda = [&](const double* X) {
            const int k = tid;
#pragma unroll
            for (int s = 0; s < S; ++s) Lk[s] = 0.0;
            // (the table reads come from L2: eight in flight at a time)
#pragma unroll 8
            for (int i = 0; i < nt; ++i) {
                const double p = phi[k * M + i];
#pragma unroll
                for (int s = 0; s < S; ++s) Lk[s] += X[i * S + s] * p;
            }
        };
        // gradient at X into gout (state-major e = l*nt + i); returns ||g||_2
        auto gradient = [&](const double* X, double* gout) -> double {
            // phase: gradient
            if (tid < Q) {
                const int k = tid;
                double Uv[S];
                point_lambda(X);
                closure_U<KIND>(c, Lk, Uv);
#pragma unroll
                for (int s = 0; s < S; ++s) Ub[s * Q + k] = Uv[s];
            }
            __syncthreads();
            project(gout);
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
        // out = H v for the FP64 Hessian at the iterate of the last gradient
        // (Lambda in Lk of the threads over points), without forming H: per
        // point, w(k) = DU_k (v^T phi_k), then out = <w phiW> -- the Hessian
        // of closure.cpp:255-271 applied to v (v and out state-major).
        // Overwrites Ub.
        auto hessvec = [&](const double* v, double* out) {
            if (tid < Q) {
                const int k = tid;
                double Ld[S], y[P];
#pragma unroll
                for (int s = 0; s < S; ++s) Ld[s] = 0.0;
#pragma unroll 8
                for (int i = 0; i < nt; ++i) {
                    const double p = phi[k * M + i];
#pragma unroll
                    for (int s = 0; s < S; ++s) Ld[s] += v[s * nt + i] * p;
                }
                closure_DU<KIND>(c, Lk, y);
#pragma unroll
                for (int l = 0; l < S; ++l) {
                    double w = 0.0;
#pragma unroll
                    for (int m = 0; m < S; ++m) w += y[l >= m ? l * (l + 1) / 2 + m : m * (m + 1) / 2 + l] * Ld[m];
                    Ub[l * Q + k] = w;
                }
            }
            __syncthreads();
            project(out);
            __syncthreads();
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

        // U^T y = dv, then U x = y, in place (U packed upper in H, 1/U_rr in
        // rdiag), in blocks of 16 unknowns: the 16 x 16 diagonal block is
        // solved by lanes 0..15 of wavefront 0 (SGPR broadcasts), the rest of
        // the vector is updated by all threads.  dv is ready in LDS (after a
        // barrier) on entry and on return.
        auto forward_solve = [&]() {
            const int nb = (n + 15) / 16;
            for (int blk = 0; blk < nb; ++blk) {
                const int i0 = 16 * blk, i1 = i0 + 16 < n ? i0 + 16 : n;
                if (wave == 0) {
                    // lane t owns y_i0+t; its column of the block and the
                    // pivots' 1/U_ii sit in registers, read ahead of the chain
                    const int col = i0 + (lane & 15);
                    const bool own = lane < 16 && col < i1;
                    double y = own ? dv[col] : 0.0;
                    double uh[16], rd[16];
#pragma unroll
                    for (int t = 0; t < 16; ++t) {
                        uh[t] = (own && i0 + t < col) ? H[bpidx(i0 + t, col)] : 0.0;
                        rd[t] = i0 + t < i1 ? rdiag[i0 + t] : 0.0;
                    }
#pragma unroll
                    for (int t = 0; t < 16; ++t) {
                        if (i0 + t >= i1) continue;  // uniform
                        const double yi = lane_bcast(y, t) * rd[t];
                        if (own && col > i0 + t) y -= uh[t] * yi;
                        if (lane == t) y = yi;
                    }
                    if (own) dv[col] = y;
                }
                __syncthreads();
                for (int cc = i1 + tid; cc < n; cc += kBlockThreads) {
                    double acc = dv[cc];
                    for (int r = i0; r < i1; ++r) acc -= H[bpidx(r, cc)] * dv[r];
                    dv[cc] = acc;
                }
                __syncthreads();
            }
        };
        auto back_solve = [&]() {
            const int nb = (n + 15) / 16;
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
        };
        // The same two solves with the kept inverse W = U^-1 (packed upper
        // in H): y = W^T v, then x = W y, each a matrix-vector product,
        // thread c (r) summing column c (row r) of W in four partial sums;
        // dv in and out.
        auto solve_kept = [&]() {
            double y = 0.0;
            if (tid < n) {
                const int c = tid;
                double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
                int r = 0;
                for (; r + 3 <= c; r += 4) {
                    s0 = fma(H[bpidx(r, c)], dv[r], s0);
                    s1 = fma(H[bpidx(r + 1, c)], dv[r + 1], s1);
                    s2 = fma(H[bpidx(r + 2, c)], dv[r + 2], s2);
                    s3 = fma(H[bpidx(r + 3, c)], dv[r + 3], s3);
                }
                for (; r <= c; ++r) s0 = fma(H[bpidx(r, c)], dv[r], s0);
                y = (s0 + s1) + (s2 + s3);
            }
            __syncthreads();
            if (tid < n) dv[tid] = y;
            __syncthreads();
            double x = 0.0;
            if (tid < n) {
                const int r = tid;
                double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
                int c = r;
                for (; c + 3 < n; c += 4) {
                    s0 = fma(H[bpidx(r, c)], dv[c], s0);
                    s1 = fma(H[bpidx(r, c + 1)], dv[c + 1], s1);
                    s2 = fma(H[bpidx(r, c + 2)], dv[c + 2], s2);
                    s3 = fma(H[bpidx(r, c + 3)], dv[c + 3], s3);
                }
                for (; c < n; ++c) s0 = fma(H[bpidx(r, c)], dv[c], s0);
                x = (s0 + s1) + (s2 + s3);
            }
            __syncthreads();
            if (tid < n) dv[tid] = x;
            __syncthreads();
        };

        // W = U^-1 in place over the factor in H (1/U_rr in rdiag).
        auto invert_factor = [&]() {
            // W_II = U_II^-1 in place, by the 16 threads 16 I + i of
            // block I (one wavefront), a column at a time (W U = I):
            // W_jj = 1/U_jj, W_ij = -(sum_{i <= k < j} W_ik U_kj) / U_jj
            const int b0 = 16 * (tid >> 4), il = tid & 15, ii = b0 + il;
            for (int j = 0; j < 16; ++j) {
                const int jc = b0 + j;
                const bool col = jc < n;
                double t = 0.0;
                if (col && il < j)
                    for (int k = il; k < j; ++k) t = fma(H[bpidx(ii, b0 + k)], H[bpidx(b0 + k, jc)], t);
                wave_sync();
                if (col && il < j)
                    H[bpidx(ii, jc)] = -t * rdiag[jc];
                else if (col && il == j)
                    H[bpidx(jc, jc)] = rdiag[jc];
                wave_sync();
            }
            __syncthreads();
            // the blocks above the diagonal, a block column J at a
            // time (W U = I): W_IJ = -T_I W_JJ, T_I = sum_{I<=K<J}
            // W_IK U_KJ, as 16 x 16 x 4 FP64 MFMA products (wavefront
            // w takes the rows I = w, w + 4).  W_IK (K < J) are final
            // and U_KJ still in place while every T_I is summed; after
            // a barrier T_I goes through the wavefront's part of rowbuf
            // into the A layout and W_IJ replaces U_IJ.
            {
                const int nbk = (n + 15) / 16, r16 = lane & 15, kq4 = lane >> 4;
                double* scr = rowbuf + 256 * wave;
                for (int J = 1; J < nbk; ++J) {
                    const int gcol = 16 * J + r16;
                    f64x4 tp[2];
#pragma unroll
                    for (int p = 0; p < 2; ++p) {
                        const int I = wave + 4 * p;
                        f64x4 t = {0.0, 0.0, 0.0, 0.0};
                        if (I < J) {
                            const int gi = 16 * I + r16;
                            for (int K = I; K < J; ++K)
#pragma unroll
                                for (int ks = 0; ks < 4; ++ks) {
                                    const int kk = 16 * K + 4 * ks + kq4;
                                    const double av = gi <= kk ? H[bpidx(gi, kk)] : 0.0;
                                    const double bv = gcol < n ? H[bpidx(kk, gcol)] : 0.0;
                                    t = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, t, 0, 0, 0);
                                }
                        }
                        tp[p] = t;
                    }
                    __syncthreads();
#pragma unroll
                    for (int p = 0; p < 2; ++p) {
                        const int I = wave + 4 * p;
                        if (I >= J) continue;
#pragma unroll
                        for (int e = 0; e < 4; ++e) scr[(4 * e + kq4) * 16 + r16] = tp[p][e];
                        wave_sync();
                        f64x4 w = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
                        for (int ks = 0; ks < 4; ++ks) {
                            const int k = 4 * ks + kq4;
                            const double av = scr[r16 * 16 + k];
                            const double bv = (k <= r16 && gcol < n) ? H[bpidx(16 * J + k, gcol)] : 0.0;
                            w = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, w, 0, 0, 0);
                        }
                        if (gcol < n) {
#pragma unroll
                            for (int e = 0; e < 4; ++e) H[bpidx(16 * I + 4 * e + kq4, gcol)] = -w[e];
                        }
                        wave_sync();
                    }
                    __syncthreads();
                }
            }
        };

        // d <- H^{-1} g (U^T U = H, upper, packed); false if not SPD
        auto cholesky_solve = [&](const double* g) -> bool {
            // phase: factor
            // [H | g] (upper; with a free column the right-hand side rides
            // along as column n, so its elimination is the forward
            // substitution U^T y = g) in 16 x 16 tiles of the FP64 MFMA
            // accumulator layout: element e of lane l of tile (I, J), I <= J,
            // is row 16 I + 4 e + l/16, column 16 J + l%16.  Tile t =
            // J(J+1)/2 + I lives in wavefront t % 4, register slot t / 4.  A
            // step eliminates the four pivots j0..j0+3 (element e0 of tile row
            // I0 = j0/16):
            //   1. every wavefront factors the 4x4 diagonal block from SGPR
            //      broadcasts of its copy of tile (I0, I0), H_dd = U_dd^T U_dd,
            //      and forms W = U_dd^-T;
            //   2. the holders of tiles (I0, J) form the four rows of U,
            //      P_J = W H(j0..j0+3, 16J..16J+15), as one MFMA 16x16x4 each
            //      (A = W, B = element e0 of the tile as it stands) and
            //      publish them;
            //   3. every trailing tile takes H(I1, I2) -= P_I1^T P_I2, one MFMA
            //      each (A = -P_I1, B = P_I2, both read in their lane layout).
            // One barrier per four pivots, one more per tile row.
            const bool aug = n < kBlockMaxN;
            const int NT = (n + (aug ? 1 : 0) + 15) >> 4;
            const int ntile = NT * (NT + 1) / 2;
            const int wv = __builtin_amdgcn_readfirstlane(wave);
            const int r16 = lane & 15, g4 = lane >> 4;
            int sI[kSlots], sJ[kSlots];
            bool sv[kSlots];
#pragma unroll
            for (int sl = 0; sl < kSlots; ++sl) {
                const int t = 4 * sl + wv;
                int J = 0;
                while ((J + 1) * (J + 2) / 2 <= t) ++J;
                sJ[sl] = J;
                sI[sl] = t - J * (J + 1) / 2;
                sv[sl] = t < ntile;
            }
            f64x4 T[kSlots];
#pragma unroll
            for (int sl = 0; sl < kSlots; ++sl)
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const int r = 16 * sI[sl] + 4 * e + g4, cc = 16 * sJ[sl] + r16;
                    T[sl][e] = !sv[sl] ? 0.0
                               : (r <= cc && cc < n) ? H[bpidx(r, cc)]
                               : (aug && cc == n && r < n) ? g[r] : 0.0;
                }
            // A operand slot of this lane in W (row l%16 < 4, column l/16 <= row)
            const int wslot = (r16 < 4 && g4 <= r16) ? r16 * (r16 + 1) / 2 + g4 : -1;
            // the not-SPD verdict is the same in every wavefront (same
            // values); a failed factorisation runs to the end on NaNs and
            // reports false there
            bool spd = true;
            for (int I0 = 0; 16 * I0 < n; ++I0) {
                // every wavefront copies the diagonal tile (I0, I0) and keeps
                // its copy current with the MFMA the holder applies, so each
                // factors the 4x4 diagonal blocks itself (same values, same
                // verdict) and no barrier separates them from the panel
                const int td = I0 * (I0 + 1) / 2 + I0;
                if (wv == (td & 3)) {
#pragma unroll
                    for (int sl = 0; sl < kSlots; ++sl)
                        if (4 * sl + wv == td)
#pragma unroll
                            for (int e = 0; e < 4; ++e) dbuf[64 * e + lane] = T[sl][e];
                }
                __syncthreads();
                f64x4 Dg;
#pragma unroll
                for (int e = 0; e < 4; ++e) Dg[e] = dbuf[64 * e + lane];
                // the four steps of the tile row, unrolled: e0 and the broadcast
                // lanes are constants; the panel buffer alternates per step
#pragma unroll
                for (int e0 = 0; e0 < 4; ++e0) {
                    const int j0 = 16 * I0 + 4 * e0;
                    if (j0 >= n) break;
                    const int np = n - j0 < 4 ? n - j0 : 4;  // real pivots; the rest become identity rows
                    // P_J of lane l at pb[2 (64 J + l)]: the two buffers interleave, so
                    // both use the same addresses (the parity is an immediate offset)
                    double* pb = rowbuf + (e0 & 1);
                    double h[4][4];
#pragma unroll
                    for (int x = 0; x < 4; ++x)
#pragma unroll
                        for (int y = x; y < 4; ++y)
                            h[x][y] = y < np ? lane_bcast(Dg[e0], 16 * x + 4 * e0 + y) : (x == y ? 1.0 : 0.0);
                    // two pivots at a time: 1/U_11 = sqrt(h00) / sqrt(det)
                    const double det01 = fma(h[0][0], h[1][1], -h[0][1] * h[0][1]);
                    const double r0 = block_rsqrt(h[0][0]);
                    const double rd01 = block_rsqrt(det01);
                    const double u01 = h[0][1] * r0, u02 = h[0][2] * r0, u03 = h[0][3] * r0;
                    const double r1 = rd01 * (h[0][0] * r0);
                    const double u12 = (h[1][2] - u01 * u02) * r1, u13 = (h[1][3] - u01 * u03) * r1;
                    const double g22 = h[2][2] - u02 * u02 - u12 * u12;
                    const double g23 = h[2][3] - u02 * u03 - u12 * u13;
                    const double g33 = h[3][3] - u03 * u03 - u13 * u13;
                    const double det23 = fma(g22, g33, -g23 * g23);
                    spd = spd && h[0][0] > 0.0 && det01 > 0.0 && g22 > 0.0 && det23 > 0.0;
                    const double r2 = block_rsqrt(g22);
                    const double rd23 = block_rsqrt(det23);
                    const double u23 = g23 * r2;
                    const double r3 = rd23 * (g22 * r2);
                    // W = U_dd^-T, lower triangular, packed by rows
                    double w[10];
                    w[0] = r0;
                    w[2] = r1;
                    w[1] = -r1 * u01 * r0;
                    w[5] = r2;
                    w[4] = -r2 * u12 * r1;
                    w[3] = -r2 * (u02 * r0 + u12 * w[1]);
                    w[9] = r3;
                    w[8] = -r3 * u23 * r2;
                    w[7] = -r3 * (u13 * r1 + u23 * w[4]);
                    w[6] = -r3 * (u03 * r0 + u13 * w[1] + u23 * w[3]);
                    if (tid == 0) {
                        rdiag[j0] = r0;
                        if (np > 1) rdiag[j0 + 1] = r1;
                        if (np > 2) rdiag[j0 + 2] = r2;
                        if (np > 3) rdiag[j0 + 3] = r3;
                    }
                    double wop = 0.0;
#pragma unroll
                    for (int t = 0; t < 10; ++t) wop = wslot == t ? w[t] : wop;
#pragma unroll
                    for (int sl = 0; sl < kSlots; ++sl) {
                        if (!sv[sl] || sI[sl] != I0) continue;
                        const f64x4 z = {0.0, 0.0, 0.0, 0.0};
                        const f64x4 pr = __builtin_amdgcn_mfma_f64_16x16x4f64(wop, T[sl][e0], z, 0, 0, 0);
                        // zero left of the diagonal
                        pb[2 * (64 * sJ[sl] + lane)] = 16 * sJ[sl] + r16 < j0 + g4 ? 0.0 : pr[0];
                    }
                    __syncthreads();
                    if (j0 + 4 < n) {
                        // trailing tiles, all slots in one basic block so the
                        // scheduler overlaps the operand loads with the MFMAs:
                        // a slot without rows past the pivots gets zero
                        // operands (both, so stale LDS cannot bring a NaN),
                        // which adds an exact zero; sI, sJ < 8 keep the reads
                        // inside the buffer
#pragma unroll
                        for (int sl = 0; sl < kSlots; ++sl) {
                            const int I1 = sI[sl], I2 = sJ[sl];
                            const bool act = sv[sl] && I1 >= I0 && 16 * I1 + 15 >= j0 + 4;
                            const double ra = pb[2 * (64 * I1 + lane)], rb = pb[2 * (64 * I2 + lane)];
                            T[sl] = __builtin_amdgcn_mfma_f64_16x16x4f64(act ? -ra : 0.0, act ? rb : 0.0, T[sl], 0, 0, 0);
                        }
                        if (e0 < 3) {
                            const double pd = pb[2 * (64 * I0 + lane)];
                            Dg = __builtin_amdgcn_mfma_f64_16x16x4f64(-pd, pd, Dg, 0, 0, 0);
                        }
                    }
                    // the panel's rows of U into their tiles
#pragma unroll
                    for (int sl = 0; sl < kSlots; ++sl) {
                        if (!sv[sl] || sI[sl] != I0) continue;
                        T[sl][e0] = pb[2 * (64 * sJ[sl] + lane)];
                    }
                }
            }
            if (!spd) return false;
            // U back to LDS (packed) for the triangular solves
#pragma unroll
            for (int sl = 0; sl < kSlots; ++sl) {
                if (!sv[sl]) continue;
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const int r = 16 * sI[sl] + 4 * e + g4, cc = 16 * sJ[sl] + r16;
                    if (r <= cc && cc < n)
                        H[bpidx(r, cc)] = T[sl][e];
                    else if (aug && cc == n && r < n)
                        dv[r] = T[sl][e];
                }
            }
            if (!aug)
                for (int e = tid; e < n; e += kBlockThreads) dv[e] = g[e];
            __syncthreads();
            // phase end: factor
            // phase: backsub
            // U^T y = g (unless eliminated with the factor) then U x = y
            if (!aug) forward_solve();
            back_solve();
            // phase end: backsub
            return true;
        };

        double* g = gA;
        double* gnew = gB;
        // gradients evaluated here (the residual check counts the one it gives)
        int ngrad = a.grad_at_lam ? 0 : 1;
        // Ub and uproj hold U and <U phiW> at lam, the outputs' source
        bool ub_lam = true;
        double ng;
        if (a.grad_at_lam) {
            // the residual check's gradient at lambda (its projection summed
            // in point order); Lambda at the points for the first Hessian
            if (tid < Q) point_lambda(lam);
            double sq = 0.0;
            if (tid < n) {
                const int l = tid / nt, i = tid - l * nt;
                const double gv = a.grad_at_lam[cell * SM + i * S + l];
                g[tid] = gv;
                sq = gv * gv;
            }
            ng = sqrt(block_allsum(sq, red));
            ub_lam = false;
        } else {
            ng = gradient(lam, g);
        }
        int status = -1;
        int iters = 0;
        if (!c.safe && ng < c.eps) status = UQK_STATUS_CONVERGED;
        // trial point lambda - step alpha d (d in dv): its gradient into gnew
        auto trial = [&](double step) -> double {
            if (tid < n) {
                const int l = tid / nt, i = tid - l * nt;
                lamnew[i * S + l] = lam[i * S + l] - step * c.alpha * dv[tid];
            }
            __syncthreads();
            ++ngrad;
            ub_lam = false;
            return gradient(lamnew, gnew);
        };
        for (int it = 0; status < 0 && it < c.max_iter; ++it) {
            if (it > 0) {
                double* t = g;
                g = gnew;
                gnew = t;
            }
            ++iters;
            if constexpr (PREC == 2) {
                if (reuse && !refactor) {
                    // a step on the kept factor: d = P g + P (g - H P g) with
                    // P = (U^T U)^-1 of the kept FP32 factor and the FP64
                    // Hessian at lambda; kept when the full step lowers
                    // ||grad L|| (the FP64 accept test), else lambda stays and
                    // the next iteration factors its own Hessian
                    if (!flds) {
                        // the slot holds U, or W once a step has used it
                        float* fs = a.pinv + cell * fslot;
                        const int hp = n * (n + 1) / 2;
                        for (int t = tid; t < hp; t += kBlockThreads) H[t] = (double)fs[t];
                        if (!(kept_tag & kKeptInverted)) {
                            __syncthreads();
                            for (int t = tid; t < n; t += kBlockThreads) rdiag[t] = 1.0 / H[bpidx(t, t)];
                            __syncthreads();
                            invert_factor();
                            for (int t = tid; t < hp; t += kBlockThreads) fs[t] = (float)H[t];
                            if (tid == 0) a.pinv_ok[cell] = nt | kKeptInverted;
                        }
                        flds = true;
                    }
                    for (int t = tid; t < n; t += kBlockThreads) dv[t] = g[t];
                    __syncthreads();
                    solve_kept();
                    const double d1 = tid < n ? dv[tid] : 0.0;
                    hessvec(dv, gnew);  // (the trial gradient rewrites gnew)
                    if (tid < n) dv[tid] = g[tid] - gnew[tid];
                    __syncthreads();
                    solve_kept();
                    if (tid < n) dv[tid] += d1;
                    __syncthreads();
                    double Lk0[S];
#pragma unroll
                    for (int s = 0; s < S; ++s) Lk0[s] = Lk[s];
                    const double nnew = trial(1.0);
                    if (nnew < ng) {
                        ++w_reuse;
                        double* t = lam;
                        lam = lamnew;
                        lamnew = t;
                        ub_lam = true;
                        const double ratio = nnew / ng;
                        ng = nnew;
                        // contraction clearly worse than a (damped) Newton
                        // step's: factor afresh at the next Newton step.
                        // From n = 64 on, a factorisation (Hessian O(Q n^2),
                        // Cholesky O(n^3)) costs over ten steps on the kept
                        // factor (O(n^2 + Q n)), and a step that contracts
                        // by 0.05 is still worth keeping
                        const double slack = n >= 64 ? 0.05 : 1e-3;
                        if (!(ratio <= 1.2 * (1.0 - c.alpha) + slack)) refactor = true;
                        if (nnew < c.eps) status = UQK_STATUS_CONVERGED;
                    } else {
#pragma unroll
                        for (int s = 0; s < S; ++s) Lk[s] = Lk0[s];
                        // g stays: the swap at the next iteration's start
                        // brings it back
                        double* t = g;
                        g = gnew;
                        gnew = t;
                        refactor = true;
                    }
                    continue;
                }
            }
            hessian();
            if (!cholesky_solve(g)) {
                status = UQK_STATUS_NOT_SPD;
                break;
            }
            ++w_factor;
            factored = reuse;
            if (c.max_iter == 1) {  // one-shot IPM (closure.cpp:151-154)
                if (tid < n) {
                    const int l = tid / nt, i = tid - l * nt;
                    lam[i * S + l] -= c.alpha * dv[tid];
                }
                __syncthreads();
                ub_lam = false;
                status = UQK_STATUS_CONVERGED;
                break;
            }
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
                    ub_lam = true;
                }
                break;
            }
            double* t = lam;
            lam = lamnew;
            lamnew = t;
            ub_lam = true;
            ng = nnew;
            if (!isfinite(nnew)) {
                status = UQK_STATUS_NONFINITE;
                break;
            }
            if (nnew < c.eps) status = UQK_STATUS_CONVERGED;
        }
        if (status < 0) status = UQK_STATUS_MAX_ITERATIONS;
        if (a.uq_out && !ub_lam) {
            // the outputs come from U at the returned lambda: evaluated again
            // where the last gradient was another point's (a rejected step, a
            // failed cell, or the check's gradient alone)
            ++ngrad;
            gradient(lam, gnew);
        }
        if constexpr (PREC == 2) {
            // the slot for the next steps: this call's last factor U (in H),
            // rounded to FP32 and tagged with the cell's order; the first
            // step that uses it inverts it (W = U^-1, P = W W^T) and keeps W
            // in its place.  No factor after a step on the kept one did
            // poorly.
            if (reuse) {
                const bool keep = factored && status != UQK_STATUS_NOT_SPD;
                if (keep) {
                    float* fs = a.pinv + cell * fslot;
                    const int hp = n * (n + 1) / 2;
                    for (int t = tid; t < hp; t += kBlockThreads) fs[t] = (float)H[t];
                }
                if (tid == 0 && (keep || refactor)) a.pinv_ok[cell] = keep ? nt : 0;
            }
        }
        for (int t = tid; t < SM; t += kBlockThreads) lg[t] = lam[t];
        if (a.uq_out) {
            // the step's next two sweeps from the gradient at the returned
            // lambda: uQ = U(lambda phi) (momentsolver.cpp:100)
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
                const double wm = wave_min(m);
                w_dt = fmin(w_dt, wm);  // per wavefront; combined below
                if (a.dt_cell) {  // the cell's ComputeDt, for the next step's check
                    if (lane == 0) red[4 + wave] = wm;
                    __syncthreads();
                    if (tid == 0) a.dt_cell[cell] = fmin(fmin(red[4], red[5]), fmin(red[6], red[7]));
                }
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
        if (a.nfactor) {
            if (w_factor) atomicAdd(a.nfactor, w_factor);
            if (w_reuse) atomicAdd(a.nfactor + 1, w_reuse);
        }
    }
    // positive doubles order as their bit patterns; one atomic per wavefront
    if (a.dt_min && lane == 0 && w_dt < 1e10) atomicMin(a.dt_min, (unsigned long long)__double_as_longlong(w_dt));
}

size_t newton_block_lds(int S, int M, int Q) {
    const int SM = S * M, P = S * (S + 1) / 2;
    return sizeof(double) * ((size_t)SM * (SM + 1) / 2 + 6 * (size_t)SM + (size_t)S * Q + (size_t)Q * P + SM + 8 + 256 + 1024 + 2);
}

// the instantiation for a launch: FP64, mixed, or mixed with kept factors
inline int block_prec(int prec, int pinv_mode) { return prec == UQK_FP64 ? 0 : (pinv_mode == 1 ? 2 : 1); }

template <int KIND>
static int block_occupancy(int M, int Q, int prec, int pinv_mode = 0) {
    constexpr int S = KindInfo<KIND>::S;
    const size_t shmem = newton_block_lds(S, M, Q);
    int occ = 0;
    hipError_t e;
    switch (block_prec(prec, pinv_mode)) {
        case 0: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, newton_block_kernel<KIND, 0>, kBlockThreads, shmem); break;
        case 1: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, newton_block_kernel<KIND, 1>, kBlockThreads, shmem); break;
        default: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, newton_block_kernel<KIND, 2>, kBlockThreads, shmem);
    }
    return e == hipSuccess && occ > 0 ? occ : 1;
}

template <int KIND>
static hipError_t block_go(const NewtonArgs& args, int max_lds, int num_cu, hipStream_t st, const LaunchShape& shape) {
    constexpr int S = KindInfo<KIND>::S;
    const size_t shmem = newton_block_lds(S, args.M, args.Q);
    if (shmem > (size_t)max_lds || S * args.M > kBlockMaxN || args.Q > kBlockThreads) return hipErrorInvalidValue;
    const int occ = block_occupancy<KIND>(args.M, args.Q, args.prec, args.pinv_mode);
    // persistent workgroups per CU: the resident maximum, or the tuned count
    const int per_cu = shape.blocks_per_cu > 0 && shape.blocks_per_cu < occ ? shape.blocks_per_cu : occ;
    long blocks = args.ncell < (long)per_cu * num_cu ? args.ncell : (long)per_cu * num_cu;
    if (blocks < 1) blocks = 1;
    NewtonArgs a = args;
    // the queue counts workgroups here: one cell per workgroup at a time
    set_queue(a, blocks);
    switch (block_prec(a.prec, a.pinv_mode)) {
        case 0: hipLaunchKernelGGL((newton_block_kernel<KIND, 0>), dim3((unsigned)blocks), dim3(kBlockThreads), shmem, st, a); break;
        case 1: hipLaunchKernelGGL((newton_block_kernel<KIND, 1>), dim3((unsigned)blocks), dim3(kBlockThreads), shmem, st, a); break;
        default: hipLaunchKernelGGL((newton_block_kernel<KIND, 2>), dim3((unsigned)blocks), dim3(kBlockThreads), shmem, st, a);
    }
    return hipGetLastError();
}

bool newton_block_fits(int S, int M, int Q, int max_lds) {
    return S * M <= kBlockMaxN && Q <= kBlockThreads && newton_block_lds(S, M, Q) <= (size_t)max_lds;
}

hipError_t newton_block_launch(int kind, const NewtonArgs& a, int max_lds, int num_cu, hipStream_t st,
                               const LaunchShape& shape) {
    switch (kind) {
        case UQK_BOUNDED_BARRIER: return block_go<UQK_BOUNDED_BARRIER>(a, max_lds, num_cu, st, shape);
        case UQK_EULER1D: return block_go<UQK_EULER1D>(a, max_lds, num_cu, st, shape);
        case UQK_EULER2D: return block_go<UQK_EULER2D>(a, max_lds, num_cu, st, shape);
        case UQK_SHALLOW_WATER2D: return block_go<UQK_SHALLOW_WATER2D>(a, max_lds, num_cu, st, shape);
        case UQK_SHALLOW_WATER2D_BARRIER: return block_go<UQK_SHALLOW_WATER2D_BARRIER>(a, max_lds, num_cu, st, shape);
    }
    return hipErrorInvalidValue;
}

int newton_block_occupancy(int kind, int M, int Q, int prec) {
    switch (kind) {
        case UQK_BOUNDED_BARRIER: return block_occupancy<UQK_BOUNDED_BARRIER>(M, Q, prec);
        case UQK_EULER1D: return block_occupancy<UQK_EULER1D>(M, Q, prec);
        case UQK_EULER2D: return block_occupancy<UQK_EULER2D>(M, Q, prec);
        case UQK_SHALLOW_WATER2D: return block_occupancy<UQK_SHALLOW_WATER2D>(M, Q, prec);
        case UQK_SHALLOW_WATER2D_BARRIER: return block_occupancy<UQK_SHALLOW_WATER2D_BARRIER>(M, Q, prec);
    }
    return 1;
}

}  // namespace uqk
