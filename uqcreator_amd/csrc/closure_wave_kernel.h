// closure_wave_kernel.h -- the one-wavefront Newton closure kernel template
// and its launch helpers, shared by closure_wave.hip (FP64 and mixed
// instantiations, the dispatch) and closure_wave_reuse.hip (the mixed
// closure with the reuse of the factorisation across time steps, compiled on
// its own).
#pragma once
// closure_wave.hip -- batched FP64 Newton dual closure, one wavefront per
// cell, specialised at compile time for the (closure, M, Q) of a config with
// n = S*M <= 63 duals and Q <= 64 quadrature points.
//
// Same algorithm as closure_fp64.hip (Closure::SolveClosure / SolveClosureSafe,
// closure.cpp:133-193, :297-339), laid out for wave64:
//   * lane k owns quadrature point k: it keeps phi(k, :) in registers and
//     evaluates Lambda(k) = lambda^T phi(k), U and DU there;
//   * lane e owns dual e = l*M + i of the gradient, trial gradient and Newton
//     direction (registers), so the line search never touches LDS for them;
//   * Hessian (closure.cpp:255-271): lanes own basis pairs (j <= i) and sum all
//     S(S+1)/2 state blocks over k; DU_k is computed by lane k and reaches the
//     other lanes as SGPR operands (v_readlane), which costs no LDS bandwidth;
//   * Cholesky (dposv 'U', vectorspace.cpp:152-171): lane c holds column c of
//     the upper triangle in registers, lane n holds the right-hand side, so the
//     right-looking elimination also performs the forward substitution; row j
//     of U is broadcast through SGPRs (v_readlane) at every step;
//   * back substitution: lane r holds y_r, column i of U is read from LDS.
#include <hip/hip_runtime.h>

#include "uqk_device.h"
#include "uqk_internal.h"

namespace uqk {

__device__ __forceinline__ double readlane_f64(double v, int lane) {
    long long x = __double_as_longlong(v);
    int lo = __builtin_amdgcn_readlane((int)x, lane);
    int hi = __builtin_amdgcn_readlane((int)(x >> 32), lane);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// offset of H(r, c), r <= c, in the packed upper triangle
__device__ __forceinline__ int hidx(int r, int c) { return c * (c + 1) / 2 + r; }

// 1/sqrt(h) from v_rsq_f64 refined by two Newton steps
__device__ __forceinline__ double rsqrt_f64(double h) {
    double r = __builtin_amdgcn_rsq(h);
    r = r * (1.5 - 0.5 * h * r * r);
    r = r * (1.5 - 0.5 * h * r * r);
    return r;
}


template <int KIND, int M, int Q>
struct WaveShape {
    static constexpr int S = KindInfo<KIND>::S;
    static constexpr int N = S * M;          // duals per cell
    static constexpr int P = S * (S + 1) / 2; // distinct DU entries
    static constexpr int QP = (Q + 1) & ~1;   // padded row length (b128 reads)
    static constexpr int NPAIR = M * (M + 1) / 2;
    static constexpr int HP = N * (N + 1) / 2; // packed upper triangle
    // per-wave LDS doubles
    static constexpr int LAM = 0;
    static constexpr int LAMNEW = LAM + S * M;
    static constexpr int UU = LAMNEW + S * M;
    static constexpr int UB = UU + S * M;        // U values, S rows of QP
    static constexpr int HS = UB + S * QP;       // packed H / U factor
    static constexpr int YB = HS + HP;           // 64
    static constexpr int DIAG = YB + 64;         // 64 reciprocal pivots
    static constexpr int ROW = (DIAG + 64 + 1) & ~1;  // 2 x 128: rows j, j+1 of U, double buffered (16-byte aligned)
    static constexpr int WAVE_DOUBLES = ROW + 256;
    static constexpr int TABLE_DOUBLES = 2 * M * QP;  // phiT, phiwT (i-major)
};

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double d2 __attribute__((ext_vector_type(2)));
typedef double f64x4 __attribute__((ext_vector_type(4)));

// PREC 0: FP64 Hessian on the vector units.  PREC 1: Hessian on MFMA
// (v_mfma_f32_16x16x4f32, FP32 operands and accumulation); gradient, residual,
// accept test, Cholesky and solve stay FP64.  PREC 2: PREC 1 with the reuse
// of the factorisation across time steps (the time loop's mixed closure; its
// own instantiation, so that the single-call closure keeps its registers).
// PREC 4: PREC 1 with the reuse of the factorisation within one closure call
// (a cell's later Newton steps solve with its last factor while they
// contract).
template <int KIND, int M, int Q, int PREC>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) newton_wave_kernel(NewtonArgs a) {
    using W = WaveShape<KIND, M, Q>;
    constexpr int S = W::S, N = W::N, P = W::P, QP = W::QP;
    static_assert(N + 1 <= 64 && Q <= 64, "wave kernel needs n < 64 and Q <= 64");
    static_assert(M <= 16 && S <= 16, "the gradient's MFMA product takes one 16 x 16 tile");
    extern __shared__ double smem[];
    if (a.active && *a.active == 0) return;
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    double* phiT = smem;              // phiT[i*QP + k]  = phi(k, i)
    double* phiwT = smem + M * QP;    // phiwT[i*QP + k] = phiw(k, i)
    for (int t = threadIdx.x; t < M * QP; t += blockDim.x) {
        int i = t / QP, k = t - i * QP;
        phiT[t] = k < Q ? a.phi[k * M + i] : 0.0;
        phiwT[t] = k < Q ? a.phiw[k * M + i] : 0.0;
    }
    __syncthreads();
    double* wb = smem + W::TABLE_DOUBLES + wave * W::WAVE_DOUBLES;
    double* lamA = wb + W::LAM;
    double* lamB = wb + W::LAMNEW;
    double* uu = wb + W::UU;
    double* Ub = wb + W::UB;
    double* Hs = wb + W::HS;
    double* yb = wb + W::YB;
    double* rdiag = wb + W::DIAG;
    double* rowb = wb + W::ROW;

    // lane k evaluates quadrature point k (lanes past Q repeat the last point)
    const int kq = lane < Q ? lane : Q - 1;

    // this lane's dual index e = l*M + i (valid for lane < N)
    const int el = lane < N ? lane / M : 0;
    const int ei = lane < N ? lane - el * M : 0;
    // this lane's basis pairs (j <= i), pair = lane + 64 * pass
    int pj[2], pi[2];
#pragma unroll
    for (int pass = 0; pass < 2; ++pass) {
        int e = lane + 64 * pass;
        if (e >= W::NPAIR) e = 0;
        int i = (int)((sqrt(8.0 * e + 1.0) - 1.0) * 0.5);
        while ((i + 1) * (i + 2) / 2 <= e) ++i;
        while (i * (i + 1) / 2 > e) --i;
        pi[pass] = i;
        pj[pass] = e - i * (i + 1) / 2;
    }
    const ClosureParams& c = a.c;
    const int SM = S * M;

    // MFMA operands of the Hessian blocks H_lm = Phi^T diag(DU_lm) PhiW:
    // A(j, k) = phi(k, j) and the constant factor PhiW(k, i) of B(k, i), in the
    // 16x16x4 lane layout (lane l: row/column l % 16, k = 4 ks + l / 16).
    constexpr int KS = (Q + 3) / 4;
    constexpr int TM = (M + 15) / 16;
    // mop[(t*KS + ks)*64 + lane] = phi, mop[TM*KS*64 + ...] = phiw, FP32
    float* mop = reinterpret_cast<float*>(smem + W::TABLE_DOUBLES + (blockDim.x >> 6) * W::WAVE_DOUBLES);
    if constexpr (PREC >= 1) {
        for (int t2 = threadIdx.x; t2 < TM * KS * 64; t2 += blockDim.x) {
            const int ln = t2 & 63, ks = (t2 >> 6) % KS, t = (t2 >> 6) / KS;
            const int k = 4 * ks + (ln >> 4), j = 16 * t + (ln & 15);
            const bool in = k < Q && j < M;
            mop[t2] = in ? (float)a.phi[(in ? k : 0) * M + (in ? j : 0)] : 0.f;
            mop[TM * KS * 64 + t2] = in ? (float)a.phiw[(in ? k : 0) * M + (in ? j : 0)] : 0.f;
        }
        __syncthreads();
    }

    // Work queue, guided: a wave takes (remaining / 2 waves) queue positions
    // with one atomicAdd, so chunks start large (few atomics) and shrink to
    // single positions at the end (balance).  Position t of the 2^qbits queue
    // is cell bitreverse(t) (skipped when >= ncell): neighbouring costly cells
    // go to different chunks.  After a residual check the queue runs over
    // the *cell_count listed cells instead: position t is
    // cell_list[bitreverse(t)] (skipped past the count).  Newton and failure
    // counts are summed per wave (wave-uniform, scalar) and added once at the
    // end.  The next cell is taken from the queue while the current one
    // starts, and its moments, duals and flags are loaded before the current
    // cell's results are stored, so neither latency stands between cells.
    unsigned long long w_fail = 0, w_newton = 0, w_grad = 0, w_took = 0, w_factor = 0, w_reuse = 0;
    double w_dt = 1e10;  // min ComputeDt over this wave's cells (fused path)
    const long nlist = a.cell_list ? (long)__builtin_amdgcn_readfirstlane((int)*a.cell_count) : 0;
    int lbits = 0;
    while ((1L << lbits) < nlist) ++lbits;
    const int qb = a.cell_list ? lbits : a.qbits;
    const long qsize = 1L << qb;
    long qpos = 0, qend = 0;
    // the next cell of the queue, -1 when it is exhausted (listed cells are
    // valid by construction)
    auto fetch = [&]() -> long {
        for (;;) {
            if (qpos == qend) {
                long want = (qsize - qend) / (2 * a.waves);
                if (want < 1) want = 1;
                unsigned long long got = 0;
                if (lane == 0) got = atomicAdd(a.counter, (unsigned long long)want);
                const long base =
                    (long)(((unsigned long long)__builtin_amdgcn_readfirstlane((unsigned)(got >> 32)) << 32) |
                           (unsigned)__builtin_amdgcn_readfirstlane((unsigned)got));
                if (base >= qsize) return -1;
                qpos = base;
                qend = base + want < qsize ? base + want : qsize;
            }
            const unsigned pos = (unsigned)qpos;
            ++qpos;
            const long qcell = qb ? (long)(__builtin_bitreverse32(pos) >> (32 - qb)) : 0;
            if (a.cell_list) {
                if (qcell >= nlist) continue;
                return (long)__builtin_amdgcn_readfirstlane(a.cell_list[qcell]);
            }
            if (qcell < a.ncell) return qcell;
        }
    };
    // a cell's inputs, loaded ahead (lane t < n holds entry t of the moments,
    // the duals and the learned start; lane e = s M + i the residual check's
    // gradient entry e): the kept-inverse flag and the ComputeDt length are
    // wave-uniform
    double u_n = 0.0, l_n = 0.0, la_n = 0.0, dx_n = 0.0, g_n = 0.0;
    int kf_n = 0;
    auto prefetch = [&](long cl) {
        if (cl < 0) return;
        if (lane < SM) {
            u_n = a.u[cl * SM + lane];
            l_n = a.lam[cl * SM + lane];
            if (a.lam_alt) la_n = a.lam_alt[cl * SM + lane];
        }
        if (a.grad_at_lam && lane < N) g_n = a.grad_at_lam[cl * SM + ei * S + el];
        if (a.uq_out && a.dt_min) dx_n = a.dx_param[cl];
        if constexpr (PREC == 2) kf_n = (a.pinv_mode == 1 && cl < a.pinv_slots) ? a.pinv_ok[cl] : 0;
    };
    long next = fetch();
    prefetch(next);
    while (next >= 0) {
        const long cell = next;
        const double la_c = la_n, dx_c = dx_n, g_c = g_n;
        const int kf_c = kf_n;
        double* lg = a.lam + cell * SM;
        if (lane < SM) {
            uu[lane] = u_n;
            lamA[lane] = l_n;
        }
        next = fetch();
        wave_sync();
        const double ue = lane < N ? uu[ei * S + el] : 0.0;

        double Lk[S];  // Lambda at this lane's point for the last gradient
        double uproj = 0.0;  // <U phiW> of this lane's dual at the last gradient
        // sum_k Ub(s, k) phiW(k, i) for this lane's dual e = s M + i, as one
        // FP64 MFMA product (S x Q)(Q x M): A(s, k) = Ub(s, k), B(k, i) =
        // phiW(k, i) in 16 x 16 x 4 steps (lane l: row/column l % 16,
        // k = 4 ks + l / 16); D(s, i) is element e of lane 16 (s - 4e) + i for
        // s = 4e + l / 16, and goes to lane s M + i through the solve buffer yb.
        auto project_ub = [&]() -> double {
            f64x4 pacc = {0.0, 0.0, 0.0, 0.0};
            {
                const int r16 = lane & 15, kq4 = lane >> 4;
#pragma unroll
                for (int ks = 0; ks < KS; ++ks) {
                    const int k = 4 * ks + kq4;
                    const double av = (r16 < S && k < Q) ? Ub[r16 * QP + k] : 0.0;
                    const double bv = (r16 < M && k < Q) ? phiwT[r16 * QP + k] : 0.0;
                    pacc = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, pacc, 0, 0, 0);
                }
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const int sr = 4 * e + kq4;  // f64 layout: element e at (4e + l/16, l%16)
                    if (sr < S && r16 < M) yb[sr * M + r16] = pacc[e];
                }
            }
            wave_sync();
            const double acc = lane < N ? yb[lane] : 0.0;
            wave_sync();
            return acc;
        };
        // Lambda at this lane's point for the duals X (LDS)
        auto point_lambda = [&](const double* X, double* L) {
#pragma unroll
            for (int s = 0; s < S; ++s) L[s] = 0.0;
#pragma unroll
            for (int i = 0; i < M; ++i) {
                const double f = phiT[i * QP + kq];
#pragma unroll
                for (int s = 0; s < S; ++s) L[s] += X[i * S + s] * f;
            }
        };
        // gradient at X (LDS); returns ||g||, leaves g_e in gout, Lambda in Lk
        auto gradient = [&](const double* X, double& gout) -> double {
            // phase: gradient
            double L[S], Uv[S];
            point_lambda(X, L);
            closure_U<KIND>(c, L, Uv);
            if (lane < Q) {
#pragma unroll
                for (int s = 0; s < S; ++s) Ub[s * QP + lane] = Uv[s];
            }
#pragma unroll
            for (int s = 0; s < S; ++s) Lk[s] = L[s];
            wave_sync();
            // <U phiW> of all duals (the solve buffer yb is idle here)
            const double acc = project_ub();
            uproj = acc;
            double g = lane < N ? acc - ue : 0.0;
            gout = g;
            double nrm = sqrt(wave_sum(g * g));
            wave_sync();
            // phase end: gradient
            return nrm;
        };

        // Hessian at the iterate of the last gradient -> Hs (packed upper).
        // DU at point k lives in lane k; it reaches every lane through
        // v_readlane (an SGPR operand), not through LDS.
        auto hessian_mfma = [&]() {
            // phase: hessian
            double y[P];
            closure_DU<KIND>(c, Lk, y);
            float yf[P];
#pragma unroll
            for (int p = 0; p < P; ++p) yf[p] = (float)y[p];
            f32x4 acc[P][TM][TM];
#pragma unroll
            for (int p = 0; p < P; ++p)
#pragma unroll
                for (int tj = 0; tj < TM; ++tj)
#pragma unroll
                    for (int ti = 0; ti < TM; ++ti) acc[p][tj][ti] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int ks = 0; ks < KS; ++ks) {
                const int kk = 4 * ks + (lane >> 4);  // < 64: DU of point kk lives in lane kk
                float yk[P], aop[TM], bw[TM];
#pragma unroll
                for (int p = 0; p < P; ++p) yk[p] = __shfl(yf[p], kk, 64);
#pragma unroll
                for (int t = 0; t < TM; ++t) {
                    aop[t] = mop[(t * KS + ks) * 64 + lane];
                    bw[t] = mop[TM * KS * 64 + (t * KS + ks) * 64 + lane];
                }
#pragma unroll
                for (int p = 0; p < P; ++p)
#pragma unroll
                    for (int ti = 0; ti < TM; ++ti) {
                        const float bop = yk[p] * bw[ti];
#pragma unroll
                        for (int tj = 0; tj < TM; ++tj)
                            acc[p][tj][ti] =
                                __builtin_amdgcn_mfma_f32_16x16x4f32(aop[tj], bop, acc[p][tj][ti], 0, 0, 0);
                    }
            }
            // D(j, i) at lane l: j = 16 tj + 4 (l / 16) + e, i = 16 ti + l % 16
#pragma unroll
            for (int l = 0; l < S; ++l)
#pragma unroll
                for (int m = 0; m <= l; ++m)
#pragma unroll
                    for (int tj = 0; tj < TM; ++tj)
#pragma unroll
                        for (int ti = 0; ti < TM; ++ti)
#pragma unroll
                            for (int e = 0; e < 4; ++e) {
                                const int j = 16 * tj + 4 * (lane >> 4) + e, i = 16 * ti + (lane & 15);
                                if (j < M && i < M && (m < l || j <= i)) {
                                    const int r = m * M + j, cc = l * M + i;
                                    Hs[hidx(r, cc)] = (double)acc[l * (l + 1) / 2 + m][tj][ti][e];
                                }
                            }
            wave_sync();
            // phase end: hessian
        };

        auto hessian = [&]() {
            if constexpr (PREC >= 1) {
                hessian_mfma();
                return;
            }
            // phase: hessian
            double y[P];
            closure_DU<KIND>(c, Lk, y);
            constexpr int PASSES = W::NPAIR > 64 ? 2 : 1;
            double acc[PASSES][P];
#pragma unroll
            for (int q = 0; q < PASSES; ++q)
#pragma unroll
                for (int p = 0; p < P; ++p) acc[q][p] = 0.0;
            const double* fj0 = phiT + pj[0] * QP;
            const double* wi0 = phiwT + pi[0] * QP;
            const double* fj1 = phiT + pj[PASSES - 1] * QP;
            const double* wi1 = phiwT + pi[PASSES - 1] * QP;
#pragma unroll 7
            for (int k = 0; k < Q; ++k) {
                double du[P];
#pragma unroll
                for (int p = 0; p < P; ++p) du[p] = readlane_f64(y[p], k);
                const double pk0 = fj0[k] * wi0[k];
#pragma unroll
                for (int p = 0; p < P; ++p) acc[0][p] += pk0 * du[p];
                if constexpr (PASSES == 2) {
                    const double pk1 = fj1[k] * wi1[k];
#pragma unroll
                    for (int p = 0; p < P; ++p) acc[PASSES - 1][p] += pk1 * du[p];
                }
            }
#pragma unroll
            for (int q = 0; q < PASSES; ++q) {
                const int e = lane + 64 * q;
                const int j = pj[q], i = pi[q];
                if (e < W::NPAIR) {
#pragma unroll
                    for (int l = 0; l < S; ++l)
#pragma unroll
                        for (int m = 0; m <= l; ++m) {
                            double v = acc[q][l * (l + 1) / 2 + m];
                            int r1 = m * M + j, c1 = l * M + i;
                            Hs[hidx(r1, c1)] = v;
                            if (m < l) {
                                int r2 = m * M + i, c2 = l * M + j;
                                Hs[hidx(r2, c2)] = v;
                            }
                        }
                }
            }
            wave_sync();
            // phase end: hessian
        };

        // d <- H^{-1} g; returns false if H is not positive definite
        auto cholesky_solve = [&](double g, double& d) -> bool {
            // phase: factor
            // [H | g] (upper triangle and the right-hand side as column N) is
            // spread over the wave in registers: lane 8p + q holds rows
            // 16 ra + 2p + i (i = 0, 1) of columns 8b + q in T[ra][i][b], i.e.
            // consecutive row pairs cyclically over the 8 row groups and
            // columns cyclically over the 8 column groups.  A step eliminates
            // two pivots: their 2x2 diagonal block is factored in scalar
            // registers, the 8 owner lanes of the two rows scale them and
            // publish them in an LDS row buffer, and every lane applies the
            // rank-2 update to its tiles (the two rank-1 updates in order).
            // Per lane this is ~n^3/(6*64) FMAs and one LDS round trip per two
            // pivots.  Elimination of column N is the forward substitution.
            constexpr int NS = (N + 1 + 7) / 8;   // column slots
            constexpr int RA = (N + 1 + 15) / 16; // row-pair slots
            const int p8 = lane >> 3, q8 = lane & 7;
            double T[RA][2][NS];
            yb[lane] = g;
            wave_sync();
#pragma unroll
            for (int ra = 0; ra < RA; ++ra)
#pragma unroll
                for (int i = 0; i < 2; ++i)
#pragma unroll
                    for (int tb = 0; tb < NS; ++tb) {
                        const int r = 16 * ra + 2 * p8 + i, cc = 8 * tb + q8;
                        double v = 0.0;
                        if (r <= cc && cc < N) v = Hs[cc * (cc + 1) / 2 + r];
                        else if (cc == N && r < N) v = yb[r];
                        T[ra][i][tb] = v;
                    }
            bool ok = true;
#pragma unroll
            for (int j = 0; j < N; j += 2) {
                const bool two = j + 1 < N;          // the last step of an odd n has one pivot
                const int pj = (j & 15) >> 1, aj = j >> 4;
                const int l0 = 8 * pj + (j & 7), l1 = 8 * pj + ((j + 1) & 7);
                // 2x2 pivot block [h00 h01; h01 h11]: U_j+1,j+1^2 = det / h00,
                // so 1/U_j+1,j+1 = sqrt(h00) / sqrt(det) and both square roots
                // start at once instead of one after the other
                const double h00 = readlane_f64(T[aj][0][j >> 3], l0);
                if (!(h00 > 0.0)) {
                    ok = false;
                    break;
                }
                double r1 = 0.0, u01 = 0.0, rdet = 0.0;
                double h01 = 0.0, det = 0.0;
                if (two) {
                    h01 = readlane_f64(T[aj][0][(j + 1) >> 3], l1);
                    det = fma(h00, readlane_f64(T[aj][1][(j + 1) >> 3], l1), -h01 * h01);
                    if (!(det > 0.0)) {
                        ok = false;
                        break;
                    }
                    rdet = rsqrt_f64(det);
                }
                const double r0 = rsqrt_f64(h00);
                if (two) {
                    u01 = h01 * r0;
                    r1 = rdet * (h00 * r0);
                }
                if (lane == 0) {
                    rdiag[j] = r0;
                    if (two) rdiag[j + 1] = r1;
                }
                // rows j, j+1 interleaved: (U_j,c, U_j+1,c) at urow[2c], 16-byte reads
                d2* urow = reinterpret_cast<d2*>(rowb + 128 * ((j >> 1) & 1));
                if (p8 == pj) {
                    // the same formulas give the diagonal (U_jj = h00 r0,
                    // U_j+1,j+1 = (h11 - u01^2) r1); columns left of the pivots
                    // get values below the diagonal, which nothing reads
#pragma unroll
                    for (int tb = j >> 3; tb < NS; ++tb) {
                        const int cc = 8 * tb + q8;
                        const double v0 = T[aj][0][tb] * r0;
                        T[aj][0][tb] = v0;
                        double v1 = 0.0;
                        if (two) {
                            v1 = (T[aj][1][tb] - u01 * v0) * r1;
                            T[aj][1][tb] = v1;
                        }
                        urow[cc] = d2{v0, v1};
                    }
                }
                wave_sync();
                double ur0[RA][2], ur1[RA][2], uc0[NS], uc1[NS];
#pragma unroll
                for (int ra = aj; ra < RA; ++ra)
#pragma unroll
                    for (int i = 0; i < 2; ++i) {
                        const d2 v = urow[16 * ra + 2 * p8 + i];
                        ur0[ra][i] = v.x;
                        ur1[ra][i] = v.y;
                    }
#pragma unroll
                for (int tb = j >> 3; tb < NS; ++tb) {
                    const d2 v = urow[8 * tb + q8];
                    uc0[tb] = v.x;
                    uc1[tb] = v.y;
                }
                // Only rows past the pivots must keep their values; entries
                // below the diagonal or past row N-1 / column N are never read
                // (pivots, published rows and the solve take r <= c), so they
                // are updated along without a per-entry test.
                const int jl = two ? j + 1 : j;  // last eliminated pivot
#pragma unroll
                for (int ra = aj; ra < RA; ++ra)
#pragma unroll
                    for (int i = 0; i < 2; ++i) {
                        const int r = 16 * ra + 2 * p8 + i;
                        if (r > jl) {
#pragma unroll
                            for (int tb = j >> 3; tb < NS; ++tb) {
                                if (8 * tb + 7 < 16 * ra) continue;  // slot wholly below the diagonal
                                double t = T[ra][i][tb] - ur0[ra][i] * uc0[tb];
                                if (two) t -= ur1[ra][i] * uc1[tb];
                                T[ra][i][tb] = t;
                            }
                        }
                    }
            }
            if (!ok) return false;
            // phase end: factor
            // phase: backsub
            // U -> Hs (packed), y -> yb
#pragma unroll
            for (int ra = 0; ra < RA; ++ra)
#pragma unroll
                for (int i = 0; i < 2; ++i)
#pragma unroll
                    for (int tb = 0; tb < NS; ++tb) {
                        const int r = 16 * ra + 2 * p8 + i, cc = 8 * tb + q8;
                        if (r <= cc && cc < N) Hs[cc * (cc + 1) / 2 + r] = T[ra][i][tb];
                        else if (cc == N && r < N) yb[r] = T[ra][i][tb];
                    }
            wave_sync();
            // U x = y with lane r owning x_r: the chain of a step is one
            // multiply, one SGPR broadcast and one FMA; column i of U comes
            // from LDS, read ahead of the chain (its loads do not depend on it)
            double yr = lane < N ? yb[lane] : 0.0;
            const double rd = lane < N ? rdiag[lane] : 0.0;
#pragma unroll
            for (int i = N - 1; i >= 0; --i) {
                const double ui = lane < i ? Hs[i * (i + 1) / 2 + lane] : 0.0;
                const double xi = readlane_f64(yr * rd, i);
                yr = lane == i ? xi : fma(-ui, xi, yr);
            }
            d = yr;
            wave_sync();
            // phase end: backsub
            return true;
        };

        // d <- (U^T U)^{-1} v with the factor of the launch's last Cholesky
        // still in Hs and rdiag (PREC 4): U^T y = v, then U x = y, lane r
        // owning y_r / x_r; each step broadcasts x_i through SGPRs and reads
        // row i of U (forward) or column i (back) from LDS, off the chain
        auto factor_resolve = [&](double v) -> double {
            // phase: matvec
            double yr = lane < N ? v : 0.0;
            const double rd = lane < N ? rdiag[lane] : 0.0;
#pragma unroll
            for (int i = 0; i < N; ++i) {
                const double ui = (lane > i && lane < N) ? Hs[lane * (lane + 1) / 2 + i] : 0.0;
                const double xi = readlane_f64(yr * rd, i);
                yr = lane == i ? xi : fma(-ui, xi, yr);
            }
#pragma unroll
            for (int i = N - 1; i >= 0; --i) {
                const double ui = lane < i ? Hs[i * (i + 1) / 2 + lane] : 0.0;
                const double xi = readlane_f64(yr * rd, i);
                yr = lane == i ? xi : fma(-ui, xi, yr);
            }
            // phase end: matvec
            return yr;
        };

        // ---- reuse of the factorisation (mixed precision) ----------------
        // The kept approximation of the inverse Hessian is P = H^{-1} of an
        // earlier step's factorisation (the FP32-assembled Hessian, factored
        // in FP64), computed by factor_inverse_kernel after the launch that
        // factored, rounded once to FP32: a full N x N matrix with row stride
        // N + 1 -- in LDS it fills exactly the bytes of the packed FP64 factor
        // in Hs.  A factored step leaves its factor U and 1/U_rr in the cell's
        // slot for that kernel.
        float* const Pl = reinterpret_cast<float*>(Hs);
        auto load_inverse = [&](const float* Pin) {
            // phase: inverse
            // N (N + 1) is even: 8-byte pieces, all loads in flight before
            // the first LDS store (the slot is 8-byte aligned)
            constexpr int NV = N * (N + 1) / 2, R = (NV + 63) / 64, B = 7;
            const float2* src = reinterpret_cast<const float2*>(Pin);
            float2* dst = reinterpret_cast<float2*>(Pl);
#pragma unroll
            for (int r0 = 0; r0 < R; r0 += B) {
                float2 part[B];
#pragma unroll
                for (int r = r0; r < r0 + B && r < R; ++r)
                    if (r * 64 + lane < NV) part[r - r0] = src[r * 64 + lane];
#pragma unroll
                for (int r = r0; r < r0 + B && r < R; ++r)
                    if (r * 64 + lane < NV) dst[r * 64 + lane] = part[r - r0];
            }
            wave_sync();
            // phase end: inverse
        };
        auto store_factor = [&](double* Uout) {
            // phase: inverse
            for (int t = lane; t < W::HP; t += 64) Uout[t] = Hs[t];
            if (lane < N) Uout[W::HP + lane] = rdiag[lane];
            // phase end: inverse
        };
        // (P v)_e for v held per lane (lane e: v_e, zero past n): P is
        // symmetric, so lane e reads column e (consecutive lanes, consecutive
        // words: no bank conflicts) and v_c arrives through SGPRs
        // (v_readlane); FP64 accumulation in two partial sums
        auto pmatvec = [&](double v) -> double {
            // phase: matvec
            const float* col = Pl + (lane < N ? lane : 0);
            double s0 = 0.0, s1 = 0.0;
#pragma unroll
            for (int c2 = 0; c2 < N; c2 += 2) {
                s0 = fma((double)col[c2 * (N + 1)], readlane_f64(v, c2), s0);
                if (c2 + 1 < N) s1 = fma((double)col[(c2 + 1) * (N + 1)], readlane_f64(v, c2 + 1), s1);
            }
            // phase end: matvec
            return lane < N ? s0 + s1 : 0.0;
        };
        // (H v)_e of the FP64 Hessian at the iterate of the last gradient
        // (Lambda in Lk), without forming H: per point k, w = DU_k (v^T phi_k),
        // then sum_k w_s(k) phiW(k, i) -- the Hessian of closure.cpp:255-271
        // applied to v.  Overwrites Ub (the next gradient rewrites it).
        auto hessvec = [&](double v) -> double {
            // phase: hessvec
            // Ld = v^T phi_k: v_{sM+i} reaches every lane through SGPRs
            double Ld[S];
#pragma unroll
            for (int s = 0; s < S; ++s) Ld[s] = 0.0;
#pragma unroll
            for (int i = 0; i < M; ++i) {
                const double f = phiT[i * QP + kq];
#pragma unroll
                for (int s = 0; s < S; ++s) Ld[s] = fma(readlane_f64(v, s * M + i), f, Ld[s]);
            }
            double y[P];
            closure_DU<KIND>(c, Lk, y);
            wave_sync();
            if (lane < Q) {
#pragma unroll
                for (int l = 0; l < S; ++l) {
                    double w = 0.0;
#pragma unroll
                    for (int m = 0; m < S; ++m) w += y[l >= m ? l * (l + 1) / 2 + m : m * (m + 1) / 2 + l] * Ld[m];
                    Ub[l * QP + lane] = w;
                }
            }
            wave_sync();
            const double hv = project_ub();
            // phase end: hessvec
            return hv;
        };

        double* lam = lamA;
        double* lamnew = lamB;
        double g = 0.0, gnew = 0.0, d = 0.0;
        // gradients evaluated here (the residual check counts the one it gives)
        int ngrad = a.grad_at_lam ? 0 : 1;
        // Ub and uproj hold U and <U phiW> at lam, the outputs' source
        bool ub_lam = true;
        double ng;
        if (a.grad_at_lam) {
            // the residual check's gradient at lambda (its projection summed in
            // point order); Lambda at this lane's point for the first Hessian
            point_lambda(lam, Lk);
            g = lane < N ? g_c : 0.0;
            ng = sqrt(wave_sum(g * g));
            ub_lam = false;
        } else {
            ng = gradient(lam, g);
        }
        int status = -1;
        int iters = 0;
        if (a.lam_alt && (c.safe || !(ng < c.eps))) {
            // learned start: Newton starts from the candidate with the smaller
            // ||grad L|| (a NaN norm fails the comparison and never wins)
            if (lane < SM) lamnew[lane] = la_c;
            double Lk0[S];
#pragma unroll
            for (int s = 0; s < S; ++s) Lk0[s] = Lk[s];
            const double g0 = g;
            double galt = 0.0;
            wave_sync();
            ++ngrad;
            const double nalt = gradient(lamnew, galt);
            if (nalt < ng) {
                double* t = lam;
                lam = lamnew;
                lamnew = t;
                g = galt;
                ng = nalt;
                ++w_took;
                ub_lam = true;
            } else {
                ub_lam = false;
                g = g0;
#pragma unroll
                for (int s = 0; s < S; ++s) Lk[s] = Lk0[s];
            }
        }
        if (!c.safe && ng < c.eps) status = UQK_STATUS_CONVERGED;
        // reuse of the factorisation (mixed precision): the cell's kept
        // inverse P (in memory, Pg; in Hs while vlds) is usable when
        // pinv_ok[cell] == 1; a factored step stores its factor (Ug) and
        // marks the cell 2 (inverse pending) for factor_inverse_kernel
        bool reuse = false, vlds = false, refactor = true, factored = false;
        if constexpr (PREC == 2) {
            if (a.pinv_mode == 1 && c.max_iter != 1 && cell < a.pinv_slots) {
                reuse = true;
                refactor = kf_c != 1;
            }
        }
        // PREC 4 (one closure call): the factor of this cell's last Cholesky
        // stays in Hs, and the next steps solve with it (lagged steps) until
        // one contracts poorly or is rejected
        if constexpr (PREC == 4) reuse = c.max_iter != 1;
        for (int it = 0; status < 0 && it < c.max_iter; ++it) {
            if (it > 0) g = gnew;
            ++iters;
            // a lagged step (mixed precision, kept inverse P): d = P g with one
            // step of iterative refinement against the FP64 Hessian at lambda,
            // d += P (g - H d); kept when the full step lowers ||grad L|| (the
            // FP64 accept test), else lambda stays and the next iteration
            // factors its own Hessian and runs the reference Newton step
            bool lagged = false;
            double Lk0[S];
            if constexpr (PREC == 2 || PREC == 4) lagged = reuse && !refactor;
            if (lagged) {
                if constexpr (PREC == 4) {
                    d = factor_resolve(g);
#pragma unroll
                    for (int s = 0; s < S; ++s) Lk0[s] = Lk[s];
                }
                if constexpr (PREC == 2) {
                    if (!vlds) {
                        load_inverse(a.pinv + cell * (long)(N * (N + 1)));
                        vlds = true;
                    }
                    const double d1 = pmatvec(g);
                    const double hv = hessvec(d1);
                    d = d1 + pmatvec(g - hv);
#pragma unroll
                    for (int s = 0; s < S; ++s) Lk0[s] = Lk[s];
                }
            } else {
                hessian();
                vlds = false;
                if (!cholesky_solve(g, d)) {
                    status = UQK_STATUS_NOT_SPD;
                    break;
                }
                ++w_factor;
                if constexpr (PREC == 4) refactor = false;
                if constexpr (PREC == 2) {
                    if (reuse) {
                        store_factor(a.ufac + cell * (long)(W::HP + N));
                        factored = true;
                    }
                }
                if (c.max_iter == 1) {  // one-shot IPM (closure.cpp:151-154)
                    if (lane < N) lam[ei * S + el] -= c.alpha * d;
                    ub_lam = false;
                    wave_sync();
                    status = UQK_STATUS_CONVERGED;
                    break;
                }
            }
            // line search (closure.cpp:159-184): halve the step until ||grad||
            // does not grow (SolveClosureSafe: and is finite); one trial site
            double step = 1.0, nnew = 0.0;
            int halvings = 0;
            bool done = false;
            for (;;) {
                if (lane < N) lamnew[ei * S + el] = lam[ei * S + el] - step * c.alpha * d;
                wave_sync();
                ++ngrad;
                nnew = gradient(lamnew, gnew);
                ub_lam = false;
                if (halvings > 0) {
                    if (nnew < c.eps) {
                        status = UQK_STATUS_CONVERGED;
                        done = true;
                        break;
                    }
                    if (halvings > c.max_halvings) {
                        status = UQK_STATUS_LINE_SEARCH;
                        done = true;
                        break;
                    }
                }
                if (!(ng < nnew || (c.safe && !isfinite(nnew)))) break;
                if (lagged) break;  // a lagged direction is not halved
                step *= 0.5;
                ++halvings;
            }
            if (lagged) {
                if (nnew < ng) {
                    ++w_reuse;
                    double* t = lam;
                    lam = lamnew;
                    lamnew = t;
                    ub_lam = true;
                    const double ratio = nnew / ng;
                    ng = nnew;
                    // contraction clearly worse than a (damped) Newton step's:
                    // factor afresh at the next Newton step
                    if (!(ratio <= 1.2 * (1.0 - c.alpha) + 1e-3)) refactor = true;
                    if (nnew < c.eps) status = UQK_STATUS_CONVERGED;
                } else {
#pragma unroll
                    for (int s = 0; s < S; ++s) Lk[s] = Lk0[s];
                    gnew = g;
                    refactor = true;
                }
                continue;
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
            double gt;
            ++ngrad;
            gradient(lam, gt);
        }
        prefetch(next);
        // the cell's slot for the next launch: a new factor to invert (2),
        // the kept inverse still fresh (1), or none (0)
        if (PREC == 2 && reuse && lane == 0) {
            a.pinv_ok[cell] = factored ? 2 : (refactor ? 0 : 1);
            if (factored) a.plist[atomicAdd(a.plist_count, 1ull)] = (int)cell;
        }
        for (int t = lane; t < SM; t += 64) lg[t] = lam[t];
        if (a.uq_out) {
            // the step's next two sweeps, from the last gradient at the
            // returned lambda: uQ = U(lambda phi) (momentsolver.cpp:100) and
            // u = <uQ phiW> (:120-124)
            double* uqg = a.uq_out + cell * (long)Q * S;
            for (int t = lane; t < Q * S; t += 64) {
                const int k = t / S, s = t - k * S;
                uqg[t] = Ub[s * QP + k];
            }
            if (lane < N) a.u_out[cell * SM + ei * S + el] = uproj;
            if (a.dt_min) {  // momentsolver.cpp:127-133 for this cell
                double m = 1e10;
                if (lane < Q) {
                    double uk[S];
#pragma unroll
                    for (int s = 0; s < S; ++s) uk[s] = Ub[s * QP + lane];
                    m = point_dt<S>(a.dtp, uk, dx_c);
                }
                const double mc = wave_min(m);
                w_dt = fmin(w_dt, mc);
                if (a.dt_cell && lane == 0) a.dt_cell[cell] = mc;
            }
        }
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
        if (w_took && a.took) atomicAdd(a.took, w_took);
        if (a.nfactor) {
            if (w_factor) atomicAdd(a.nfactor, w_factor);
            if (w_reuse) atomicAdd(a.nfactor + 1, w_reuse);
        }
        // positive doubles order as their bit patterns
        if (a.dt_min && w_dt < 1e10) atomicMin(a.dt_min, (unsigned long long)__double_as_longlong(w_dt));
    }
}

// ---------------------------------------------------------------------------
// dispatch over the compiled shapes
// ---------------------------------------------------------------------------
template <int KIND, int M, int Q, int PREC>
static size_t wave_lds(int waves, size_t* table_bytes) {
    using W = WaveShape<KIND, M, Q>;
    const size_t mop = PREC >= 1 ? sizeof(float) * 2 * ((M + 15) / 16) * ((Q + 3) / 4) * 64 : 0;
    const size_t tb = sizeof(double) * W::TABLE_DOUBLES + mop;
    if (table_bytes) *table_bytes = tb;
    return tb + (size_t)waves * sizeof(double) * W::WAVE_DOUBLES;
}

template <int KIND, int M, int Q, int PREC>
static int wave_max_waves(int max_lds) {
    int waves = 4;
    while (waves > 1 && wave_lds<KIND, M, Q, PREC>(waves, nullptr) > (size_t)max_lds) --waves;
    return waves;
}

template <int KIND, int M, int Q, int PREC>
static int wave_occupancy(int waves) {
    int occ = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, newton_wave_kernel<KIND, M, Q, PREC>, 64 * waves,
                                                     wave_lds<KIND, M, Q, PREC>(waves, nullptr)) != hipSuccess ||
        occ <= 0)
        occ = 1;
    return occ;
}

template <int KIND, int M, int Q, int PREC>
static void wave_go(const NewtonArgs& a, int max_lds, int num_cu, hipStream_t st, hipError_t* err,
                    const LaunchShape& shape) {
    const int wmax = wave_max_waves<KIND, M, Q, PREC>(max_lds);
    int waves = shape.waves > 0 ? (shape.waves < wmax ? shape.waves : wmax) : wmax;
    const int threads = 64 * waves;
    const size_t shmem = wave_lds<KIND, M, Q, PREC>(waves, nullptr);
    const int occ = wave_occupancy<KIND, M, Q, PREC>(waves);
    const int per_cu = shape.blocks_per_cu > 0 && shape.blocks_per_cu < occ ? shape.blocks_per_cu : occ;
    long want = (a.ncell + waves - 1) / waves;
    int blocks = (int)(want < (long)per_cu * num_cu ? want : (long)per_cu * num_cu);
    if (blocks < 1) blocks = 1;
    NewtonArgs b = a;
    set_queue(b, (long)blocks * waves);
    hipLaunchKernelGGL((newton_wave_kernel<KIND, M, Q, PREC>), dim3(blocks), dim3(threads), shmem, st, b);
    *err = hipGetLastError();
}

template <int KIND, int M, int Q>
static bool wave_fits(int kind, int m, int q, int max_lds) {
    if (kind != KIND || m != M || q != Q) return false;
    using W = WaveShape<KIND, M, Q>;
    return sizeof(double) * (W::TABLE_DOUBLES + W::WAVE_DOUBLES) <= (size_t)max_lds;
}

}  // namespace uqk
