// synth.h -- synthetic closure workload, identical code on host and device.
//
// Cell j's true duals lambda* are drawn from a counter-based Philox4x32-10
// stream (Salmon et al., SC'11) with key = seed and counter = (j, attempt,
// draw), so every cell is generated independently and reproducibly:
//   rho ~ U[0.5, 2], v ~ U[-1, 1], p ~ U[0.5, 2]   (a Maxwellian)
//   lambda*(:, 0)   = EulerClosure::DS(rho, rho v, p/(gamma-1) + rho v^2/2)
//                     (eulerclosure.cpp:70-80), exact for the constant moment
//   lambda*(:, i>0) ~ N(0, 0.05^2)                   (Box-Muller)
// A draw is kept only if the temperature dual is negative at every quadrature
// point and varies there by at most 3x, and the density varies by at most 100x;
// otherwise the cell redraws with the next attempt counter.  The moments are
// u = <U(lambda*^T phi) phi> and the Newton start is lambda_0 = (DS(u_0), 0, ..)
// as in MomentSolver::SetDuals (momentsolver.cpp:473-484).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace uqk {

struct Philox4x32 {
    uint32_t v[4];
};

__host__ __device__ inline void mulhilo32(uint32_t a, uint32_t b, uint32_t* hi, uint32_t* lo) {
    uint64_t p = (uint64_t)a * (uint64_t)b;
    *hi = (uint32_t)(p >> 32);
    *lo = (uint32_t)p;
}

// Philox4x32 with 10 rounds.
__host__ __device__ inline Philox4x32 philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0,
                                                    uint32_t k1) {
    for (int r = 0; r < 10; ++r) {
        uint32_t hi0, lo0, hi1, lo1;
        mulhilo32(0xD2511F53u, c0, &hi0, &lo0);
        mulhilo32(0xCD9E8D57u, c2, &hi1, &lo1);
        uint32_t n0 = hi1 ^ c1 ^ k0, n1 = lo1, n2 = hi0 ^ c3 ^ k1, n3 = lo0;
        c0 = n0;
        c1 = n1;
        c2 = n2;
        c3 = n3;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    Philox4x32 out;
    out.v[0] = c0;
    out.v[1] = c1;
    out.v[2] = c2;
    out.v[3] = c3;
    return out;
}

// 53-bit uniform in (0, 1) from two 32-bit words.
__host__ __device__ inline double u01(uint32_t a, uint32_t b) {
    uint64_t x = ((uint64_t)a << 21) ^ (uint64_t)(b >> 11);
    x &= ((uint64_t)1 << 53) - 1;
    return ((double)x + 0.5) * (1.0 / 9007199254740992.0);
}

struct SynthParams {
    int M, Q;
    double gamma;
    double sigma;       // std of the higher duals (0.05)
    double temp_ratio;  // min over max of the temperature dual (1/3)
    double rho_ratio;   // max over min of the density (100)
    int max_attempts;
};

__host__ __device__ inline double synth_rho(double v0, double v1, double v2, double g) {
    return exp((2.0 * v0 * v2 - 2.0 * v2 * log(-v2) - 2.0 * v2 * g - v1 * v1) / (2.0 * v2 * (g - 1.0)));
}

__host__ __device__ inline void synth_ds(const double* u, double g, double* ds) {
    const double gm = -g;
    double rho = u[0], ru = u[1], ru2 = ru * ru, rE = u[2];
    ds[0] = (ru2 + gm * (2 * rho * rE - ru2)) / (-2 * rho * rE + ru2) - log(pow(rho, gm) * (rE - ru2 / (2 * rho)));
    ds[1] = -((2 * rho * ru) / (-2 * rho * rE + ru2));
    ds[2] = -(rho / (rE - ru2 / (2 * rho)));
}

// Generates one cell into lam (M x 3, reference layout), u and lam0.
// Returns the number of attempts used (max_attempts + 1 on failure).
__host__ __device__ inline int synth_cell(const SynthParams& p, const double* phi, const double* phiw, uint64_t seed,
                                          uint64_t cell, double* lam, double* u, double* lam0) {
    const int M = p.M, Q = p.Q;
    const uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
    const uint32_t c0 = (uint32_t)cell, c1 = (uint32_t)(cell >> 32);
    const double g = p.gamma;
    int attempt = 0;
    for (; attempt < p.max_attempts; ++attempt) {
        uint32_t draw = 0;
        Philox4x32 r = philox4x32_10(c0, c1, (uint32_t)attempt, draw++, k0, k1);
        double rho = 0.5 + 1.5 * u01(r.v[0], r.v[1]);
        double vel = -1.0 + 2.0 * u01(r.v[2], r.v[3]);
        r = philox4x32_10(c0, c1, (uint32_t)attempt, draw++, k0, k1);
        double pr = 0.5 + 1.5 * u01(r.v[0], r.v[1]);
        double cons[3] = {rho, rho * vel, pr / (g - 1.0) + 0.5 * rho * vel * vel};
        double ds[3];
        synth_ds(cons, g, ds);
        for (int s = 0; s < 3; ++s) lam[s] = ds[s];
        const int nh = (M - 1) * 3;
        for (int t = 0; t < nh; t += 2) {
            r = philox4x32_10(c0, c1, (uint32_t)attempt, draw++, k0, k1);
            double a = u01(r.v[0], r.v[1]), b = u01(r.v[2], r.v[3]);
            double rad = sqrt(-2.0 * log(a));
            double z0 = rad * cos(6.283185307179586 * b), z1 = rad * sin(6.283185307179586 * b);
            lam[3 + t] = p.sigma * z0;
            if (t + 1 < nh) lam[3 + t + 1] = p.sigma * z1;
        }
        // acceptance: temperature dual < 0 everywhere, ratio >= temp_ratio; density ratio <= rho_ratio
        double tmax = -1e300, tmin = 1e300, rmax = 0.0, rmin = 1e300;
        bool ok = true;
        for (int k = 0; k < Q && ok; ++k) {
            double L[3] = {0.0, 0.0, 0.0};
            for (int i = 0; i < M; ++i)
                for (int s = 0; s < 3; ++s) L[s] += lam[i * 3 + s] * phi[k * M + i];
            if (!(L[2] < 0.0)) {
                ok = false;
                break;
            }
            tmax = fmax(tmax, L[2]);
            tmin = fmin(tmin, L[2]);
            double rr = synth_rho(L[0], L[1], L[2], g);
            rmax = fmax(rmax, rr);
            rmin = fmin(rmin, rr);
        }
        if (!ok || !(tmax / tmin >= p.temp_ratio) || !(rmax / rmin <= p.rho_ratio)) continue;
        // moments u = sum_k U(L_k) phiw(k, :)
        for (int e = 0; e < 3 * M; ++e) u[e] = 0.0;
        for (int k = 0; k < Q; ++k) {
            double L[3] = {0.0, 0.0, 0.0};
            for (int i = 0; i < M; ++i)
                for (int s = 0; s < 3; ++s) L[s] += lam[i * 3 + s] * phi[k * M + i];
            double r0 = synth_rho(L[0], L[1], L[2], g);
            double U[3] = {r0, -(L[1] / L[2]) * r0, ((L[1] * L[1] - 2.0 * L[2]) / (2.0 * L[2] * L[2])) * r0};
            for (int i = 0; i < M; ++i)
                for (int s = 0; s < 3; ++s) u[i * 3 + s] += U[s] * phiw[k * M + i];
        }
        double d0[3];
        synth_ds(u, g, d0);
        for (int e = 0; e < 3 * M; ++e) lam0[e] = 0.0;
        for (int s = 0; s < 3; ++s) lam0[s] = d0[s];
        return attempt + 1;
    }
    return p.max_attempts + 1;
}

}  // namespace uqk
