// uqk_device.h -- device-side closed forms shared by the closure and sweep
// kernels: entropy closures (U, DU), point fluxes, wave-level helpers.
// All arithmetic here is FP64.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "uqk.h"

namespace uqk {

struct ClosureParams {
    double gamma;      // positive adiabatic exponent
    double u_minus, u_plus;
    double g;
    double barrier;
    double alpha;
    double eps;
    int max_iter;
    int max_halvings;
    int du_exact;
    int safe;
};

struct ProblemParams {
    int problem;
    int flux;
    double cfl;
    double gamma;
    double g;
    double friction;
    int friction_dim;
};

template <int KIND> struct KindInfo;
template <> struct KindInfo<UQK_BOUNDED_BARRIER> { static constexpr int S = 1; };
template <> struct KindInfo<UQK_EULER1D> { static constexpr int S = 3; };
template <> struct KindInfo<UQK_EULER2D> { static constexpr int S = 4; };
template <> struct KindInfo<UQK_SHALLOW_WATER2D> { static constexpr int S = 3; };
template <> struct KindInfo<UQK_SHALLOW_WATER2D_BARRIER> { static constexpr int S = 3; };

// ---------------------------------------------------------------------------
// Entropy closures.  U(Lambda) = (grad s)^{-1}(Lambda) per quadrature point.
// ---------------------------------------------------------------------------

// Logistic map onto (u-, u+), evaluated on the branch whose exponential cannot
// overflow (boundedbarrier.cpp:19-31).
__device__ __forceinline__ double bb_u(double lam, double um, double up) {
    double ep = exp(lam);
    double en = 1.0 / ep;
    if (lam > 0.0) return up / (en + 1.0) + um * en / (1.0 + en);
    return um / (ep + 1.0) + up * ep / (1.0 + ep);
}

// d/dLambda of bb_u (boundedbarrier.cpp:67-80).
__device__ __forceinline__ double bb_du(double lam, double um, double up) {
    double ep = exp(lam);
    double en = 1.0 / ep;
    if (lam > 0.0) return en * (up - um) / (exp(-2.0 * lam) + 2.0 * en + 1.0);
    return ep * (up - um) / (1.0 + 2.0 * ep + exp(2.0 * lam));
}

// Euler 1D density from the duals (v0, v1, v2 < 0) (eulerclosure.cpp:7-12).
__device__ __forceinline__ double e1_rho(double v0, double v1, double v2, double gm) {
    return exp((2.0 * v0 * v2 - 2.0 * v2 * log(-v2) - 2.0 * v2 * gm - v1 * v1) / (2.0 * v2 * (gm - 1.0)));
}

// Vector form of U (the one Closure::Gradient uses).
template <int KIND>
__device__ __forceinline__ void closure_U(const ClosureParams& c, const double* L, double* out) {
    if constexpr (KIND == UQK_BOUNDED_BARRIER) {
        out[0] = bb_u(L[0], c.u_minus, c.u_plus);
    } else if constexpr (KIND == UQK_EULER1D) {
        double r = e1_rho(L[0], L[1], L[2], c.gamma);
        out[0] = r;
        out[1] = -(L[1] / L[2]) * r;
        out[2] = ((L[1] * L[1] - 2.0 * L[2]) / (2.0 * L[2] * L[2])) * r;
    } else if constexpr (KIND == UQK_EULER2D) {
        // eulerclosure2d.cpp:7-17 (gamma stored negated there):
        // exp(a / (1 + gm)) (-v4)^(1 / (1 + gm)) taken as one exponential,
        // exp((a + log(-v4)) / (1 + gm)): equal up to rounding, no pow
        const double gm = -c.gamma;
        double v1 = L[0], v2 = L[1], v3 = L[2], v4 = L[3];
        double e = exp(((v2 * v2 + v3 * v3 - 2.0 * v1 * v4 - 2.0 * v4 * gm) / (2.0 * v4) + log(-v4)) / (1.0 + gm));
        out[0] = e;
        out[1] = -(v2 * e / v4);
        out[2] = -(v3 * e / v4);
        out[3] = -(e * (-v2 * v2 - v3 * v3 + 2.0 * v4)) / (2.0 * v4 * v4);
    } else if constexpr (KIND == UQK_SHALLOW_WATER2D) {
        // shallowwaterclosure2d.cpp:10-14
        double c0 = 2.0 * L[0] + L[1] * L[1] + L[2] * L[2];
        out[0] = 0.5 * c0 / c.g;
        out[1] = 0.5 * L[1] * c0 / c.g;
        out[2] = 0.5 * (2.0 * L[0] * L[2] + L[1] * L[1] * L[2] + L[2] * L[2] * L[2]) / c.g;
    } else {
        // entropy |q|^2/(2h) + g h^2/2 - eps log h: h solves g h - eps/h = c
        double cc = L[0] + 0.5 * (L[1] * L[1] + L[2] * L[2]);
        double r = sqrt(cc * cc + 4.0 * c.g * c.barrier);
        double h = cc > 0.0 ? (cc + r) / (2.0 * c.g) : (2.0 * c.barrier) / (r - cc);
        out[0] = h;
        out[1] = h * L[1];
        out[2] = h * L[2];
    }
}

// Matrix form of U (U(const Matrix&), used for uQ).  Only Euler2D differs:
// pow(-exp(a) v4, 1/(1+gamma)) (eulerclosure2d.cpp:19-32).
template <int KIND>
__device__ __forceinline__ void closure_U_mat(const ClosureParams& c, const double* L, double* out) {
    if constexpr (KIND == UQK_EULER2D) {
        const double gm = -c.gamma;
        double v1 = L[0], v2 = L[1], v3 = L[2], v4 = L[3];
        double e = pow(-exp((v2 * v2 + v3 * v3 - 2.0 * v1 * v4 - 2.0 * v4 * gm) / (2.0 * v4)) * v4, 1.0 / (1.0 + gm));
        out[0] = e;
        out[1] = -(v2 * e / v4);
        out[2] = -(v3 * e / v4);
        out[3] = -(e * (-v2 * v2 - v3 * v3 + 2.0 * v4)) / (2.0 * v4 * v4);
    } else {
        closure_U<KIND>(c, L, out);
    }
}

// Packed lower triangle of DU: entry p(l, m) = l*(l+1)/2 + m for m <= l holds
// dU_l/dLambda_m.  The Hessian's upper triangle only reads these
// (closure.cpp:265 with dposv 'U').
template <int KIND>
__device__ __forceinline__ void closure_DU(const ClosureParams& c, const double* L, double* y) {
    if constexpr (KIND == UQK_BOUNDED_BARRIER) {
        y[0] = bb_du(L[0], c.u_minus, c.u_plus);
    } else if constexpr (KIND == UQK_EULER1D) {
        // eulerclosure.cpp:37-56; DU(2,2) is the exact derivative unless du_exact == 0
        const double gm = c.gamma;
        double v0 = L[0], v1 = L[1], v2 = L[2];
        double lg = log(-v2);
        double E = exp((2.0 * v0 * v2 - 2.0 * v2 * lg - 2.0 * v2 * gm - v1 * v1) / (2.0 * v2 * (gm - 1.0)));
        double d0 = E / (gm - 1.0);
        double d1 = -(v1 / (v2 * (gm - 1.0))) * E;
        double d2 = ((-2.0 * gm + 2.0 * v0 - 2.0 * lg - 2.0) / (2.0 * (gm - 1.0) * v2) -
                     (-2.0 * gm * v2 + 2.0 * v0 * v2 - v1 * v1 - 2.0 * v2 * lg) / (2.0 * (gm - 1.0) * v2 * v2)) *
                    E;
        double a = (v1 * v1 - 2.0 * v2) / (2.0 * v2 * v2);
        y[0] = d0;                                   // (0,0)
        y[1] = -(v1 / v2) * d0;                      // (1,0)
        y[2] = -(1.0 / v2) * E - (v1 / v2) * d1;     // (1,1)
        y[3] = a * d0;                               // (2,0)
        y[4] = a * d1 + (v1 / (v2 * v2)) * E;        // (2,1)
        y[5] = c.du_exact ? (1.0 / (v2 * v2) - v1 * v1 / (v2 * v2 * v2)) * E + a * d2
                          : (v1 * v1 * (v2 * v2 * v2) + 1.0 / (v2 * v2)) * E + a * d2;  // (2,2)
    } else if constexpr (KIND == UQK_EULER2D) {
        // eulerclosure2d.cpp:54-88
        const double gm = -c.gamma;
        double v1 = L[0], v2 = L[1], v3 = L[2], v4 = L[3];
        double v22 = v2 * v2, v32 = v3 * v3, v42 = v4 * v4;
        // the density factor as in closure_U (one exponential, no pow)
        double e = exp(((v22 + v32 - 2.0 * v1 * v4 - 2.0 * v4 * gm) / (2.0 * v4) + log(-v4)) / (1.0 + gm)) / (1.0 + gm);
        double vt = e * (v22 + v32 + 2.0 * v4 * gm) / (2.0 * v42 * v4);
        y[0] = -e;                                                  // (0,0)
        y[1] = v2 * e / v4;                                         // (1,0)
        y[2] = -(e * (v22 + (gm + 1.0) * v4)) / v42;                // (1,1)
        y[3] = v3 * e / v4;                                         // (2,0)
        y[4] = -(v2 * v3 * e) / v42;                                // (2,1)
        y[5] = -(e * (v32 + (gm + 1.0) * v4)) / v42;                // (2,2)
        y[6] = -(0.5 * (v22 + v32 - 2.0 * v4) * e) / v42;           // (3,0)
        y[7] = v2 * vt;                                             // (3,1)
        y[8] = v3 * vt;                                             // (3,2)
        y[9] = -(e * (v22 * v22 + v32 * v32 + 4.0 * v32 * v4 * gm - 4.0 * v42 * gm + 2.0 * v22 * (v32 + 2.0 * v4 * gm))) /
               (4.0 * v42 * v42);                                   // (3,3)
    } else if constexpr (KIND == UQK_SHALLOW_WATER2D) {
        // shallowwaterclosure2d.cpp:39-49
        const double g = c.g;
        y[0] = 1.0 / g;
        y[1] = L[1] / g;
        y[2] = (L[0] + 1.5 * L[1] * L[1] + 0.5 * L[2] * L[2]) / g;
        y[3] = L[2] / g;
        y[4] = L[1] * L[2] / g;
        y[5] = 0.5 * (2.0 * L[0] + L[1] * L[1] + 3.0 * L[2] * L[2]) / g;
    } else {
        // DU = h' a a^T + h diag(0,1,1), a = (1, L1, L2), h' = dh/dc = h / r
        double cc = L[0] + 0.5 * (L[1] * L[1] + L[2] * L[2]);
        double r = sqrt(cc * cc + 4.0 * c.g * c.barrier);
        double h = cc > 0.0 ? (cc + r) / (2.0 * c.g) : (2.0 * c.barrier) / (r - cc);
        double hp = h / r;
        y[0] = hp;
        y[1] = hp * L[1];
        y[2] = hp * L[1] * L[1] + h;
        y[3] = hp * L[2];
        y[4] = hp * L[1] * L[2];
        y[5] = hp * L[2] * L[2] + h;
    }
}

// ---------------------------------------------------------------------------
// Point fluxes G(u, v, nUnit, n): the outward flux of the cell holding u
// through an edge with unit normal nu and scaled normal n.
// ---------------------------------------------------------------------------
template <int S>
__device__ __forceinline__ void point_flux(const ProblemParams& p, const double* u, const double* v, double nux,
                                           double nuy, double nx, double ny, double* out) {
    if (p.problem == UQK_BURGERS) {
        if (p.flux == UQK_FLUX_REFERENCE) {
            // burgers.cpp:21-28
            double w = (u[0] * nux > 0) ? u[0] : v[0];
            out[0] = 0.5 * w * w * nx;
        } else {
            // local Lax-Friedrichs (julia/Tensorization/IPMSolver.jl:39-42)
            double beta = fmax(fabs(u[0]), fabs(v[0]));
            out[0] = 0.5 * (0.5 * u[0] * u[0] + 0.5 * v[0] * v[0]) * nx - 0.5 * beta * (v[0] - u[0]) * fabs(nx);
        }
        return;
    }
    if constexpr (S >= 3) {
        double nn = sqrt(nx * nx + ny * ny);
        double Fu[S], Fv[S];  // F . n
        double unu, unv, au, av;
        if (p.problem == UQK_EULER2D_PROBLEM) {
            if constexpr (S == 4) {
                const double gm = p.gamma;
                double ri = 1.0 / u[0];
                double uu = u[1] * ri, vu = u[2] * ri;
                double pu = (gm - 1.0) * (u[3] - 0.5 * u[0] * (uu * uu + vu * vu));
                au = sqrt(gm * pu * ri);
                double rv = 1.0 / v[0];
                double uv = v[1] * rv, vv = v[2] * rv;
                double pv = (gm - 1.0) * (v[3] - 0.5 * v[0] * (uv * uv + vv * vv));
                av = sqrt(gm * pv * rv);
                unu = nux * uu + nuy * vu;
                unv = nux * uv + nuy * vv;
                // euler2d.cpp:95-112
                Fu[0] = u[1] * nx + u[2] * ny;
                Fu[1] = (u[1] * uu + pu) * nx + u[2] * uu * ny;
                Fu[2] = u[1] * vu * nx + (u[2] * vu + pu) * ny;
                Fu[3] = (u[3] + pu) * uu * nx + (u[3] + pu) * vu * ny;
                Fv[0] = v[1] * nx + v[2] * ny;
                Fv[1] = (v[1] * uv + pv) * nx + v[2] * uv * ny;
                Fv[2] = v[1] * vv * nx + (v[2] * vv + pv) * ny;
                Fv[3] = (v[3] + pv) * uv * nx + (v[3] + pv) * vv * ny;
            }
        } else {
            if constexpr (S == 3) {
                // shallowwater2d.cpp:21-65
                const double g = p.g;
                double uu = u[1] / u[0], vu = u[2] / u[0], uv = v[1] / v[0], vv = v[2] / v[0];
                au = sqrt(g * u[0]);
                av = sqrt(g * v[0]);
                unu = nux * uu + nuy * vu;
                unv = nux * uv + nuy * vv;
                Fu[0] = u[1] * nx + u[2] * ny;
                Fu[1] = (u[1] * u[1] / u[0] + 0.5 * g * u[0] * u[0]) * nx + u[1] * u[2] / u[0] * ny;
                Fu[2] = u[1] * u[2] / u[0] * nx + (u[2] * u[2] / u[0] + 0.5 * g * u[0] * u[0]) * ny;
                Fv[0] = v[1] * nx + v[2] * ny;
                Fv[1] = (v[1] * v[1] / v[0] + 0.5 * g * v[0] * v[0]) * nx + v[1] * v[2] / v[0] * ny;
                Fv[2] = v[1] * v[2] / v[0] * nx + (v[2] * v[2] / v[0] + 0.5 * g * v[0] * v[0]) * ny;
            }
        }
        if (p.flux == UQK_FLUX_REFERENCE) {
            // HLL with Davis wave speeds (euler2d.cpp:52-87)
            double lmin = unu - au, lmax = unv + av;
            if (lmin >= 0) {
#pragma unroll
                for (int s = 0; s < S; ++s) out[s] = Fu[s];
            } else if (lmax <= 0) {
#pragma unroll
                for (int s = 0; s < S; ++s) out[s] = Fv[s];
            } else {
                double f = 1.0 / (lmax - lmin);
#pragma unroll
                for (int s = 0; s < S; ++s) out[s] = f * ((lmax * Fu[s] - lmin * Fv[s]) + lmax * lmin * (v[s] - u[s]) * nn);
            }
        } else {
            double smax = fmax(fabs(unu) + au, fabs(unv) + av);
#pragma unroll
            for (int s = 0; s < S; ++s) out[s] = 0.5 * (Fu[s] + Fv[s]) - 0.5 * smax * (v[s] - u[s]) * nn;
        }
    }
}

// Problem::ComputeDt for one quadrature point (burgers.cpp:114-122,
// euler2d.cpp:114-134; SW2D mirrors Euler2D with c = sqrt(g h)).
template <int S>
__device__ __forceinline__ double point_dt(const ProblemParams& p, const double* u, double dx) {
    if (p.problem == UQK_BURGERS) return p.cfl * dx / fabs(u[0]);
    if constexpr (S >= 3) {
        double ri = 1.0 / u[0];
        double vx = u[1] * ri, vy = u[2] * ri, a;
        if constexpr (S == 4) {
            double pr = (p.gamma - 1.0) * (u[3] - 0.5 * u[0] * (vx * vx + vy * vy));
            a = sqrt(p.gamma * pr * ri);
        } else {
            a = sqrt(p.g * u[0]);
        }
        return (p.cfl / dx) * fmin(fmin(fabs(1.0 / (vy - a)), fabs(1.0 / (vy + a))),
                                   fmin(fabs(1.0 / (vx + a)), fabs(1.0 / (vx - a))));
    }
    return 1e10;
}

// ---------------------------------------------------------------------------
// Wave helpers (wave64).
// ---------------------------------------------------------------------------
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

__device__ __forceinline__ double wave_min(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = fmin(v, __shfl_xor(v, off, 64));
    return v;
}

}  // namespace uqk
