"""Roofline models of the closure and advance kernels on one MI355X.

Algorithmic flops of the Newton closure per step kind, the time-weighted peak
of the units they run on, and the algorithmic HBM bytes of the fused advance.
Shared by bench.py and the configuration runner (uqcreator_amd.configs).
"""
import numpy as np

# dense peaks of one MI355X
FP64_PEAK_TFLOPS = 78.6
FP32_PEAK_TFLOPS = 157.3
HBM_PEAK_GBS = 8000.0


def flop_model(S, M, Q):
    """Algorithmic flops, n = S M duals:
    F_H    the symmetric Hessian of one factored Newton step, Q n (n+1);
    F_C    its Cholesky and solve, n^3/3 + 2 n^2;
    F_inv  the inverse the time loop keeps after a factorisation, n^3/3;
    F_lag  one time-loop step on a kept inverse: two applications of the
           inverse (4 n^2) and the FP64 Hessian-vector product
           (4 Q S M + 2 Q S^2);
    F_tri  one step of a single closure call on its kept factor: the two
           triangular solves, 2 n^2;
    F_g    one gradient evaluation, 4 Q S M."""
    n = S * M
    return dict(F_H=Q * n * (n + 1), F_C=n ** 3 / 3.0 + 2.0 * n * n, F_inv=n ** 3 / 3.0,
                F_lag=4.0 * n * n + 4.0 * Q * S * M + 2.0 * Q * S * S, F_tri=2.0 * n * n, F_g=4.0 * Q * S * M)


def advance_bytes_per_cell(S, M, Q, edges, dim):
    """Algorithmic HBM bytes of the fused advance per cell: uQ of the cell
    (the neighbours' uQ are the same array, read again through L2), the
    reconstructed moments, the new moments, and the cell's geometry
    (neighbour ids, unit and scaled normals, area, boundary type and normal)."""
    geometry = edges * 4 + 2 * edges * dim * 8 + 8 + 4 + dim * 8
    return 8 * S * Q + 8 * S * M + 8 * S * M + geometry


def closure_roofline(kernel, kernel_ms, newton, grads, fm, precision_mixed, factors=None, lagged=0.0,
                     reuse="inverse"):
    """Roofline of a closure launch: algorithmic flops over its HIP-event
    time; the peak is time-weighted over the units the flops run on (Hessian
    on FP64 or FP32 MFMA, the rest FP64).  factors / lagged: factored Newton
    steps and steps on a kept factorisation (mixed closure; by default every
    Newton step is factored).  reuse: "inverse" (the time loop keeps an
    inverse per factorisation -- the wave kernel's H^-1, the workgroup
    kernel's U^-1 formed at its first use -- F_inv + F_lag) or "factor" (a
    single call solves with its last factor, F_tri)."""
    if factors is None:
        factors = newton
    f_hess = factors * fm["F_H"]
    f_fp64 = factors * fm["F_C"] + grads * fm["F_g"]
    if lagged:
        if reuse == "inverse":
            f_fp64 += factors * fm["F_inv"] + lagged * fm["F_lag"]
        else:
            f_fp64 += lagged * fm["F_tri"]
    flops = f_hess + f_fp64
    hess_peak = FP32_PEAK_TFLOPS if precision_mixed else FP64_PEAK_TFLOPS
    t_min = f_hess / (hess_peak * 1e12) + f_fp64 / (FP64_PEAK_TFLOPS * 1e12)
    achieved = flops / (kernel_ms * 1e-3) / 1e12 if kernel_ms > 0 else 0.0
    peak = flops / t_min / 1e12 if t_min > 0 else FP64_PEAK_TFLOPS
    return {"bound": "mfma", "achieved": achieved, "peak": peak, "unit": "TFLOP/s", "frac": achieved / peak,
            "traffic": None, "kernel": kernel, "kernel_ms": kernel_ms, "flops_per_launch": flops,
            "hessian_flops_per_launch": f_hess, "newton_steps_per_launch": newton,
            "factorizations_per_launch": factors, "lagged_steps_per_launch": lagged,
            "gradient_evals_per_launch": grads, "flop_model": fm}


def advance_roofline(kernel, kernel_ms, S, M, Q, edges, dim, n_cells):
    """The advance against HBM: algorithmic bytes per launch over its time."""
    bpc = advance_bytes_per_cell(S, M, Q, edges, dim)
    gbs = bpc * n_cells / (kernel_ms * 1e-3) / 1e9 if kernel_ms > 0 else 0.0
    return {"bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": gbs / HBM_PEAK_GBS,
            "traffic": None, "kernel": kernel, "kernel_ms": kernel_ms, "bytes_per_launch": bpc * n_cells,
            "bytes_per_cell": bpc}


def rel_l2(a, b, area=None):
    """Area-weighted relative L2 per field (momentsolver.cpp:687-749), the
    largest over the fields; a field whose norm is below 1e-10 of the largest
    is measured against that floor."""
    w = np.ones(a.shape[0]) if area is None else area
    num = ((a - b) ** 2 * w[:, None]).sum(0)
    den = (b ** 2 * w[:, None]).sum(0)
    return float(np.sqrt(num / np.maximum(den, 1e-20 * den.max())).max())
