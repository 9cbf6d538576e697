// uqk_ctx.h -- the context behind the opaque uqk_ctx handle of the C ABI:
// basis tables in HBM, closure parameters, stream, events, reduction buffers
// and the captured step graph.  Shared by the translation units that
// implement the C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "uqk.h"
#include "uqk_internal.h"

struct uqk_mlp;

struct uqk_ctx {
    int device = 0;
    int kind = 0;
    int S = 0, M = 0, Q = 0, d = 0;
    int precision = UQK_FP64;
    uqk::ClosureParams c{};
    double* d_phi = nullptr;
    double* d_phiw = nullptr;
    double* d_xi = nullptr;
    unsigned long long* d_counters = nullptr;  // [0] queue head, [1] failed cells, [2] Newton steps, [3] gradients,
                                               // [4] cells that took the learned warm start
    double* d_partial = nullptr;               // per-block partials
    double* d_scalars = nullptr;               // [0] dt, [1] residual
    hipStream_t own_stream = nullptr;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    double last_closure_ms = 0.0;
    // the last closure call's (or run's) counters 1..6: failed cells, Newton
    // steps, gradients, learned starts taken, factorisations, kept-inverse steps
    unsigned long long last_stats[6] = {0, 0, 0, 0, 0, 0};
    int kernel = UQK_KERNEL_AUTO;
    int last_kernel = UQK_KERNEL_AUTO;
    bool timed = false;  // tev0/tev1 bracket the last closure kernel
    hipEvent_t tev0 = nullptr, tev1 = nullptr;  // the pair the last closure recorded (ev0/ev1 or the step timer's)
    // per-step kernel timing (uqk_set_kernel_timing): four pool events per
    // step enqueued since the last collection (closure pair, advance pair)
    bool ktiming = false;
    std::vector<hipEvent_t> kev;
    size_t kev_used = 0;
    hipEvent_t cev0 = nullptr, cev1 = nullptr;  // closure pair of the step being enqueued
    double k_closure_ms = 0.0, k_advance_ms = 0.0;
    long long k_steps = 0;
    bool capturing = false;  // enqueueing into a graph: no events, no syncs
    uqk::RunState* d_run = nullptr;
    uqk::RunState* h_run = nullptr;  // pinned host copy of *d_run, read back by the replay loop
    hipGraphExec_t graph_exec = nullptr;
    std::string graph_key;
    size_t max_lds = 0;
    int num_cu = 0;
    uqk::LaunchShape shape;   // specialised closure kernel launch shape (autotuned or set)
    uqk_mlp* warm = nullptr;  // learned warm start run before each closure (not owned)
    uqk::RefineArgs refine{};  // adaptive levels: refine.nlev > 1 turns refinement on in the step
    unsigned char* scratch = nullptr;  // device address of the library's static staging array on `device`
    // reuse of the factorisation in the mixed wave kernel across time
    // steps: each cell's kept inverse Hessian (FP32, n x n, row stride n + 1),
    // its last factor (FP64 packed U, 1/U_rr) and state (0 none, 1 inverse
    // kept, 2 factor stored, inverse pending)
    bool factor_reuse = true;
    float* pinv_cell = nullptr;
    double* ufac_cell = nullptr;
    int* pinv_ok = nullptr;
    int* pinv_list = nullptr;  // cells with a new factor, per launch
    // what the buffers hold: 1 the wave kernel's kept inverses (with ufac
    // and pinv_list), 2 the workgroup kernel's kept FP32 factors; 0 none
    int factor_mode = 0;
    // residual check before the wave kernel in the time loop (on by
    // default): the cells it leaves to the Newton kernel
    bool residual_check = true;
    int* check_list = nullptr;
    long check_cap = 0;
    // and their gradients at lambda (n x S x M), and every cell's ComputeDt
    // (the steps of a run after the first)
    double* check_grad = nullptr;
    double* check_dt = nullptr;
    long check_grad_cap = 0;
    // the last launch_closure left uQ, the moments and each cell's ComputeDt
    // at lambda for every cell (a run's step then sets RunState::recon)
    bool step_recon = false;
    // within uqk_run_steps: 1 the step's residual check evaluates U, 2 it
    // takes the gradient from the last step's outputs; 0 outside a run
    int check_mode = 0;
    long pinv_cell_cap = 0;
    std::string err;
};

namespace uqk {
// Drops the captured step graph.  Every setter that changes what the step
// enqueues (kernel choice, launch shape, warm start, refinement, stream)
// calls it, so the next uqk_run_steps captures afresh instead of replaying
// launches that name a detached network, freed buffers or an old shape.
inline void invalidate_graph(uqk_ctx* ctx) {
    if (ctx->graph_exec) {
        (void)hipStreamSynchronize(ctx->stream);
        (void)hipGraphExecDestroy(ctx->graph_exec);
    }
    ctx->graph_exec = nullptr;
    ctx->graph_key.clear();
}
// Grows the attached network's buffers for a warm start over ncell cells
// (outside graph capture, which cannot allocate).
int prepare_warm_start(uqk_ctx* ctx, long ncell);
// learned warm start (mlp.hip): the attached network's starting duals for the
// moments u into its own buffer (*lam_alt; null without a network); enqueue
// only.  nt: per-cell orders (moments past a cell's order read as zero, its
// duals there are zero), or null.
int enqueue_warm_predict(uqk_ctx* ctx, long ncell, const double* u, const double** lam_alt, const int* nt);
// per cell, the start with the smaller ||grad L|| (over the cell's nt[c]
// moments when nt is given) into lambda
int enqueue_warm_select(uqk_ctx* ctx, long ncell, const double* u, double* lambda, const double* lambda_alt,
                        const int* active, const int* nt);
}  // namespace uqk
