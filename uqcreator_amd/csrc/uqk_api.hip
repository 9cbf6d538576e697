// uqk_api.hip -- the C ABI declared in include/uqk.h.
//
// A context is the GPU-side Closure object: it holds the basis tables in HBM,
// the closure parameters (Closure::Create, closure.cpp:94-131), a stream and
// the small device buffers the kernels reduce into.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>

#include <mutex>

#include "uqk.h"
#include "uqk_internal.h"
#include "uqk_ctx.h"

// Device staging of uqk_closure_solve_host_noalloc: one call's closure
// counters, moments, duals, per-cell nTotal, statuses and iteration counts.
// Every GPU the library runs on holds its own copy of this array from the
// code object's load until the process ends.  Host code never names it where
// a device address is needed (on the host the name is only a stand-in, and a
// kernel given it faults): uqk_create resolves the device address once per
// device with hipGetSymbolAddress, and every use goes through that address.
// Calls that use it on one GPU at the same time race (one array per GPU).
constexpr size_t kScratchBytes = size_t(64) << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

constexpr int kScratchDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kScratchDevices] = {};  // device address of `scratch`, per device ordinal

int fail(uqk_ctx* ctx, int code, const std::string& msg) {
    if (ctx) ctx->err = msg;
    return code;
}

#define UQK_HIP(ctx, call)                                                                        \
    do {                                                                                          \
        hipError_t e_ = (call);                                                                   \
        if (e_ != hipSuccess) return fail(ctx, -2, std::string(#call ": ") + hipGetErrorString(e_)); \
    } while (0)

uqk::SweepArgs sweep_args(const uqk_ctx* ctx) {
    uqk::SweepArgs a{};
    a.c = ctx->c;
    a.S = ctx->S;
    a.M = ctx->M;
    a.Q = ctx->Q;
    a.d = ctx->d;
    a.phi = ctx->d_phi;
    a.phiw = ctx->d_phiw;
    a.xi = ctx->d_xi;
    a.partial = ctx->d_partial;
    a.npartial = uqk::sweep_partial_blocks(0);
    return a;
}

uqk::ProblemParams problem_params(const uqk_problem* p) {
    uqk::ProblemParams q{};
    q.problem = p->problem;
    q.flux = p->flux;
    q.cfl = p->cfl;
    q.gamma = p->gamma;
    q.g = p->gravity;
    q.friction = p->friction;
    q.friction_dim = p->friction_dim;
    return q;
}

void set_mesh(uqk::SweepArgs& a, const uqk_mesh* m) {
    a.nedge = m->n_edges;
    a.dim = m->dim;
    a.nbr = m->neighbor;
    a.unit_normal = m->unit_normal;
    a.normal = m->normal;
    a.area = m->area;
    a.dx_param = m->dx_param;
    a.bc = m->bc;
    a.bnormal = m->boundary_normal;
}

int expected_states(int kind) {
    switch (kind) {
        case UQK_BOUNDED_BARRIER: return 1;
        case UQK_EULER1D: return 3;
        case UQK_EULER2D: return 4;
        case UQK_SHALLOW_WATER2D: return 3;
        case UQK_SHALLOW_WATER2D_BARRIER: return 3;
    }
    return -1;
}

// The sweeps handle lines (2 edges, dim 1) and quadrilaterals (4 edges, dim 2).
int check_mesh(uqk_ctx* ctx, const uqk_mesh* m) {
    if (m->n_cells < 0 || m->n_edges < 1 || m->n_edges > 4 || m->dim < 1 || m->dim > 2)
        return fail(ctx, -1, "mesh: n_edges must be 1..4 and dim 1 or 2");
    return 0;
}

// Kept inverse factors of the mixed wave kernel across time steps: grows
// the per-cell slots (flags cleared) to n cells.  Allocates only outside
// graph capture.
// Kept factorisations of the time loop's mixed closure, by the kernel the
// launch takes: 1 the wave kernel's kept inverses, 2 the workgroup kernel's
// kept FP32 factors, 0 none.
int factor_mode(const uqk_ctx* ctx, bool per_cell_orders) {
    if (ctx->precision != UQK_MIXED || !ctx->factor_reuse || ctx->c.max_iter == 1) return 0;
    int wm = 0, wo = 0;
    if ((ctx->kernel == UQK_KERNEL_AUTO || ctx->kernel == UQK_KERNEL_WAVE) && !per_cell_orders &&
        uqk::newton_wave_limits(ctx->kind, ctx->M, ctx->Q, ctx->precision, (int)ctx->max_lds, 1, &wm, &wo))
        return uqk::newton_wave_reuse_fits(ctx->kind, ctx->M, ctx->Q, (int)ctx->max_lds) ? 1 : 0;
    if (ctx->kernel == UQK_KERNEL_WAVE) return 0;
    const bool block_ok = uqk::newton_block_fits(ctx->S, ctx->M, ctx->Q, (int)ctx->max_lds);
    if (block_ok && (ctx->kernel == UQK_KERNEL_BLOCK || (ctx->kernel == UQK_KERNEL_AUTO && ctx->S * ctx->M >= 64)))
        return 2;
    return 0;
}

// Per-cell buffers of factor mode `mode` for n_cells cells, outside graph
// capture only; a change of mode starts with no kept factor.
int ensure_factor_buffers(uqk_ctx* ctx, long n_cells, int mode) {
    const long n = (long)ctx->S * ctx->M;
    // mode 1: n x n FP32 inverse (row stride n + 1) and the FP64 factor
    // (packed U, 1/U_rr); mode 2: the FP32 factor (packed U)
    const long nn = mode == 1 ? n * (n + 1) : n * (n + 1) / 2;
    const long nu = n * (n + 1) / 2 + n;
    if (ctx->pinv_cell_cap >= n_cells && ctx->factor_mode == mode) return 0;
    if (ctx->capturing) return fail(ctx, -1, "kept factors: buffers must exist before graph capture");
    hipFree(ctx->pinv_cell);
    hipFree(ctx->ufac_cell);
    hipFree(ctx->pinv_ok);
    hipFree(ctx->pinv_list);
    ctx->pinv_cell = nullptr;
    ctx->ufac_cell = nullptr;
    ctx->pinv_ok = nullptr;
    ctx->pinv_list = nullptr;
    ctx->pinv_cell_cap = 0;
    ctx->factor_mode = 0;
    UQK_HIP(ctx, hipMalloc(&ctx->pinv_cell, sizeof(float) * (size_t)nn * n_cells));
    if (mode == 1) {
        UQK_HIP(ctx, hipMalloc(&ctx->ufac_cell, sizeof(double) * (size_t)nu * n_cells));
        UQK_HIP(ctx, hipMalloc(&ctx->pinv_list, sizeof(int) * (size_t)n_cells));
    }
    UQK_HIP(ctx, hipMalloc(&ctx->pinv_ok, sizeof(int) * (size_t)n_cells));
    UQK_HIP(ctx, hipMemsetAsync(ctx->pinv_ok, 0, sizeof(int) * (size_t)n_cells, ctx->stream));
    ctx->pinv_cell_cap = n_cells;
    ctx->factor_mode = mode;
    return 0;
}

// The list of cells the residual check leaves to the Newton kernel and their
// gradients at lambda: grown to n cells, outside graph capture only.
int ensure_check_list(uqk_ctx* ctx, long n_cells) {
    const size_t cells = (size_t)(n_cells > 0 ? n_cells : 1);
    if (ctx->check_cap < n_cells) {
        if (ctx->capturing) return fail(ctx, -1, "residual check: the cell list must exist before graph capture");
        hipFree(ctx->check_list);
        ctx->check_list = nullptr;
        ctx->check_cap = 0;
        UQK_HIP(ctx, hipMalloc(&ctx->check_list, sizeof(int) * cells));
        ctx->check_cap = n_cells;
    }
    if (ctx->check_grad_cap < n_cells) {
        if (ctx->capturing) return fail(ctx, -1, "residual check: the gradients must exist before graph capture");
        hipFree(ctx->check_grad);
        ctx->check_grad = nullptr;
        ctx->check_grad_cap = 0;
        UQK_HIP(ctx, hipMalloc(&ctx->check_grad, sizeof(double) * (size_t)ctx->S * ctx->M * cells));
        hipFree(ctx->check_dt);
        ctx->check_dt = nullptr;
        UQK_HIP(ctx, hipMalloc(&ctx->check_dt, sizeof(double) * cells));
        ctx->check_grad_cap = n_cells;
    }
    return 0;
}

// Whether the time loop's step runs the residual check before the wave
// kernel: a wave kernel shape, no per-cell orders (the wave kernel takes
// none), not SolveClosureSafe (which always takes a step).
bool wave_check(const uqk_ctx* ctx, bool per_cell_orders) {
    int wm = 0, wo = 0;
    return (ctx->kernel == UQK_KERNEL_AUTO || ctx->kernel == UQK_KERNEL_WAVE) && ctx->residual_check &&
           !ctx->c.safe && !per_cell_orders &&
           uqk::newton_wave_limits(ctx->kind, ctx->M, ctx->Q, ctx->precision, (int)ctx->max_lds, 1, &wm, &wo);
}

// Launch the Newton closure kernel; leaves the failed-cell count in counters[1]
// (counters: the context's d_counters unless the caller brings its own
// kClosureCounters).
// When uq_out/u_out are given and the specialised or the workgroup kernel runs
// (not one-shot), it also writes uQ and the reconstructed moments and *fused
// is set.
// With dt_out as well (the step), the fused kernels also reduce
// Problem::ComputeDt over the cells into *dt_out and *fused_dt is set.
int launch_closure(uqk_ctx* ctx, int64_t n, const int32_t* nt, const double* u, double* lam, int32_t* status,
                   int32_t* iters, const int* active = nullptr, double* uq_out = nullptr, double* u_out = nullptr,
                   bool* fused = nullptr, double* dt_out = nullptr, const uqk::ProblemParams* dtp = nullptr,
                   const double* dx_param = nullptr, bool* fused_dt = nullptr,
                   unsigned long long* counters = nullptr, bool prezeroed = false, bool factors_per_cell = false,
                   const int* recon_ok = nullptr) {
    if (fused) *fused = false;
    if (fused_dt) *fused_dt = false;
    ctx->timed = false;
    ctx->step_recon = false;
    if (!counters) counters = ctx->d_counters;
    // (in a replayed run the previous step's step_end_kernel already cleared
    // the counters and set dt's start value: prezeroed)
    // [0] queue head, [1] failed cells, [2] Newton steps, [3] gradients, [4] cells that took the learned start
    if (!prezeroed) UQK_HIP(ctx, uqk::launch_fill_u64(counters, uqk::kClosureCounters, ctx->stream));
    if (n <= 0) return 0;
    // learned warm start: the network's prediction (per-cell orders: masked
    // to each cell's moments); the wave kernel selects the start per cell
    // itself, the other kernels after a selection pass
    const double* lam_alt = nullptr;
    if (ctx->warm)
        if (int rc = uqk::enqueue_warm_predict(ctx, n, u, &lam_alt, nt)) return rc;
    uqk::NewtonArgs a{};
    a.c = ctx->c;
    a.M = ctx->M;
    a.Q = ctx->Q;
    a.phi = ctx->d_phi;
    a.phiw = ctx->d_phiw;
    a.ncell = n;
    a.nt_cell = nt;
    a.u = u;
    a.lam = lam;
    a.status = status;
    a.iters = iters;
    a.counter = counters;
    a.nfail = counters + 1;
    a.prec = ctx->precision;
    a.active = active;
    a.took = counters + 4;
    a.nfactor = counters + 5;
    // mixed precision: kept factorisations (per cell in the time loop: the
    // wave kernel's inverses, computed after the launch that factored, or the
    // workgroup kernel's FP32 factors; per wave for a single closure call);
    // the buffers exist before a graph is captured
    const int fmode = factors_per_cell ? factor_mode(ctx, nt != nullptr) : 0;
    if (fmode) {
        if (int rc = ensure_factor_buffers(ctx, n, fmode)) return rc;
        a.pinv_mode = 1;
        a.pinv = ctx->pinv_cell;
        a.ufac = ctx->ufac_cell;
        a.pinv_ok = ctx->pinv_ok;
        a.plist = ctx->pinv_list;
        a.plist_count = counters + 7;
        a.pinv_slots = ctx->pinv_cell_cap;
    } else if (ctx->precision == UQK_MIXED && ctx->factor_reuse && !nt && ctx->c.max_iter != 1 && !factors_per_cell &&
               ctx->c.alpha < 1.0) {
        // (wave kernel) reuse within the call, for the damped Newton only: its
        // steps contract linearly, by about 1 - alpha, with a kept factor as
        // with a fresh one; the undamped Newton would lose its quadratic rate
        a.pinv_mode = 2;
    }
    const bool want_dt = dt_out && dtp && dx_param && uq_out && u_out && ctx->c.max_iter != 1;
    if (want_dt) {
        // the start value of ComputeDt's MIN (momentsolver.cpp:127); a path
        // that does not fuse leaves dt to the dt sweep, which overwrites it
        if (!prezeroed) UQK_HIP(ctx, uqk::launch_fill_f64(dt_out, 1, 1e10, ctx->stream));
        a.dtp = *dtp;
        a.dx_param = dx_param;
        a.dt_min = reinterpret_cast<unsigned long long*>(dt_out);
    }
    const bool ev = !ctx->capturing;
    // the step timer brings its own pair (uqk_set_kernel_timing), else ev0/ev1
    ctx->tev0 = ctx->cev0 ? ctx->cev0 : ctx->ev0;
    ctx->tev1 = ctx->cev1 ? ctx->cev1 : ctx->ev1;
    if (ctx->kernel == UQK_KERNEL_AUTO || ctx->kernel == UQK_KERNEL_WAVE) {
        hipError_t e = hipSuccess;
        const bool fuse = uq_out && u_out && ctx->c.max_iter != 1;
        a.uq_out = fuse ? uq_out : nullptr;
        a.u_out = fuse ? u_out : nullptr;
        a.lam_alt = lam_alt;
        if (ev) UQK_HIP(ctx, hipEventRecord(ctx->tev0, ctx->stream));
        // the time loop's step: the residual check finishes the cells whose
        // lambda already passes the convergence test and lists the rest, with
        // their gradients, for the wave kernel
        if (fuse && wave_check(ctx, nt != nullptr)) {
            if (int rc = ensure_check_list(ctx, n)) return rc;
            uqk::NewtonArgs ca = a;
            ca.list_out = ctx->check_list;
            ca.list_count = counters + 8;
            ca.grad_at_lam = ctx->check_grad;
            // in a run every cell's ComputeDt is kept, and the steps after the
            // first take the gradient from the last step's outputs
            // (closure_recheck_kernel)
            const bool keep = want_dt && recon_ok;
            ca.recon_ok = keep && ctx->check_mode == 2 ? recon_ok : nullptr;
            ca.dt_cell = keep ? ctx->check_dt : nullptr;
            if (uqk::closure_check_launch(ctx->kind, ctx->M, ctx->Q, ca, ctx->num_cu, ctx->stream, &e)) {
                UQK_HIP(ctx, e);
                a.cell_list = ctx->check_list;
                a.cell_count = counters + 8;
                a.grad_at_lam = ctx->check_grad;
                a.dt_cell = ca.dt_cell;
            }
        }
        if (uqk::newton_wave_launch(ctx->kind, ctx->M, ctx->Q, a, (int)ctx->max_lds, ctx->num_cu, ctx->stream, &e,
                                    ctx->shape)) {
            UQK_HIP(ctx, e);
            if (a.pinv_mode == 1) {
                // the inverses of this launch's new factors, for the next steps
                if (!uqk::factor_inverse_launch(ctx->kind, ctx->M, ctx->Q, n, counters + 7, ctx->pinv_list,
                                                ctx->pinv_ok, ctx->ufac_cell, ctx->pinv_cell, active, ctx->num_cu,
                                                ctx->stream, &e))
                    return fail(ctx, -3, "no factor inverse kernel for this shape");
                UQK_HIP(ctx, e);
            }
            if (ev) UQK_HIP(ctx, hipEventRecord(ctx->tev1, ctx->stream));
            // every cell's uQ, moments and ComputeDt are now those of its lambda
            ctx->step_recon = a.dt_cell != nullptr;
            ctx->last_kernel = UQK_KERNEL_WAVE;
            ctx->timed = ev;
            if (fused) *fused = fuse;
            if (fused_dt) *fused_dt = want_dt;
            return 0;
        }
        if (ctx->kernel == UQK_KERNEL_WAVE) return fail(ctx, -3, "no specialised closure kernel for this shape");
        a.lam_alt = nullptr;
        a.cell_list = nullptr;
        a.cell_count = nullptr;
        a.grad_at_lam = nullptr;
        a.dt_cell = nullptr;
    }
    const bool block_ok = uqk::newton_block_fits(ctx->S, ctx->M, ctx->Q, (int)ctx->max_lds);
    if (ctx->kernel == UQK_KERNEL_BLOCK || (ctx->kernel == UQK_KERNEL_AUTO && ctx->S * ctx->M >= 64 && block_ok)) {
        if (!block_ok) return fail(ctx, -3, "closure too large for the workgroup kernel");
        if (a.pinv_mode == 1 && ctx->factor_mode != 2) a.pinv_mode = 0;  // (the slots hold inverses)
        const bool fuse = uq_out && u_out && ctx->c.max_iter != 1;
        a.uq_out = fuse ? uq_out : nullptr;
        a.u_out = fuse ? u_out : nullptr;
        const bool check = fuse && ctx->residual_check && !ctx->c.safe &&
                           uqk::closure_check_wave_fits(ctx->S, ctx->M, ctx->Q, (int)ctx->max_lds);
        // learned start: the residual check chooses it for the cells that
        // fail at lambda; without the check a pass over every cell does
        if (lam_alt && !check)
            if (int rc = uqk::enqueue_warm_select(ctx, n, u, lam, lam_alt, active, nt)) return rc;
        if (ev) UQK_HIP(ctx, hipEventRecord(ctx->tev0, ctx->stream));
        // the time loop's step: the residual check (one wave per cell) first
        if (check) {
            if (int rc = ensure_check_list(ctx, n)) return rc;
            uqk::NewtonArgs ca = a;
            ca.lam_alt = lam_alt;
            ca.list_out = ctx->check_list;
            ca.list_count = counters + 8;
            ca.grad_at_lam = ctx->check_grad;
            // in a run every cell's ComputeDt is kept, and the steps after the
            // first take the gradient from the last step's outputs
            const bool keep = want_dt && recon_ok;
            ca.recon_ok = keep && ctx->check_mode == 2 ? recon_ok : nullptr;
            ca.dt_cell = keep ? ctx->check_dt : nullptr;
            UQK_HIP(ctx, uqk::closure_check_wave_launch(ctx->kind, ca, (int)ctx->max_lds, ctx->num_cu, ctx->stream));
            a.cell_list = ctx->check_list;
            a.cell_count = counters + 8;
            a.grad_at_lam = ctx->check_grad;
            a.dt_cell = ca.dt_cell;
        }
        UQK_HIP(ctx, uqk::newton_block_launch(ctx->kind, a, (int)ctx->max_lds, ctx->num_cu, ctx->stream, ctx->shape));
        if (ev) UQK_HIP(ctx, hipEventRecord(ctx->tev1, ctx->stream));
        ctx->step_recon = a.dt_cell != nullptr;
        ctx->last_kernel = UQK_KERNEL_BLOCK;
        ctx->timed = ev;
        if (fused) *fused = fuse;
        if (fused_dt) *fused_dt = want_dt;
        return 0;
    }
    if (lam_alt)
        if (int rc = uqk::enqueue_warm_select(ctx, n, u, lam, lam_alt, active, nt)) return rc;
    ctx->last_kernel = UQK_KERNEL_GENERIC;
    a.uq_out = nullptr;
    a.u_out = nullptr;
    a.dt_min = nullptr;
    a.wave_doubles = uqk::newton_fp64_wave_doubles(ctx->S, ctx->M, ctx->Q);
    const size_t wave_bytes = sizeof(double) * (size_t)a.wave_doubles;
    const size_t table_bytes = sizeof(double) * 2 * (size_t)ctx->Q * ctx->M;
    bool tlds = false;
    int waves = 0;
    for (int w : {4, 2, 1}) {
        if (table_bytes + w * wave_bytes <= ctx->max_lds) {
            tlds = true;
            waves = w;
            break;
        }
    }
    if (!tlds) {
        for (int w : {4, 2, 1}) {
            if (w * wave_bytes <= ctx->max_lds) {
                waves = w;
                break;
            }
        }
    }
    if (waves == 0)
        return fail(ctx, -3, "closure too large for LDS: " + std::to_string(wave_bytes) + " bytes per cell");
    const int threads = 64 * waves;
    const size_t shmem = (tlds ? table_bytes : 0) + waves * wave_bytes;
    int occ = uqk::newton_fp64_occupancy(ctx->kind, tlds, threads, shmem);
    if (occ <= 0) occ = 1;
    long want = (n + waves - 1) / waves;
    int blocks = (int)std::min<long>(want, (long)occ * ctx->num_cu);
    if (ev) UQK_HIP(ctx, hipEventRecord(ctx->tev0, ctx->stream));
    UQK_HIP(ctx, uqk::newton_fp64_launch(ctx->kind, a, tlds, blocks, threads, shmem, ctx->stream));
    if (ev) UQK_HIP(ctx, hipEventRecord(ctx->tev1, ctx->stream));
    ctx->timed = ev;
    return 0;
}

// Enqueue one step body (momentsolver.cpp:83-152) on ctx->stream without any
// host synchronisation: closure, uQ, reconstruction, dt, advance, residual.
// in_run: a step of uqk_run_steps, whose step_end_kernel clears the counters,
// sets dt's start value and sums the residual (no fills, no residual kernel).
// With refinement on (ctx->refine.nlev > 1) the levels move between uQ and the
// reconstruction, as in momentsolver.cpp:99-124: the closure and uQ use the
// old nTotal, the reconstruction and the advance the new one.
int enqueue_step(uqk_ctx* ctx, const uqk_problem* prob, const uqk_mesh* mesh, int32_t* nt, double* u,
                 double* lam, double* uq, double* u_scratch, int32_t* status, int32_t* iters, const int* active,
                 bool in_run = false) {
    const int64_t n = mesh->n_cells;
    const bool refining = ctx->refine.nlev > 1;
    if (refining && !nt) return fail(ctx, -1, "refinement levels need ntotal_cell");
    bool fused = false, fused_dt = false;
    const uqk::ProblemParams pp = problem_params(prob);
    // per-step kernel timing: four pool events (closure and advance pairs)
    const bool kt = ctx->ktiming && !ctx->capturing && n > 0;
    hipEvent_t ea0 = nullptr, ea1 = nullptr;
    if (kt) {
        while (ctx->kev.size() < ctx->kev_used + 4) {
            hipEvent_t e = nullptr;
            UQK_HIP(ctx, hipEventCreate(&e));
            ctx->kev.push_back(e);
        }
        ctx->cev0 = ctx->kev[ctx->kev_used];
        ctx->cev1 = ctx->kev[ctx->kev_used + 1];
        ea0 = ctx->kev[ctx->kev_used + 2];
        ea1 = ctx->kev[ctx->kev_used + 3];
    }
    int rc = launch_closure(ctx, n, nt, u, lam, status, iters, active, refining ? nullptr : uq,
                            refining ? nullptr : u_scratch, &fused, ctx->d_scalars, &pp, mesh->dx_param, &fused_dt,
                            nullptr, in_run, true, in_run ? &ctx->d_run->recon : nullptr);
    ctx->cev0 = ctx->cev1 = nullptr;
    if (rc) return rc;
    uqk::SweepArgs a = sweep_args(ctx);
    a.p = pp;
    set_mesh(a, mesh);
    a.ncell = n;
    a.nt_cell = nt;
    a.lam = lam;
    a.uq_out = uq;
    a.active = active;
    if (!fused) UQK_HIP(ctx, uqk::launch_eval_uq(ctx->kind, a, ctx->stream));
    if (refining) {
        uqk::RefineArgs r = ctx->refine;
        r.ncell = n;
        r.u = u;
        r.nt_cell = nt;
        r.active = active;
        UQK_HIP(ctx, uqk::launch_refine(r, ctx->stream));
    }
    a.uq = uq;
    a.u_out = u_scratch;
    if (!fused) UQK_HIP(ctx, uqk::launch_reconstruct(a, ctx->stream));
    a.scalar_out = ctx->d_scalars;
    if (!fused_dt) UQK_HIP(ctx, uqk::launch_dt(ctx->kind, a, ctx->stream));
    a.u = u_scratch;
    a.u_out = u;
    a.dt_dev = ctx->d_scalars;
    if (kt) UQK_HIP(ctx, hipEventRecord(ea0, ctx->stream));
    UQK_HIP(ctx, uqk::launch_advance(ctx->kind, a, ctx->stream));
    if (kt) {
        UQK_HIP(ctx, hipEventRecord(ea1, ctx->stream));
        ctx->kev_used += 4;
    }
    if (!in_run) {
        a.scalar_out = ctx->d_scalars + 1;
        UQK_HIP(ctx, uqk::launch_residual(a, ctx->stream));
    }
    return 0;
}

// Adds the kernel times of the steps enqueued since the last call (their
// events have completed: the stream was synchronised) to the step timer.
int collect_step_times(uqk_ctx* ctx) {
    for (size_t i = 0; i + 4 <= ctx->kev_used; i += 4) {
        float c = 0.f, a = 0.f;
        UQK_HIP(ctx, hipEventElapsedTime(&c, ctx->kev[i], ctx->kev[i + 1]));
        UQK_HIP(ctx, hipEventElapsedTime(&a, ctx->kev[i + 2], ctx->kev[i + 3]));
        ctx->k_closure_ms += c;
        ctx->k_advance_ms += a;
        ++ctx->k_steps;
    }
    ctx->kev_used = 0;
    return 0;
}

int64_t finish_closure(uqk_ctx* ctx, const unsigned long long* counters = nullptr) {
    if (!counters) counters = ctx->d_counters;
    UQK_HIP(ctx, hipMemcpyAsync(ctx->last_stats, counters + 1, sizeof(ctx->last_stats), hipMemcpyDeviceToHost,
                                ctx->stream));  // counters 1..6
    UQK_HIP(ctx, hipStreamSynchronize(ctx->stream));
    float ms = 0.f;
    ctx->last_closure_ms = 0.0;
    if (ctx->timed) UQK_HIP(ctx, hipEventElapsedTime(&ms, ctx->tev0, ctx->tev1));
    if (ctx->timed) ctx->last_closure_ms = ms;
    return (int64_t)ctx->last_stats[0];
}

// Device address of `scratch` on ctx's device (current): looked up once per
// device and kept in g_scratch_ptr.
int resolve_scratch(uqk_ctx* ctx) {
    if (ctx->device < 0 || ctx->device >= kScratchDevices)
        return fail(ctx, -1, "device ordinal outside 0.." + std::to_string(kScratchDevices - 1));
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[ctx->device]) {
        void* p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(ctx, -2, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_ptr[ctx->device] = p;
    }
    ctx->scratch = static_cast<unsigned char*>(g_scratch_ptr[ctx->device]);
    return 0;
}

size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

// Layout of one uqk_closure_solve_host_noalloc call in `scratch`: the
// closure counters, then u, lambda, status, iters and (if given) nTotal, each
// on a 256-byte boundary.  Returns the bytes used.
size_t noalloc_layout(int64_t n, int S, int M, bool with_nt, size_t off[6]) {
    const size_t cb = sizeof(double) * (size_t)n * S * M;
    const size_t ib = sizeof(int32_t) * (size_t)n;
    size_t at = 0;
    off[0] = at, at += align256(uqk::kClosureCounters * sizeof(unsigned long long));
    off[1] = at, at += align256(cb);
    off[2] = at, at += align256(cb);
    off[3] = at, at += align256(ib);
    off[4] = at, at += align256(ib);
    off[5] = at, at += with_nt ? align256(ib) : 0;
    return at;
}

}  // namespace

extern "C" {

int uqk_create(const uqk_desc* desc, uqk_ctx** out) {
    if (!desc || !out) return -1;
    *out = nullptr;
    uqk_ctx* ctx = new uqk_ctx();
    int S = expected_states(desc->closure);
    if (S < 0 || desc->n_states != S) {
        int rc = fail(ctx, -1, "closure kind and n_states do not match");
        *out = ctx;
        return rc;
    }
    if (desc->n_moments <= 0 || desc->n_quad <= 0 || !desc->phi || !desc->phiw) {
        int rc = fail(ctx, -1, "invalid sizes or tables");
        *out = ctx;
        return rc;
    }
    if (desc->precision != UQK_FP64 && desc->precision != UQK_MIXED) {
        int rc = fail(ctx, -1, "unknown precision");
        *out = ctx;
        return rc;
    }
    *out = ctx;
    ctx->device = desc->device;
    ctx->kind = desc->closure;
    ctx->S = S;
    ctx->M = desc->n_moments;
    ctx->Q = desc->n_quad;
    ctx->d = desc->dim_xi;
    ctx->precision = desc->precision;
    ctx->c.gamma = desc->gamma;
    ctx->c.u_minus = desc->u_minus;
    ctx->c.u_plus = desc->u_plus;
    ctx->c.g = desc->gravity;
    ctx->c.barrier = desc->barrier;
    ctx->c.alpha = desc->alpha;
    ctx->c.eps = desc->epsilon;
    ctx->c.max_iter = desc->max_iterations;
    ctx->c.max_halvings = desc->max_halvings;
    ctx->c.du_exact = desc->du_exact;
    ctx->c.safe = desc->safe_line_search;
    UQK_HIP(ctx, hipSetDevice(ctx->device));
    if (int rc = resolve_scratch(ctx)) return rc;
    hipDeviceProp_t prop;
    UQK_HIP(ctx, hipGetDeviceProperties(&prop, ctx->device));
    ctx->max_lds = prop.sharedMemPerBlock;
    ctx->num_cu = prop.multiProcessorCount;
    const size_t tb = sizeof(double) * (size_t)ctx->Q * ctx->M;
    UQK_HIP(ctx, hipMalloc(&ctx->d_phi, tb));
    UQK_HIP(ctx, hipMalloc(&ctx->d_phiw, tb));
    UQK_HIP(ctx, hipMemcpy(ctx->d_phi, desc->phi, tb, hipMemcpyHostToDevice));
    UQK_HIP(ctx, hipMemcpy(ctx->d_phiw, desc->phiw, tb, hipMemcpyHostToDevice));
    if (desc->xi && desc->dim_xi > 0) {
        const size_t xb = sizeof(double) * (size_t)ctx->Q * desc->dim_xi;
        UQK_HIP(ctx, hipMalloc(&ctx->d_xi, xb));
        UQK_HIP(ctx, hipMemcpy(ctx->d_xi, desc->xi, xb, hipMemcpyHostToDevice));
    }
    UQK_HIP(ctx, hipMalloc(&ctx->d_counters, uqk::kClosureCounters * sizeof(unsigned long long)));
    UQK_HIP(ctx, hipMalloc(&ctx->d_partial, sizeof(double) * uqk::sweep_partial_blocks(0)));
    UQK_HIP(ctx, hipMalloc(&ctx->d_scalars, 4 * sizeof(double)));
    UQK_HIP(ctx, hipStreamCreate(&ctx->own_stream));
    ctx->stream = ctx->own_stream;
    UQK_HIP(ctx, hipEventCreate(&ctx->ev0));
    UQK_HIP(ctx, hipEventCreate(&ctx->ev1));
    return 0;
}

int uqk_destroy(uqk_ctx* ctx) {
    if (!ctx) return 0;
    hipSetDevice(ctx->device);
    if (ctx->own_stream) hipStreamSynchronize(ctx->own_stream);
    hipFree(ctx->d_phi);
    hipFree(ctx->d_phiw);
    hipFree(ctx->d_xi);
    hipFree(ctx->d_counters);
    hipFree(ctx->d_partial);
    hipFree(ctx->d_scalars);
    hipFree(ctx->pinv_cell);
    hipFree(ctx->ufac_cell);
    hipFree(ctx->pinv_ok);
    hipFree(ctx->pinv_list);
    hipFree(ctx->check_list);
    hipFree(ctx->check_grad);
    hipFree(ctx->check_dt);
    if (ctx->graph_exec) hipGraphExecDestroy(ctx->graph_exec);
    hipFree(ctx->d_run);
    if (ctx->h_run) hipHostFree(ctx->h_run);
    for (hipEvent_t e : ctx->kev) hipEventDestroy(e);
    if (ctx->ev0) hipEventDestroy(ctx->ev0);
    if (ctx->ev1) hipEventDestroy(ctx->ev1);
    if (ctx->own_stream) hipStreamDestroy(ctx->own_stream);
    delete ctx;
    return 0;
}

const char* uqk_last_error(const uqk_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int uqk_set_stream(uqk_ctx* ctx, void* s) {
    if (!ctx) return -1;
    uqk::invalidate_graph(ctx);
    ctx->stream = s ? (hipStream_t)s : ctx->own_stream;
    return 0;
}

int uqk_set_closure_kernel(uqk_ctx* ctx, int32_t kernel) {
    if (!ctx) return -1;
    if (kernel < UQK_KERNEL_AUTO || kernel > UQK_KERNEL_BLOCK) return fail(ctx, -1, "unknown closure kernel");
    if (ctx->kernel != kernel) uqk::invalidate_graph(ctx);
    ctx->kernel = kernel;
    return 0;
}

int32_t uqk_last_closure_kernel(const uqk_ctx* ctx) { return ctx ? ctx->last_kernel : -1; }

int uqk_synchronize(uqk_ctx* ctx) {
    if (!ctx) return -1;
    UQK_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return 0;
}

int64_t uqk_closure_solve(uqk_ctx* ctx, int64_t n, const int32_t* nt, const double* u, double* lam,
                          int32_t* status, int32_t* iters) {
    if (!ctx) return -1;
    UQK_HIP(ctx, hipSetDevice(ctx->device));
    int rc = launch_closure(ctx, n, nt, u, lam, status, iters);
    if (rc) return rc;
    return finish_closure(ctx);
}

int64_t uqk_closure_solve_host(uqk_ctx* ctx, int64_t n, const int32_t* nt, const double* u, double* lam,
                               int32_t* status, int32_t* iters) {
    if (!ctx) return -1;
    UQK_HIP(ctx, hipSetDevice(ctx->device));
    const size_t cb = sizeof(double) * (size_t)n * ctx->S * ctx->M;
    double *du = nullptr, *dl = nullptr;
    int32_t *dnt = nullptr, *dst = nullptr, *dit = nullptr;
    int64_t result = 0;
    auto cleanup = [&]() {
        hipFree(du);
        hipFree(dl);
        hipFree(dnt);
        hipFree(dst);
        hipFree(dit);
    };
#define UQK_HIP_C(call)                                                                    \
    do {                                                                                   \
        hipError_t e_ = (call);                                                            \
        if (e_ != hipSuccess) {                                                            \
            cleanup();                                                                     \
            return fail(ctx, -2, std::string(#call ": ") + hipGetErrorString(e_));         \
        }                                                                                  \
    } while (0)
    if (n > 0) {
        UQK_HIP_C(hipMalloc(&du, cb));
        UQK_HIP_C(hipMalloc(&dl, cb));
        UQK_HIP_C(hipMalloc(&dst, sizeof(int32_t) * n));
        UQK_HIP_C(hipMalloc(&dit, sizeof(int32_t) * n));
        UQK_HIP_C(hipMemcpyAsync(du, u, cb, hipMemcpyHostToDevice, ctx->stream));
        UQK_HIP_C(hipMemcpyAsync(dl, lam, cb, hipMemcpyHostToDevice, ctx->stream));
        if (nt) {
            UQK_HIP_C(hipMalloc(&dnt, sizeof(int32_t) * n));
            UQK_HIP_C(hipMemcpyAsync(dnt, nt, sizeof(int32_t) * n, hipMemcpyHostToDevice, ctx->stream));
        }
    }
    int rc = launch_closure(ctx, n, dnt, du, dl, dst, dit);
    if (rc) {
        cleanup();
        return rc;
    }
    result = finish_closure(ctx);
    if (result >= 0 && n > 0) {
        UQK_HIP_C(hipMemcpyAsync(lam, dl, cb, hipMemcpyDeviceToHost, ctx->stream));
        if (status) UQK_HIP_C(hipMemcpyAsync(status, dst, sizeof(int32_t) * n, hipMemcpyDeviceToHost, ctx->stream));
        if (iters) UQK_HIP_C(hipMemcpyAsync(iters, dit, sizeof(int32_t) * n, hipMemcpyDeviceToHost, ctx->stream));
        UQK_HIP_C(hipStreamSynchronize(ctx->stream));
    }
#undef UQK_HIP_C
    cleanup();
    return result;
}

int64_t uqk_closure_solve_host_noalloc(uqk_ctx* ctx, int64_t n, const int32_t* nt, const double* u, double* lam,
                                       int32_t* status, int32_t* iters) {
    if (!ctx) return -1;
    if (n < 0 || (n > 0 && (!u || !lam))) return fail(ctx, -1, "uqk_closure_solve_host_noalloc: invalid arguments");
    if (ctx->warm)
        return fail(ctx, -1,
                    "uqk_closure_solve_host_noalloc: a learned warm start is attached (its buffers grow with the "
                    "batch); detach it with uqk_set_warm_start(ctx, NULL)");
    size_t off[6];
    const size_t need = noalloc_layout(n, ctx->S, ctx->M, nt != nullptr, off);
    if (need > kScratchBytes) {
        size_t off1[6];
        const size_t per_cell = noalloc_layout(1 << 16, ctx->S, ctx->M, nt != nullptr, off1) >> 16;
        return fail(ctx, -1,
                    "uqk_closure_solve_host_noalloc: " + std::to_string(n) + " cells need " + std::to_string(need) +
                        " bytes of staging, more than its fixed " + std::to_string(kScratchBytes >> 20) +
                        " MiB (about " + std::to_string(kScratchBytes / per_cell) +
                        " cells of this closure); split the batch or use uqk_closure_solve_host");
    }
    UQK_HIP(ctx, hipSetDevice(ctx->device));
    unsigned char* base = ctx->scratch;  // device address of `scratch` on this device
    auto* counters = reinterpret_cast<unsigned long long*>(base + off[0]);
    auto* du = reinterpret_cast<double*>(base + off[1]);
    auto* dl = reinterpret_cast<double*>(base + off[2]);
    auto* dst = reinterpret_cast<int32_t*>(base + off[3]);
    auto* dit = reinterpret_cast<int32_t*>(base + off[4]);
    auto* dnt = nt ? reinterpret_cast<int32_t*>(base + off[5]) : nullptr;
    const size_t cb = sizeof(double) * (size_t)n * ctx->S * ctx->M;
    if (n > 0) {
        UQK_HIP(ctx, hipMemcpyAsync(du, u, cb, hipMemcpyHostToDevice, ctx->stream));
        UQK_HIP(ctx, hipMemcpyAsync(dl, lam, cb, hipMemcpyHostToDevice, ctx->stream));
        if (nt) UQK_HIP(ctx, hipMemcpyAsync(dnt, nt, sizeof(int32_t) * n, hipMemcpyHostToDevice, ctx->stream));
    }
    int rc = launch_closure(ctx, n, dnt, du, dl, dst, dit, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
                            nullptr, nullptr, counters);
    if (rc) return rc;
    const int64_t result = finish_closure(ctx, counters);
    if (result >= 0 && n > 0) {
        UQK_HIP(ctx, hipMemcpyAsync(lam, dl, cb, hipMemcpyDeviceToHost, ctx->stream));
        if (status) UQK_HIP(ctx, hipMemcpyAsync(status, dst, sizeof(int32_t) * n, hipMemcpyDeviceToHost, ctx->stream));
        if (iters) UQK_HIP(ctx, hipMemcpyAsync(iters, dit, sizeof(int32_t) * n, hipMemcpyDeviceToHost, ctx->stream));
        UQK_HIP(ctx, hipStreamSynchronize(ctx->stream));
    }
    return result;
}

int uqk_eval_uq(uqk_ctx* ctx, int64_t n, const int32_t* nt, const double* lam, double* uq) {
    if (!ctx) return -1;
    uqk::SweepArgs a = sweep_args(ctx);
    a.ncell = n;
    a.nt_cell = nt;
    a.lam = lam;
    a.uq_out = uq;
    UQK_HIP(ctx, uqk::launch_eval_uq(ctx->kind, a, ctx->stream));
    UQK_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return 0;
}

int uqk_set_refinement(uqk_ctx* ctx, int32_t n_levels, const int32_t* ntotal_per_level, double refine_above,
                       double coarsen_below) {
    if (!ctx) return -1;
    if (n_levels <= 1) {
        ctx->refine = uqk::RefineArgs{};
        uqk::invalidate_graph(ctx);
        return 0;
    }
    if (n_levels > uqk::kMaxLevels || !ntotal_per_level)
        return fail(ctx, -1, "refinement: 2.." + std::to_string(uqk::kMaxLevels) + " levels and their nTotal");
    for (int l = 0; l < n_levels; ++l) {
        const int v = ntotal_per_level[l];
        if (v < 1 || v > ctx->M || (l > 0 && v <= ntotal_per_level[l - 1]))
            return fail(ctx, -1, "refinement: nTotal per level must ascend within 1..n_moments");
    }
    if (!(coarsen_below <= refine_above)) return fail(ctx, -1, "refinement: coarsen_below > refine_above");
    uqk::RefineArgs r{};
    r.S = ctx->S;
    r.M = ctx->M;
    r.dim_xi = ctx->d;
    r.nlev = n_levels;
    r.max_level = n_levels - 1;
    for (int l = 0; l < n_levels; ++l) r.nt_level[l] = ntotal_per_level[l];
    r.refine_above = refine_above;
    r.coarsen_below = coarsen_below;
    ctx->refine = r;
    uqk::invalidate_graph(ctx);  // the captured step changes
    return 0;
}

int uqk_set_refinement_cap(uqk_ctx* ctx, int32_t max_level) {
    if (!ctx) return -1;
    if (ctx->refine.nlev < 2) return fail(ctx, -1, "refinement levels are not set");
    if (max_level < 0 || max_level >= ctx->refine.nlev) return fail(ctx, -1, "refinement cap outside the levels");
    if (ctx->refine.max_level != max_level) {
        ctx->refine.max_level = max_level;
        uqk::invalidate_graph(ctx);  // the captured step changes
    }
    return 0;
}

int uqk_refine(uqk_ctx* ctx, int64_t n, const double* u, int32_t* nt) {
    if (!ctx) return -1;
    if (ctx->refine.nlev < 2) return fail(ctx, -1, "refinement levels are not set");
    if (n > 0 && (!u || !nt)) return fail(ctx, -1, "refine: null array");
    UQK_HIP(ctx, hipSetDevice(ctx->device));
    uqk::RefineArgs r = ctx->refine;
    r.ncell = n;
    r.u = u;
    r.nt_cell = nt;
    UQK_HIP(ctx, uqk::launch_refine(r, ctx->stream));
    UQK_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return 0;
}

int uqk_reconstruct(uqk_ctx* ctx, int64_t n, const int32_t* nt, const double* uq, double* u) {
    if (!ctx) return -1;
    uqk::SweepArgs a = sweep_args(ctx);
    a.ncell = n;
    a.nt_cell = nt;
    a.uq = uq;
    a.u_out = u;
    UQK_HIP(ctx, uqk::launch_reconstruct(a, ctx->stream));
    UQK_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return 0;
}

int uqk_compute_dt(uqk_ctx* ctx, const uqk_problem* prob, const uqk_mesh* mesh, const double* uq, double* dt_host) {
    if (!ctx || !prob || !mesh) return -1;
    if (int rc = check_mesh(ctx, mesh)) return rc;
    uqk::SweepArgs a = sweep_args(ctx);
    a.p = problem_params(prob);
    set_mesh(a, mesh);
    a.ncell = mesh->n_cells;
    a.uq = uq;
    a.scalar_out = ctx->d_scalars;
    UQK_HIP(ctx, uqk::launch_dt(ctx->kind, a, ctx->stream));
    UQK_HIP(ctx, hipMemcpyAsync(dt_host, ctx->d_scalars, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    UQK_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return 0;
}

int uqk_advance(uqk_ctx* ctx, const uqk_problem* prob, const uqk_mesh* mesh, const int32_t* nt, double dt,
                const double* u, const double* uq, double* unew) {
    if (!ctx || !prob || !mesh) return -1;
    if (int rc = check_mesh(ctx, mesh)) return rc;
    uqk::SweepArgs a = sweep_args(ctx);
    a.p = problem_params(prob);
    set_mesh(a, mesh);
    a.ncell = mesh->n_cells;
    a.nt_cell = nt;
    a.u = u;
    a.uq = uq;
    a.u_out = unew;
    a.dt = dt;
    UQK_HIP(ctx, uqk::launch_advance(ctx->kind, a, ctx->stream));
    UQK_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return 0;
}

int64_t uqk_step(uqk_ctx* ctx, const uqk_problem* prob, const uqk_mesh* mesh, int32_t* nt, double* u,
                 double* lam, double* uq, double* u_scratch, int32_t* status, int32_t* iters, double* dt_out,
                 double* residual_out) {
    if (!ctx || !prob || !mesh) return -1;
    if (int rc = check_mesh(ctx, mesh)) return rc;
    UQK_HIP(ctx, hipSetDevice(ctx->device));
    int rc = enqueue_step(ctx, prob, mesh, nt, u, lam, uq, u_scratch, status, iters, nullptr);
    if (rc) return rc;
    double sc[2];
    UQK_HIP(ctx, hipMemcpyAsync(sc, ctx->d_scalars, 2 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    int64_t nfail = finish_closure(ctx);
    if (nfail >= 0)
        if (int rc2 = collect_step_times(ctx)) return rc2;
    if (dt_out) *dt_out = sc[0];
    if (residual_out) *residual_out = sc[1];
    return nfail;
}

int64_t uqk_run_steps(uqk_ctx* ctx, const uqk_problem* prob, const uqk_mesh* mesh, int32_t* nt, double* u,
                      double* lam, double* uq, double* u_scratch, int32_t* status, int32_t* iters, double t0,
                      double t_end, int64_t max_steps, int32_t use_graph, double* t_out, int64_t* steps_out,
                      double* residual_out, int64_t* newton_out) {
    if (!ctx || !prob || !mesh) return -1;
    if (int rc = check_mesh(ctx, mesh)) return rc;
    UQK_HIP(ctx, hipSetDevice(ctx->device));
    if (!ctx->d_run) UQK_HIP(ctx, hipMalloc(&ctx->d_run, sizeof(uqk::RunState)));
    if (!ctx->h_run) UQK_HIP(ctx, hipHostMalloc((void**)&ctx->h_run, sizeof(uqk::RunState), hipHostMallocDefault));
    // the warm start's and the kept inverses' buffers must exist before the
    // step is captured
    if (int rc = uqk::prepare_warm_start(ctx, mesh->n_cells)) return rc;
    if (const int fmode = factor_mode(ctx, nt != nullptr))
        if (mesh->n_cells > 0)
            if (int rc = ensure_factor_buffers(ctx, mesh->n_cells, fmode)) return rc;
    if (ctx->residual_check && mesh->n_cells > 0)
        if (int rc = ensure_check_list(ctx, mesh->n_cells)) return rc;
    UQK_HIP(ctx, hipStreamSynchronize(ctx->stream));
    *ctx->h_run = uqk::RunState{};
    ctx->h_run->t = t0;
    ctx->h_run->t_end = t_end;
    ctx->h_run->max_steps = max_steps;
    ctx->h_run->active = (t0 < t_end && max_steps != 0) ? 1 : 0;
    UQK_HIP(ctx, hipMemcpy(ctx->d_run, ctx->h_run, sizeof(uqk::RunState), hipMemcpyHostToDevice));
    // what each step's step_end_kernel leaves for the next one: zero closure
    // counters and dt's start value (momentsolver.cpp:127)
    UQK_HIP(ctx, uqk::launch_fill_u64(ctx->d_counters, uqk::kClosureCounters, ctx->stream));
    UQK_HIP(ctx, uqk::launch_fill_f64(ctx->d_scalars, 1, 1e10, ctx->stream));
    const int* active = &ctx->d_run->active;
    // check_mode: 1 the residual check evaluates U (the run's first step:
    // the caller may have changed the state since the last call), 2 it takes
    // the gradient from the last step's outputs
    auto enqueue = [&](int check_mode) -> int {
        ctx->check_mode = check_mode;
        int rc = enqueue_step(ctx, prob, mesh, nt, u, lam, uq, u_scratch, status, iters, active, true);
        ctx->check_mode = 0;
        if (rc) return rc;
        UQK_HIP(ctx, uqk::launch_step_end(ctx->d_run, ctx->d_scalars, ctx->d_counters, ctx->d_partial,
                                          ctx->step_recon ? 1 : 0, uqk::sweep_partial_blocks(0), ctx->stream));
        return 0;
    };
    const int64_t limit = max_steps < 0 ? INT64_MAX : max_steps;
    int64_t launched = 0;
    int later = 1;  // the later steps' check_mode
    if (limit > 0) {
        if (int rc = enqueue(1)) return rc;
        launched = 1;
        later = ctx->step_recon ? 2 : 1;
        UQK_HIP(ctx, hipStreamSynchronize(ctx->stream));
    }
    if (use_graph && launched < limit) {
        // the arguments and the context state the captured launches depend on
        // (the setters of that state also drop the graph)
        char key[640];
        snprintf(key, sizeof key,
                 "%d %d %.17g %.17g %.17g %.17g %d|%lld %d %d %p %p %p %p %p %p %p|%p %p %p %p %p %p %p|"
                 "%d %d %d %p %p %d",
                 prob->problem, prob->flux, prob->cfl, prob->gamma, prob->gravity, prob->friction, prob->friction_dim,
                 (long long)mesh->n_cells, mesh->n_edges, mesh->dim, (const void*)mesh->neighbor,
                 (const void*)mesh->unit_normal, (const void*)mesh->normal, (const void*)mesh->area,
                 (const void*)mesh->dx_param, (const void*)mesh->bc, (const void*)mesh->boundary_normal,
                 (const void*)nt, (void*)u, (void*)lam, (void*)uq, (void*)u_scratch, (void*)status, (void*)iters,
                 ctx->kernel, ctx->shape.waves, ctx->shape.blocks_per_cu,
                 (void*)ctx->warm, (void*)ctx->stream, later);
        if (!ctx->graph_exec || ctx->graph_key != key) {
            if (ctx->graph_exec) hipGraphExecDestroy(ctx->graph_exec);
            ctx->graph_exec = nullptr;
            hipGraph_t g = nullptr;
            UQK_HIP(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
            ctx->capturing = true;
            int rc = enqueue(later);
            ctx->capturing = false;
            hipError_t ec = hipStreamEndCapture(ctx->stream, &g);
            if (rc) {
                if (g) hipGraphDestroy(g);
                return rc;
            }
            UQK_HIP(ctx, ec);
            hipError_t ei = hipGraphInstantiate(&ctx->graph_exec, g, nullptr, nullptr, 0);
            hipGraphDestroy(g);
            UQK_HIP(ctx, ei);
            ctx->graph_key = key;
        }
    }
    // replay; every `chunk` steps read the run state back (into pinned host
    // memory owned by the context) to stop early
    uqk::RunState& rs = *ctx->h_run;
    const int64_t chunk = 64;
    for (;;) {
        const int64_t todo = std::min<int64_t>(chunk, limit - launched);
        for (int64_t i = 0; i < todo; ++i) {
            if (use_graph) {
                UQK_HIP(ctx, hipGraphLaunch(ctx->graph_exec, ctx->stream));
            } else {
                int rc = enqueue(later);
                if (rc) return rc;
            }
        }
        launched += todo;
        UQK_HIP(ctx, hipMemcpyAsync(ctx->h_run, ctx->d_run, sizeof(uqk::RunState), hipMemcpyDeviceToHost, ctx->stream));
        UQK_HIP(ctx, hipStreamSynchronize(ctx->stream));
        if (int rc = collect_step_times(ctx)) return rc;
        if (!rs.active || launched >= limit || rs.nfail) break;
    }
    ctx->timed = false;
    ctx->last_stats[0] = (unsigned long long)rs.nfail;  // uqk_last_closure_stats: the run's totals
    ctx->last_stats[1] = (unsigned long long)rs.newton;
    ctx->last_stats[2] = (unsigned long long)rs.gradients;
    ctx->last_stats[3] = 0;
    ctx->last_stats[4] = (unsigned long long)rs.factorizations;
    ctx->last_stats[5] = (unsigned long long)rs.reuse_steps;
    if (t_out) *t_out = rs.t;
    if (steps_out) *steps_out = rs.steps;
    if (residual_out) *residual_out = rs.residual;
    if (newton_out) *newton_out = rs.newton;
    return rs.nfail;
}

int uqk_mean_var(uqk_ctx* ctx, int64_t n, const int32_t* nt, const double* lam, double* mv) {
    if (!ctx) return -1;
    uqk::SweepArgs a = sweep_args(ctx);
    a.ncell = n;
    a.nt_cell = nt;
    a.lam = lam;
    a.mv_out = mv;
    UQK_HIP(ctx, uqk::launch_mean_var(ctx->kind, a, ctx->stream));
    UQK_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return 0;
}

double uqk_last_closure_ms(const uqk_ctx* ctx) { return ctx ? ctx->last_closure_ms : 0.0; }

int uqk_last_closure_stats(const uqk_ctx* ctx, int64_t* newton_steps, int64_t* gradient_evaluations) {
    if (!ctx) return -1;
    if (newton_steps) *newton_steps = (int64_t)ctx->last_stats[1];
    if (gradient_evaluations) *gradient_evaluations = (int64_t)ctx->last_stats[2];
    return 0;
}

int uqk_set_kernel_timing(uqk_ctx* ctx, int32_t on) {
    if (!ctx) return -1;
    if (ctx->kev_used) UQK_HIP(ctx, hipStreamSynchronize(ctx->stream));
    ctx->kev_used = 0;
    ctx->ktiming = on != 0;
    ctx->k_closure_ms = ctx->k_advance_ms = 0.0;
    ctx->k_steps = 0;
    return 0;
}

int uqk_set_residual_check(uqk_ctx* ctx, int32_t on) {
    if (!ctx) return -1;
    if (ctx->residual_check != (on != 0)) uqk::invalidate_graph(ctx);
    ctx->residual_check = on != 0;
    return 0;
}

int uqk_set_factor_reuse(uqk_ctx* ctx, int32_t on) {
    if (!ctx) return -1;
    if (ctx->factor_reuse != (on != 0)) uqk::invalidate_graph(ctx);
    ctx->factor_reuse = on != 0;
    return 0;
}

int uqk_invalidate_factors(uqk_ctx* ctx) {
    if (!ctx) return -1;
    if (ctx->pinv_ok && ctx->pinv_cell_cap > 0) {
        UQK_HIP(ctx, hipSetDevice(ctx->device));
        UQK_HIP(ctx, hipMemsetAsync(ctx->pinv_ok, 0, sizeof(int) * (size_t)ctx->pinv_cell_cap, ctx->stream));
        UQK_HIP(ctx, hipStreamSynchronize(ctx->stream));
    }
    return 0;
}

int uqk_last_closure_factorizations(const uqk_ctx* ctx, int64_t* factorizations, int64_t* reuse_steps) {
    if (!ctx) return -1;
    if (factorizations) *factorizations = (int64_t)ctx->last_stats[4];
    if (reuse_steps) *reuse_steps = (int64_t)ctx->last_stats[5];
    return 0;
}

int uqk_kernel_times(const uqk_ctx* ctx, double* closure_ms, double* advance_ms, int64_t* steps) {
    if (!ctx) return -1;
    if (closure_ms) *closure_ms = ctx->k_closure_ms;
    if (advance_ms) *advance_ms = ctx->k_advance_ms;
    if (steps) *steps = ctx->k_steps;
    return 0;
}

int uqk_set_launch_shape(uqk_ctx* ctx, int32_t waves, int32_t blocks_per_cu) {
    if (!ctx) return -1;
    if (waves != 0 && waves != 1 && waves != 2 && waves != 4) return fail(ctx, -1, "waves per workgroup: 0, 1, 2 or 4");
    if (blocks_per_cu < 0) return fail(ctx, -1, "workgroups per CU must be >= 0");
    if (ctx->shape.waves != waves || ctx->shape.blocks_per_cu != blocks_per_cu) uqk::invalidate_graph(ctx);
    ctx->shape.waves = waves;
    ctx->shape.blocks_per_cu = blocks_per_cu;
    return 0;
}

int uqk_get_launch_shape(const uqk_ctx* ctx, int32_t* waves, int32_t* blocks_per_cu) {
    if (!ctx) return -1;
    if (waves) *waves = ctx->shape.waves;
    if (blocks_per_cu) *blocks_per_cu = ctx->shape.blocks_per_cu;
    return 0;
}

int uqk_autotune(uqk_ctx* ctx, int64_t n, const double* u, const double* lambda, const char* json_path) {
    if (!ctx || n <= 0 || !u || !lambda) return -1;
    // the kernel the context runs: the one-wavefront kernel where it fits,
    // else the workgroup kernel (n = S*M >= 64), whose only launch parameter
    // is the number of persistent workgroups per CU (one wave count: 4)
    int wmax = 0, occ = 0;
    const bool wave = ctx->kernel != UQK_KERNEL_BLOCK &&
                      uqk::newton_wave_limits(ctx->kind, ctx->M, ctx->Q, ctx->precision, (int)ctx->max_lds, 4, &wmax,
                                              &occ);
    const bool block = !wave && (ctx->kernel == UQK_KERNEL_AUTO || ctx->kernel == UQK_KERNEL_BLOCK) &&
                       uqk::newton_block_fits(ctx->S, ctx->M, ctx->Q, (int)ctx->max_lds);
    if (!wave && !block) return fail(ctx, -3, "uqk_autotune: no specialised closure kernel for this shape");
    if (block) wmax = 4;
    const size_t bytes = sizeof(double) * (size_t)n * ctx->S * ctx->M;
    double* trial = nullptr;
    UQK_HIP(ctx, hipMalloc(&trial, bytes));
    const uqk::LaunchShape keep = ctx->shape;
    const int kernel_keep = ctx->kernel;
    ctx->kernel = wave ? UQK_KERNEL_WAVE : UQK_KERNEL_BLOCK;
    double best_ms = -1.0;
    uqk::LaunchShape best = keep;
    int timed = 0, rc = 0;
    for (int waves = block ? 4 : 1; waves <= wmax && rc == 0; waves *= 2) {
        int wm = 0, oc = 0;
        if (wave)
            uqk::newton_wave_limits(ctx->kind, ctx->M, ctx->Q, ctx->precision, (int)ctx->max_lds, waves, &wm, &oc);
        else
            oc = uqk::newton_block_occupancy(ctx->kind, ctx->M, ctx->Q, ctx->precision);
        for (int per_cu = 1; per_cu <= oc && rc == 0; ++per_cu) {
            ctx->shape.waves = waves;
            ctx->shape.blocks_per_cu = per_cu;
            double ms_min = 1e300;
            for (int rep = 0; rep < 2 && rc == 0; ++rep) {  // first run warms the caches
                if (hipMemcpyAsync(trial, lambda, bytes, hipMemcpyDeviceToDevice, ctx->stream) != hipSuccess) {
                    rc = fail(ctx, -2, "uqk_autotune: copy failed");
                    break;
                }
                rc = launch_closure(ctx, n, nullptr, u, trial, nullptr, nullptr);
                if (rc) break;
                if (finish_closure(ctx) < 0) rc = -2;
                if (rep == 1) ms_min = ctx->last_closure_ms;
            }
            if (rc) break;
            ++timed;
            if (best_ms < 0 || ms_min < best_ms) {
                best_ms = ms_min;
                best = ctx->shape;
            }
        }
    }
    hipFree(trial);
    ctx->kernel = kernel_keep;
    ctx->shape = rc ? keep : best;
    uqk::invalidate_graph(ctx);
    if (rc) return rc;
    if (json_path) {
        FILE* f = std::fopen(json_path, "a");
        if (!f) return fail(ctx, -1, std::string("uqk_autotune: cannot write ") + json_path);
        std::fprintf(f,
                     "{\"closure\": %d, \"n_states\": %d, \"n_moments\": %d, \"n_quad\": %d, \"precision\": %d, "
                     "\"waves\": %d, \"blocks_per_cu\": %d, \"ms\": %.6f, \"cells\": %lld}\n",
                     ctx->kind, ctx->S, ctx->M, ctx->Q, ctx->precision, best.waves, best.blocks_per_cu, best_ms,
                     (long long)n);
        std::fclose(f);
    }
    return timed;
}

int uqk_load_tuning(uqk_ctx* ctx, const char* json_path) {
    if (!ctx || !json_path) return -1;
    FILE* f = std::fopen(json_path, "r");
    if (!f) return fail(ctx, -1, std::string("uqk_load_tuning: cannot open ") + json_path);
    char line[512];
    int found = 0;
    while (std::fgets(line, sizeof(line), f)) {
        int k = -1, S = -1, M = -1, Q = -1, p = -1, w = 0, b = 0;
        if (std::sscanf(line,
                        "{\"closure\": %d, \"n_states\": %d, \"n_moments\": %d, \"n_quad\": %d, \"precision\": %d, "
                        "\"waves\": %d, \"blocks_per_cu\": %d",
                        &k, &S, &M, &Q, &p, &w, &b) == 7 &&
            k == ctx->kind && S == ctx->S && M == ctx->M && Q == ctx->Q && p == ctx->precision) {
            if (ctx->shape.waves != w || ctx->shape.blocks_per_cu != b) uqk::invalidate_graph(ctx);
            ctx->shape.waves = w;  // the last matching line wins
            ctx->shape.blocks_per_cu = b;
            found = 1;
        }
    }
    std::fclose(f);
    return found;
}

}  // extern "C"
