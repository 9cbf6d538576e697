/*
 * uqk.h -- C ABI of the MI355X-native IPM moment-solver hot path.
 *
 * Every entry point is extern "C", takes plain pointers and sizes, and returns
 * 0 on success or a negative error code (uqk_last_error() holds the message).
 * Device-pointer entry points expect buffers in HBM on the context's device;
 * *_host entry points accept host buffers and copy them.
 *
 * Host array layouts are the reference's:
 *   per-cell moments u / duals lambda : Matrix(S, M) column-major, state fastest,
 *                                       element (s, i) at [i*S + s]; cells contiguous
 *                                       (code/src/matrix.cpp:107, momentsolver.h:24).
 *   per-cell quadrature values uq     : Matrix(S, Q) column-major, (s, k) at [k*S + s]
 *                                       (momentsolver.cpp:50).
 *   basis tables phi, phiw            : (k, i) at [k*M + i]; phi = Closure::_phiTilde,
 *                                       phiw = Closure::_phiTildeWf (closure.cpp:45-66).
 */
#ifndef UQK_H
#define UQK_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct uqk_ctx uqk_ctx;

/* Closure kinds: Closure::Create (closure.cpp:94-131). */
enum {
    UQK_BOUNDED_BARRIER = 0,           /* BoundedBarrier (boundedbarrier.cpp)            */
    UQK_EULER1D = 1,                   /* EulerClosure (eulerclosure.cpp)                */
    UQK_EULER2D = 2,                   /* EulerClosure2D (eulerclosure2d.cpp)            */
    UQK_SHALLOW_WATER2D = 3,           /* ShallowWaterClosure2D (shallowwaterclosure2d)  */
    UQK_SHALLOW_WATER2D_BARRIER = 4    /* SW2D with a log barrier on h (no reference)    */
};

/* Arithmetic of the Hessian assembly. */
enum {
    UQK_FP64 = 0,   /* everything in FP64 (the plain path)                        */
    UQK_MIXED = 1   /* Hessian on MFMA in reduced precision, FP64 Cholesky,
                       FP64 gradient/residual/accept test                         */
};

/* Per-cell closure status (replaces the reference's exit(EXIT_FAILURE)). */
enum {
    UQK_STATUS_CONVERGED = 0,
    UQK_STATUS_MAX_ITERATIONS = 1,     /* closure.cpp:191-192                     */
    UQK_STATUS_LINE_SEARCH = 2,        /* closure.cpp:180-183                     */
    UQK_STATUS_NOT_SPD = 3,            /* vectorspace.cpp:166-170 (dposv info>0)  */
    UQK_STATUS_NONFINITE = 4           /* accepted a non-finite trial residual    */
};

/* Problem kinds for the flux / time step (Problem::Create, problem.cpp:18-42). */
enum { UQK_BURGERS = 0, UQK_EULER2D_PROBLEM = 1, UQK_SHALLOW_WATER2D_PROBLEM = 2 };
/* Flux flavour: the reference's (Burgers upwind / Euler2D, SW2D HLL) or the
 * Lax-Friedrichs (Burgers) / Rusanov (2D) kinetic flux. */
enum { UQK_FLUX_REFERENCE = 0, UQK_FLUX_LAX_FRIEDRICHS = 1 };
/* Boundary types (BoundaryType in the reference's mesh.h). */
enum { UQK_BC_NONE = 0, UQK_BC_DIRICHLET = 1, UQK_BC_NOSLIP = 2, UQK_BC_SWWALL = 3 };

typedef struct {
    int32_t closure;            /* UQK_BOUNDED_BARRIER ...                        */
    int32_t n_states;           /* S                                              */
    int32_t n_moments;          /* M = nTotal of the finest level                 */
    int32_t n_quad;             /* Q = nQTotal                                    */
    int32_t dim_xi;             /* d, number of uncertain dimensions              */
    double gamma;               /* Euler adiabatic exponent                       */
    double u_minus, u_plus;     /* bounded-barrier bounds                         */
    double gravity;             /* shallow water g                                */
    double barrier;             /* SW2D barrier strength                          */
    double alpha;               /* Newton damping (closure-specific default)      */
    double epsilon;             /* convergence threshold on ||grad L||_2          */
    int32_t max_iterations;     /* Newton iterations (1 = one-shot IPM)           */
    int32_t max_halvings;       /* line-search halvings (reference: 1000)         */
    int32_t du_exact;           /* Euler1D: 1 exact DU, 0 DU(2,2) as shipped      */
    int32_t safe_line_search;   /* 1: SolveClosureSafe semantics                  */
    int32_t precision;          /* UQK_FP64 or UQK_MIXED                          */
    const double* phi;          /* host, Q x M                                    */
    const double* phiw;         /* host, Q x M                                    */
    const double* xi;           /* host, Q x d quadrature nodes (for sources)     */
    int32_t device;             /* HIP device ordinal                             */
} uqk_desc;

typedef struct {
    int32_t problem;            /* UQK_BURGERS ...                                */
    int32_t flux;               /* UQK_FLUX_REFERENCE or UQK_FLUX_LAX_FRIEDRICHS  */
    double cfl;
    double gamma;
    double gravity;
    double friction;            /* SW2D bed friction base coefficient (0 = off)   */
    int32_t friction_dim;       /* xi dimension scaling the friction, -1 = none   */
} uqk_problem;

/* Mesh arrays on the device (see uqk_mesh_upload for host arrays).
 * neighbor id == n_cells marks a boundary edge (mesh2d.cpp:194). */
typedef struct {
    int64_t n_cells;
    int32_t n_edges;            /* edges per cell: 2 (1D lines) or 4 (quads)      */
    int32_t dim;                /* 1 or 2                                         */
    const int32_t* neighbor;    /* n_cells * n_edges                              */
    const double* unit_normal;  /* n_cells * n_edges * dim                        */
    const double* normal;       /* n_cells * n_edges * dim, scaled by edge length */
    const double* area;         /* n_cells                                        */
    const double* dx_param;     /* n_cells: maxEdge / area (ComputeDt argument)   */
    const int32_t* bc;          /* n_cells, UQK_BC_*                              */
    const double* boundary_normal; /* n_cells * dim                               */
} uqk_mesh;

/* Context: Closure::Create + the Closure constructor tables
 * (closure.cpp:16-84, :94-131).  Uploads phi/phiw, picks kernels. */
int uqk_create(const uqk_desc* desc, uqk_ctx** out);
int uqk_destroy(uqk_ctx* ctx);
const char* uqk_last_error(const uqk_ctx* ctx);
/* Launch on a caller-provided hipStream_t (NULL = the context's own stream). */
int uqk_set_stream(uqk_ctx* ctx, void* hip_stream);
int uqk_synchronize(uqk_ctx* ctx);

/* Closure kernel selection: AUTO picks the compile-time specialised wave
 * kernel when one exists for (closure, M, Q) and the batch has one order,
 * else the workgroup-per-cell kernel for systems of 64..256 duals, else the
 * generic wave-per-cell kernel (any size, per-cell orders, FP64 Hessian;
 * below 64 duals it outruns the workgroup kernel's mixed precision, which
 * UQK_KERNEL_BLOCK still selects). */
enum { UQK_KERNEL_AUTO = 0, UQK_KERNEL_GENERIC = 1, UQK_KERNEL_WAVE = 2, UQK_KERNEL_BLOCK = 3 };
int uqk_set_closure_kernel(uqk_ctx* ctx, int32_t kernel);
/* Kernel used by the last closure call (UQK_KERNEL_GENERIC / _WAVE / _BLOCK). */
int32_t uqk_last_closure_kernel(const uqk_ctx* ctx);

/* Closure::SolveClosure / SolveClosureSafe (closure.cpp:133-193, :297-339),
 * batched over cells: the per-cell loop of momentsolver.cpp:83-89.
 * lambda: in = warm start, out = solution.  ntotal_cell: per-cell nTotal
 * (adaptive levels, momentsolver.cpp:88) or NULL for M.  status/iters may be NULL.
 * Device pointers.  Returns the number of cells with nonzero status (>= 0). */
int64_t uqk_closure_solve(uqk_ctx* ctx, int64_t n_cells, const int32_t* ntotal_cell, const double* u,
                          double* lambda, int32_t* status, int32_t* iters);
/* Same with host pointers (copies in and out). */
int64_t uqk_closure_solve_host(uqk_ctx* ctx, int64_t n_cells, const int32_t* ntotal_cell, const double* u,
                               double* lambda, int32_t* status, int32_t* iters);
/* uqk_closure_solve_host without any device allocation.  No counterpart in
 * the reference: it is for callers that cannot allocate device memory in
 * their loop (serving loops, harnesses that forbid hipMalloc).  The device
 * staging of the call (counters, u, lambda, status, iters, nTotal) lives in
 * one fixed 64 MiB array per GPU, held from the library's load until the
 * process ends; batches that do not fit (about 50,000 cells at 40 duals) are
 * rejected with an error.  Calls on one GPU share that array: calls that
 * overlap on one GPU (other threads, other contexts) race.  A context with a
 * learned warm start attached is rejected.  Same results as
 * uqk_closure_solve_host. */
int64_t uqk_closure_solve_host_noalloc(uqk_ctx* ctx, int64_t n_cells, const int32_t* ntotal_cell, const double* u,
                                       double* lambda, int32_t* status, int32_t* iters);

/* uQ_j = U(EvaluateLambdaOnPE(lambda_j)) (momentsolver.cpp:99-101,
 * closure.cpp:217-230 with the U(Matrix) form).  Device pointers. */
int uqk_eval_uq(uqk_ctx* ctx, int64_t n_cells, const int32_t* ntotal_cell, const double* lambda, double* uq);
/* u_j = uQ_j * phiw over the cell's nTotal moments, rest zero
 * (momentsolver.cpp:120-124, vectorspace.cpp:174-185).  Device pointers. */
int uqk_reconstruct(uqk_ctx* ctx, int64_t n_cells, const int32_t* ntotal_cell, const double* uq, double* u);
/* Adaptive refinement levels (momentsolver.cpp:104-117 with
 * ComputeRefIndicator, :271-289).  ntotal_per_level: nTotal of each of the
 * n_levels levels, ascending, within 1..n_moments.  A cell at level l has
 * the indicator |u(0,nt-1)| + |u(0,nt-2)| for one uncertain dimension, else
 * sum |u(0,i)| over the moments level l adds (from nTotal(l-1), or 1 at level
 * 0); it moves up a level above refine_above and down below coarsen_below (the
 * reference hard-codes 0.02 and 0.004).  With n_levels > 1, uqk_step and
 * uqk_run_steps refine every step between uQ and the reconstruction, and
 * update ntotal_cell (then required) in place; n_levels <= 1 turns it off. */
int uqk_set_refinement(uqk_ctx* ctx, int32_t n_levels, const int32_t* ntotal_per_level, double refine_above,
                       double coarsen_below);
/* Refinement retardation (Paper1/adaptivity.tex:61-86, not in the reference
 * code): no cell moves above level max_level (0 .. n_levels-1; a cell above it
 * comes down to it at its next refinement pass).  uqk_set_refinement resets
 * the cap to the finest level. */
int uqk_set_refinement_cap(uqk_ctx* ctx, int32_t max_level);
/* One refinement pass over n_cells: ntotal_cell (device, in/out) from the
 * moments u (device).  Cells whose ntotal_cell is no level's nTotal are left
 * unchanged.  Synchronous. */
int uqk_refine(uqk_ctx* ctx, int64_t n_cells, const double* u, int32_t* ntotal_cell);
/* dt = min_j Problem::ComputeDt(uQ_j, maxEdge_j/area_j) (momentsolver.cpp:127-133).
 * Result written to *dt_host. */
int uqk_compute_dt(uqk_ctx* ctx, const uqk_problem* prob, const uqk_mesh* mesh, const double* uq, double* dt_host);
/* ExplicitEuler::Advance (expliciteuler.cpp:8-75) with MomentSolver::numFlux
 * (momentsolver.cpp:413-416): unew = u - dt/A * sum_edges G(uQ_j, uQ_nb) phiw. */
int uqk_advance(uqk_ctx* ctx, const uqk_problem* prob, const uqk_mesh* mesh, const int32_t* ntotal_cell,
                double dt, const double* u, const double* uq, double* unew);
/* One MomentSolver::Solve step body (momentsolver.cpp:83-152): closure, uQ,
 * reconstruction, dt, advance, residual.  u, lambda updated in place (device);
 * ntotal_cell too when refinement levels are set (uqk_set_refinement).
 * Writes dt and the residual of the state-0 mean.  Returns the number of
 * cells whose closure failed (>= 0) or an error code. */
int64_t uqk_step(uqk_ctx* ctx, const uqk_problem* prob, const uqk_mesh* mesh, int32_t* ntotal_cell,
                 double* u, double* lambda, double* uq, double* u_scratch, int32_t* status, int32_t* iters,
                 double* dt_out, double* residual_out);
/* The time loop of MomentSolver::Solve (momentsolver.cpp:79-162) on the
 * device: while t < t_end and fewer than max_steps steps (max_steps < 0: no
 * limit) and no closure failure, run the step body.  The call's first step
 * is enqueued on its own; with use_graph the later steps' body is captured
 * once into a hipGraph (re-captured when any argument changes) and
 * replayed; t, the step count and failures live on the device and are read
 * back every 64 steps, so there is no host synchronisation inside a step.
 * Returns the number of failed cells of the last step (>= 0) or an error. */
int64_t uqk_run_steps(uqk_ctx* ctx, const uqk_problem* prob, const uqk_mesh* mesh, int32_t* ntotal_cell,
                      double* u, double* lambda, double* uq, double* u_scratch, int32_t* status, int32_t* iters,
                      double t0, double t_end, int64_t max_steps, int32_t use_graph, double* t_out, int64_t* steps_out,
                      double* residual_out, int64_t* newton_steps_out);
/* WriteMeanAndVar (momentsolver.cpp:302-328): mv is (2S) x n_cells,
 * (f, j) at [j*2S + f].  Device pointers. */
int uqk_mean_var(uqk_ctx* ctx, int64_t n_cells, const int32_t* ntotal_cell, const double* lambda, double* mv);

/* ---- host utilities (no GPU work) ------------------------------------- */
/* Gauss-Legendre rule, nodes ascending (legendre.cpp:9-28, polynomial.cpp:7-16). */
int uqk_gauss_legendre(int32_t n, double* x, double* w);
/* Multi-index set of the moment basis (settings.cpp:242-266): idx is
 * (count x d), dimension 0 fastest; ntotal_per_level[l] = nTotal of level l.
 * Returns the count, or -1 if it exceeds cap. */
int32_t uqk_index_set(int32_t d, int32_t n_levels, const int32_t* degrees, int32_t max_degree, int32_t cap,
                      int32_t* idx, int32_t* ntotal_per_level);
/* Tensor Gauss-Legendre grid and the Closure tables phi / phiw
 * (tensorizedquadrature.cpp:27-74, closure.cpp:45-66).  Returns Q. */
int32_t uqk_basis_tables(int32_t d, int32_t q1, int32_t m, const int32_t* idx, double* xi, double* wq, double* phi,
                         double* phiw);

/* ---- synthetic workload and device memory ------------------------------ */
/* Batched-closure benchmark input (no reference counterpart): 1D Euler cells
 * with seeded random duals lambda* around a Maxwellian (counter-based Philox,
 * key = seed, counter = (cell, attempt, draw)), moments u = <U(lambda* phi) phi>
 * and the Newton start lambda0 = (DS(u_0), 0, ...).  phi/phiw are host tables
 * (Q x M); u, lam_star, lam0 are (ncell x M x 3).  The first form writes device
 * buffers, the _host form host buffers.  Returns the number of cells that
 * found no admissible draw (0 expected), or a negative error. */
int64_t uqk_synth_euler1d_closure(int64_t ncell, uint64_t seed, int32_t M, int32_t Q, const double* phi,
                                  const double* phiw, double* u, double* lam_star, double* lam0);
int64_t uqk_synth_euler1d_closure_host(int64_t ncell, uint64_t seed, int32_t M, int32_t Q, const double* phi,
                                       const double* phiw, double* u, double* lam_star, double* lam0);
/* Device memory on the current device, for callers without an allocator.
 * uqk_memcpy kind: 0 host->device, 1 device->host, 2 device->device. */
void* uqk_device_alloc(int64_t bytes);
int uqk_device_free(void* ptr);
int uqk_memcpy(void* dst, const void* src, int64_t bytes, int32_t kind);
int uqk_device_synchronize(void);
int32_t uqk_device_count(void);

/* Milliseconds spent in the closure kernel by the last uqk_closure_solve /
 * uqk_step call, timed with HIP events on the launch stream. */
double uqk_last_closure_ms(const uqk_ctx* ctx);
/* Totals of the last closure call: Newton steps (those on a reused
 * factorisation included) and gradient evaluations (line-search trials included) over all cells; after
 * uqk_run_steps, the totals over the run's steps. */
int uqk_last_closure_stats(const uqk_ctx* ctx, int64_t* newton_steps, int64_t* gradient_evaluations);
/* Reuse of the factorisation in the mixed closure (no counterpart in the
 * reference; on by default for UQK_MIXED contexts).  In the time loop, after
 * a cell's FP64 Cholesky of its (FP32-assembled) Hessian in one step, the
 * library keeps P = H^{-1} of the factored matrix: the wave kernel (uniform
 * orders, n < 64) as an explicit inverse computed after the launch (FP64,
 * kept in FP32), the workgroup kernel (n up to 128, per-cell orders too) as
 * the factor U rounded to FP32 and tagged with the cell's order, replaced
 * by U^{-1} (P = U^{-1} U^{-T}) the first time a later step uses it.  The
 * cell's Newton steps in later time steps take d = P g refined once with
 * the FP64 Hessian-vector product (d += P (g - H d)) and keep it when the
 * full step lowers ||grad L|| (the FP64 accept test); otherwise, and after
 * a contraction clearly worse than a damped Newton step's (above
 * 1.2 (1 - alpha) + 1e-3; + 0.05 from n = 64 on, where a factorisation costs
 * over ten steps on the kept one), the cell's current Hessian is factored
 * and the reference Newton step runs.  Memory per cell: n (n + 1) floats and
 * n (n + 3) / 2 doubles (wave kernel), n (n + 1) / 2 floats (workgroup
 * kernel), n = S M.  In a single closure call (wave kernel, uniform
 * orders, damped Newton: alpha < 1, whose steps contract linearly anyway)
 * a cell's Newton steps after a factorisation
 * solve with that factor (d = (U^T U)^{-1} g) under the same accept and
 * contraction tests, with no memory beyond the call.  Results match the
 * plain mixed closure to the Newton tolerance, not bit for bit.
 * uqk_invalidate_factors forgets the kept factorisations (after replacing the
 * duals wholesale). */
int uqk_set_factor_reuse(uqk_ctx* ctx, int32_t on);
/* Residual check before the Newton kernel in the time loop (no counterpart
 * in the reference; on by default; not with SolveClosureSafe).
 * Closure::SolveClosure stops at once when ||grad L||_2 < epsilon at the
 * cell's lambda (closure.cpp:133-193); a check kernel evaluates that
 * gradient for every cell -- one lane per cell for the one-wavefront
 * kernel's shapes, one wave per cell for the workgroup kernel's (n up to
 * 128, per-cell orders too) -- finishes the cells that pass (uQ, the
 * moments, ComputeDt, status, zero iterations) and lists the others, which
 * the Newton kernel then takes, starting from the gradient the check
 * found.  Within one uqk_run_steps call, the steps after the first take
 * that gradient from the previous step's outputs instead (the moments
 * <U(lambda phi) phiW> it left in u_scratch, minus u), and a cell that
 * passes keeps its uQ, moments and ComputeDt: the first step of every call
 * evaluates U, since the caller may have changed u, lambda or u_scratch
 * between calls.  Lambda and U at the points are the Newton kernel's; the
 * results agree with the check off to rounding. */
int uqk_set_residual_check(uqk_ctx* ctx, int32_t on);
int uqk_invalidate_factors(uqk_ctx* ctx);
/* Cholesky factorisations and steps on a kept inverse of the last closure
 * call (after uqk_run_steps: of the run). */
int uqk_last_closure_factorizations(const uqk_ctx* ctx, int64_t* factorizations, int64_t* reuse_steps);
/* Per-step kernel timing (no counterpart in the reference): while on, every
 * step that uqk_step, or uqk_run_steps without a graph, enqueues brackets its
 * closure kernel and its advance kernel with HIP events on the context's
 * stream, and uqk_kernel_times returns the sums of their durations and the
 * number of steps timed.  Turning it on or off resets the sums.  Steps
 * replayed from a captured graph are not timed. */
int uqk_set_kernel_timing(uqk_ctx* ctx, int32_t on);
int uqk_kernel_times(const uqk_ctx* ctx, double* closure_ms, double* advance_ms, int64_t* steps);

/* ---- launch shapes and the autotuner (no counterpart in the reference) --- */
/* Launch shape of the specialised closure kernels: wavefronts (cells in
 * flight) per workgroup of the one-wavefront kernel (1, 2, 4; 0 = as many as
 * LDS allows) and persistent workgroups per CU of either kernel (0 = every
 * resident slot; the workgroup kernel for n = S*M >= 64 takes only this). */
int uqk_set_launch_shape(uqk_ctx* ctx, int32_t waves, int32_t blocks_per_cu);
int uqk_get_launch_shape(const uqk_ctx* ctx, int32_t* waves, int32_t* blocks_per_cu);
/* Times the specialised closure kernel the context runs (the one-wavefront
 * kernel, or the workgroup kernel for n = S*M >= 64) on n cells (device u and
 * starting lambda; lambda is not modified) for every launch shape the LDS and
 * the occupancy allow, keeps the fastest for ctx, and appends it as one JSON
 * line
 * {closure, n_states, n_moments, n_quad, precision, waves, blocks_per_cu,
 * ms, cells} to json_path when not NULL.  Returns the number of shapes timed. */
int uqk_autotune(uqk_ctx* ctx, int64_t n, const double* u, const double* lambda, const char* json_path);
/* Applies the last line of json_path that matches ctx's closure, sizes and
 * precision.  Returns 1 if one matched, 0 if none. */
int uqk_load_tuning(uqk_ctx* ctx, const char* json_path);

/* ---- learned warm start (no counterpart in the reference) --------------- */
/* A network that maps a cell's moments to a starting dual for the Newton
 * closure of an Euler or shallow-water context (UQK_EULER1D, UQK_EULER2D,
 * UQK_SHALLOW_WATER2D, UQK_SHALLOW_WATER2D_BARRIER).  Inputs are the moments
 * scaled by the cell's mean state (Euler: rho0, rho0 sqrt(T0), rho0 T0;
 * shallow water: h0, h0 sqrt(g h0)), outputs the correction to the
 * linearised start (the plain start (DS(u_0), 0, ...) and one Newton step)
 * made dimensionless by the same mean state; four hidden layers of `hidden` units with GELU, BF16 weights on
 * MFMA with FP32 accumulation, trained on the GPU with Adam.  All array
 * arguments are device pointers unless named host_. */
typedef struct uqk_mlp uqk_mlp;
int uqk_mlp_create(uqk_ctx* ctx, int32_t hidden, uint64_t seed, uqk_mlp** out);
int uqk_mlp_destroy(uqk_mlp* net);
/* dims[0..5] = inputs, hidden x 4, outputs (each S*M rounded up to 32);
 * returns the number of layers (5). */
int uqk_mlp_shape(const uqk_mlp* net, int32_t* dims, int64_t* n_params);
/* FP32 parameters [W_1 (out x in, row-major), b_1, ..., W_5, b_5] and the
 * normalisation [mu_x, sd_x, mu_t, sd_t] (host buffers; either may be NULL). */
int uqk_mlp_get_params(uqk_mlp* net, float* host_params, float* host_norm);
int uqk_mlp_set_params(uqk_mlp* net, const float* host_params, const float* host_norm);
/* Per-feature mean and std of inputs and targets from (u, lambda) pairs. */
int uqk_mlp_fit_normalization(uqk_mlp* net, int64_t n_pairs, const double* u, const double* lambda);
/* Features x (n x inputs) and, when lambda is not NULL, targets t
 * (n x outputs), raw or standardised. */
int uqk_mlp_features(uqk_mlp* net, int64_t n, const double* u, const double* lambda, int32_t standardise, float* x,
                     float* t);
/* Network output y (rows x outputs) for standardised features x. */
int uqk_mlp_forward(uqk_mlp* net, int64_t rows, const float* x, float* y);
/* Mean squared error over the first n_real outputs and its gradient (same
 * layout as the parameters) for one batch (rows a multiple of 32). */
int uqk_mlp_loss_grad(uqk_mlp* net, int64_t rows, const float* x, const float* t, int32_t n_real, double* host_loss,
                      float* grad);
/* `steps` Adam steps on minibatches of `batch` pairs (a multiple of 64) drawn
 * with replacement from the n_pairs (u, lambda) pairs; fits the
 * normalisation first if it was never set.  host_mean_loss: mean batch loss. */
int uqk_mlp_train(uqk_mlp* net, int64_t n_pairs, const double* u, const double* lambda, int32_t batch, int32_t steps,
                  double lr, uint64_t seed, double* host_mean_loss);
/* lambda_out (n x M x S) = the network's start for moments u. */
int uqk_mlp_predict(uqk_mlp* net, int64_t n, const double* u, double* lambda_out);
/* Per cell keep in lambda whichever of lambda and lambda_alt has the smaller
 * ||grad L||_2 (non-finite never wins).  Returns the number of cells that
 * took lambda_alt. */
int64_t uqk_warm_start_select(uqk_ctx* ctx, int64_t n, const double* u, double* lambda, const double* lambda_alt);
/* Attach (or detach with NULL) a network: every later closure of ctx first
 * predicts a start and keeps the better one per cell.  Under per-cell orders
 * the moments past a cell's order enter the network as zero, the predicted
 * duals there are zero, and the choice compares ||grad L|| over the cell's
 * own moments; training pairs from such runs are zero-padded the same way. */
int uqk_set_warm_start(uqk_ctx* ctx, uqk_mlp* net);

/* ---- whole runs from an input file ------------------------------------- */
typedef struct {
    int64_t n_cells, steps, newton_steps;
    double t, residual, runtime_s, cell_steps_per_s;
    int32_t n_states, n_moments, n_quad;
} uqk_run_report;

/* Runs the input file end to end on device 0: Settings from the TOML keys of
 * settings.cpp:22-323 (plus mesh `type` line|nacaOGrid|quadGrid, problem `ic`,
 * `flux`, `maxSteps`, moment_system `precision` fp64|mixed), SetupIC and
 * SetDuals (momentsolver.cpp:463-563), the time loop of MomentSolver::Solve
 * (momentsolver.cpp:79-181) and, when mean_var_csv is not NULL, the
 * expectation and variance of every cell (WriteMeanAndVar,
 * momentsolver.cpp:302-328; one row per cell: S means then S variances).
 * Returns 0, or -1 with the message in err. */
int uqk_run_config(const char* toml_path, const char* mean_var_csv, uqk_run_report* report, char* err,
                   int32_t errlen);
/* uqk_run_config that also writes the run's files when output_prefix is not
 * NULL (the end of MomentSolver::Solve, momentsolver.cpp:190-228):
 * <prefix>.vtk, the E/Var fields as cell and point data (Mesh2D::Export,
 * mesh2d.cpp:399-615; 1D: Mesh1D::Export, mesh1d.cpp:90-220, plus
 * <prefix>ExpectedValue and <prefix>Variance); <prefix>, a run log echoing the
 * input (main.cpp:160-180); <prefix>_moments and <prefix>_duals, one line per
 * cell (MomentSolver::Export, momentsolver.cpp:565-588).  An input with
 * [general] restartFile = <prefix> (and importDualState = true for the duals)
 * continues such a run from its previous tEnd (momentsolver.cpp:418-523),
 * with the same moments and quadrature. */
int uqk_run_config_output(const char* toml_path, const char* mean_var_csv, const char* output_prefix,
                          uqk_run_report* report, char* err, int32_t errlen);
/* SU2 mesh reader (Mesh2D::LoadSU2MeshFromFile, mesh2d.cpp:68-255, with the
 * geometry of triangle.cpp / quadrangle.cpp / cell.cpp).  bc_spec types the
 * markers: "tag=type;tag=type" with type dirichlet | noslip | swwall |
 * neumann | periodic (the last two act as untyped boundaries, as in
 * expliciteuler.cpp).  Triangle or quadrilateral meshes (not mixed), element
 * ids 0..NELEM-1 in file order.  Returns the number of cells (>= 0) and the
 * edges per cell in *n_edges; the uqk_mesh arrays (n_cells x n_edges, x 2 for
 * normals) and center (n_cells x 2) are filled where not NULL, so a first call
 * with NULL arrays sizes them.  -1 with the message in err on failure. */
int64_t uqk_su2_mesh(const char* path, const char* bc_spec, int32_t* n_edges, int32_t* neighbor,
                     double* unit_normal, double* normal, double* area, double* dx_param, int32_t* bc,
                     double* boundary_normal, double* center, char* err, int32_t errlen);
/* Host-only part of uqk_run_config: parses the file and builds the mesh and
 * the initial moments without touching a GPU.  info receives (closure, S, M,
 * Q, d, edges per cell, mesh dim, precision); u and lambda (n_cells x M x S)
 * and center (n_cells x dim) are written when not NULL.  Returns n_cells, or
 * -1 with the message in err. */
int64_t uqk_config_setup(const char* toml_path, int32_t* info, double* u, double* lambda, double* center, char* err,
                         int32_t errlen);

#ifdef __cplusplus
}
#endif
#endif
