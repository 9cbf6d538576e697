// uqk_internal.h -- kernel argument blocks and launchers shared between the
// kernel translation units and the C-ABI layer (uqk_api.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "uqk.h"
#include "uqk_device.h"

namespace uqk {

struct NewtonArgs {
    ClosureParams c;
    int M;                       // table width (nTotal of the finest level)
    int Q;                       // quadrature points
    const double* phi;           // Q x M, device
    const double* phiw;          // Q x M, device
    long ncell;
    const int* nt_cell;          // per-cell nTotal or nullptr
    const double* u;             // ncell x S x M
    double* lam;                 // ncell x S x M, in/out
    int* status;                 // ncell or nullptr
    int* iters;                  // ncell or nullptr
    unsigned long long* counter; // work-queue head, zero at launch
    unsigned long long* nfail;   // [0] failed cells, [1] Newton steps, [2] gradient evaluations; zero at launch
    int wave_doubles;            // LDS doubles per wave
    int prec;                    // UQK_FP64 or UQK_MIXED (Hessian arithmetic)
    const int* active;           // device flag: 0 skips the launch (graph replays past t_end)
    double* uq_out;              // wave kernel, when not null: U at the quadrature points of the
    double* u_out;               // returned lambda and its moments <U phiW> (fused uQ + reconstruction)
    long waves;                  // resident waves of the launch (work-queue chunking), set by the launcher
    int qbits;                   // work queue of 2^qbits positions >= ncell, set by the launcher
    // with uq_out, when dt_min is not null: also min over cells and points of
    // Problem::ComputeDt(uQ, dx_param) as the bits of a positive double,
    // combined with atomicMin (the caller sets *dt_min = 1e10 first)
    ProblemParams dtp;
    const double* dx_param;
    unsigned long long* dt_min;
    // learned warm start, wave kernel: when lam_alt is not null, a cell whose
    // lambda fails the first convergence test also evaluates ||grad L|| at
    // lam_alt and starts Newton from whichever is smaller (non-finite never
    // wins); *took += cells that started from lam_alt
    const double* lam_alt;
    unsigned long long* took;
    // residual check before the Newton kernel (time loop): the check kernel
    // finishes every cell that passes ||grad L|| < eps at its lambda and
    // appends the others to list_out[(*list_count)++]; the wave kernel then
    // takes its cells from cell_list[0 .. *cell_count) (null: every cell)
    int* list_out;
    unsigned long long* list_count;
    const int* cell_list;
    const unsigned long long* cell_count;
    // and, when not null, the check kernel stores the gradient at lambda of
    // every cell it lists (cell-major, entry (i, s) at i * S + s), which the
    // Newton kernel takes in place of evaluating it again
    double* grad_at_lam;
    // time loop, wave kernel shapes: each cell's ComputeDt (the fused
    // kernels store it), and the run's flag that the last step's outputs
    // (uQ, u_out, dt_cell) are those of the current lambda -- then the
    // residual check needs no evaluation of U: the gradient at lambda is
    // u_out - u (u_out = <U(lambda phi) phiW>), and a cell that passes keeps
    // its uQ, moments and ComputeDt
    double* dt_cell;
    const int* recon_ok;
    // mixed precision, wave kernel: reuse of the factorisation.
    // pinv_mode 1 (time loop, across steps): a cell whose pinv_ok is 1 holds
    // a kept inverse Hessian in pinv (FP32, n x n, row stride n + 1) and
    // takes Newton steps d = P g refined once with the FP64 Hessian-vector
    // product; a factored step stores its factor in ufac (packed upper U,
    // then 1/U_rr, FP64), sets pinv_ok to 2 and appends the cell to
    // plist[(*plist_count)++], and factor_inverse_kernel turns the listed
    // factors into the next inverses.  pinv_mode 2 (one closure call): a
    // cell's later Newton steps solve with the factor of its last Cholesky
    // while they contract.  nfactor[0] += Cholesky factorisations,
    // nfactor[1] += accepted steps on a kept factor or inverse.
    float* pinv;
    double* ufac;
    int* pinv_ok;
    int* plist;
    unsigned long long* plist_count;
    int pinv_mode;
    long pinv_slots;
    unsigned long long* nfactor;
};

// Device-side state of a multi-step run replayed from a captured graph.  The
// host writes t, t_end, max_steps and the first step's active flag; each
// step's step_end_kernel sets the flag for the next step, so the captured
// launches do not depend on the run's length.
struct RunState {
    double t;
    double residual;
    long long steps;
    long long newton;
    long long gradients;
    long long nfail;
    long long factorizations;
    long long reuse_steps;
    double t_end;
    long long max_steps;  // < 0: no limit
    int active;
    // 1 once a step's closure has left uQ, the reconstructed moments and
    // every cell's ComputeDt at the current lambda (zero at the run's start):
    // the next step's residual check then takes the gradient from them
    int recon;
};

// Closure counters of a launch: [0] work-queue head, [1] failed cells, [2]
// Newton steps, [3] gradient evaluations, [4] cells that started from the
// learned warm start, [5] Cholesky factorisations, [6] steps on a kept
// inverse, [7] cells listed for factor_inverse_kernel, [8] cells the
// residual check left to the Newton kernel; all zero at launch.
constexpr int kClosureCounters = 9;

// Launch shape of the specialised closure kernel; 0 = automatic (as many
// waves per workgroup as LDS allows, every resident workgroup slot filled).
struct LaunchShape {
    int waves = 0;          // wavefronts (cells in flight) per workgroup: 1, 2 or 4
    int blocks_per_cu = 0;  // persistent workgroups per CU (capped by occupancy)
};

// Work-queue parameters of a closure launch with `resident_waves` waves.
inline void set_queue(NewtonArgs& a, long resident_waves) {
    a.waves = resident_waves > 0 ? resident_waves : 1;
    int bits = 0;
    while ((1L << bits) < a.ncell) ++bits;
    a.qbits = bits;
}

bool newton_wave_launch(int kind, int M, int Q, const NewtonArgs& a, int max_lds, int num_cu, hipStream_t st,
                        hipError_t* err, const LaunchShape& shape = LaunchShape());
// Residual check of the wave kernel's shapes (closure_check.hip): a.list_out
// and a.list_count receive the cells left to the Newton kernel; false when
// the shape has no check kernel.
bool closure_check_launch(int kind, int M, int Q, const NewtonArgs& a, int num_cu, hipStream_t st, hipError_t* err);
// The same check for the workgroup kernel (n = S*M <= 128, per-cell orders):
// one wave per cell (closure_check.hip).
bool closure_check_wave_fits(int S, int M, int Q, int max_lds);
hipError_t closure_check_wave_launch(int kind, const NewtonArgs& a, int max_lds, int num_cu, hipStream_t st);
// Whether (kind, M, Q) has the mixed wave kernel with kept inverses.
bool newton_wave_reuse_fits(int kind, int M, int Q, int max_lds);
// Inverses of the factors the last mixed wave launch stored and listed
// (list[0 .. *count), flags 2 -> 1), up to ncell of them; false when there
// is no such kernel for (kind, M, Q).
bool factor_inverse_launch(int kind, int M, int Q, long ncell, const unsigned long long* count, const int* list,
                           int* flags, const double* ufac, float* pinv, const int* active, int num_cu,
                           hipStream_t st, hipError_t* err);
// Largest waves per workgroup the LDS allows and the resident workgroups per
// CU at `waves` for the specialised kernel of (kind, M, Q, prec); false when
// there is none.
bool newton_wave_limits(int kind, int M, int Q, int prec, int max_lds, int waves, int* max_waves, int* occupancy);
int newton_fp64_wave_doubles(int S, int M, int Q);
// workgroup-per-cell kernel for n = S*M up to 256 (closure_block.hip)
bool newton_block_fits(int S, int M, int Q, int max_lds);
// shape.blocks_per_cu > 0 caps the persistent workgroups per CU
hipError_t newton_block_launch(int kind, const NewtonArgs& a, int max_lds, int num_cu, hipStream_t st,
                               const LaunchShape& shape = LaunchShape());
// resident workgroups per CU of the workgroup kernel for (kind, M, Q, prec)
int newton_block_occupancy(int kind, int M, int Q, int prec);
int newton_fp64_occupancy(int kind, bool tlds, int threads, size_t shmem);
hipError_t newton_fp64_launch(int kind, const NewtonArgs& a, bool tlds, int blocks, int threads, size_t shmem,
                              hipStream_t st);

struct SweepArgs {
    ClosureParams c;
    ProblemParams p;
    int S, M, Q, d;
    const double* phi;
    const double* phiw;
    const double* xi;            // Q x d
    long ncell;
    const int* nt_cell;
    // mesh
    int nedge, dim;
    const int* nbr;
    const double* unit_normal;
    const double* normal;
    const double* area;
    const double* dx_param;
    const int* bc;
    const double* bnormal;
    // data
    const double* lam;
    const double* u;
    const double* uq;
    double* uq_out;
    double* u_out;
    double* mv_out;
    double dt;
    const double* dt_dev;        // if non-null, dt is read from here
    double* partial;             // per-block partial results
    int npartial;
    double* scalar_out;          // final reduction result
    const int* active;           // device flag: 0 skips the launch
    int table_lds;               // advance: PhiW staged in LDS (set by the launcher)
};

// Refinement step of the adaptive moment solver: per cell the level whose
// nTotal equals nt_cell[j], its indicator from the moments u, one level up or
// down, nt_cell[j] rewritten.
constexpr int kMaxLevels = 16;
struct RefineArgs {
    long ncell;
    int S, M, dim_xi;
    int nlev;
    int max_level;             // cap on the level (refinement retardation), nlev - 1 without
    int nt_level[kMaxLevels];  // nTotal per level, ascending
    double refine_above, coarsen_below;
    const double* u;  // moments, (s, i) of cell j at j*S*M + i*S + s
    int* nt_cell;
    const int* active;  // device flag: 0 skips the launch
};

hipError_t launch_eval_uq(int kind, const SweepArgs& a, hipStream_t st);
hipError_t launch_refine(const RefineArgs& a, hipStream_t st);
hipError_t launch_reconstruct(const SweepArgs& a, hipStream_t st);
hipError_t launch_dt(int kind, const SweepArgs& a, hipStream_t st);
hipError_t launch_advance(int kind, const SweepArgs& a, hipStream_t st);
hipError_t launch_residual(const SweepArgs& a, hipStream_t st);
hipError_t launch_mean_var(int kind, const SweepArgs& a, hipStream_t st);
hipError_t launch_fill_u64(unsigned long long* p, int n, hipStream_t st);
hipError_t launch_fill_f64(double* p, int n, double v, hipStream_t st);
int sweep_partial_blocks(long ncell);
// End of a replayed step (momentsolver.cpp:134-152): residual = sum of the
// advance's per-block partials, t += dt, the step's closure counts into the
// run totals; then clears the closure counters, sets the fused dt's start value
// and the next step's active flag.
hipError_t launch_step_end(RunState* rs, double* scalars, unsigned long long* counters, const double* partial,
                           int set_recon, int npartial, hipStream_t st);

}  // namespace uqk
