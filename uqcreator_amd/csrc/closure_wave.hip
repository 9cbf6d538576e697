// closure_wave.hip -- instantiations and dispatch of the one-wavefront
// Newton closure (closure_wave_kernel.h): FP64 (PREC 0) and mixed (PREC 1);
// the mixed closure with kept inverses (PREC 2) and the kernel that turns a
// time step's new Cholesky factors into those inverses live in
// closure_wave_reuse.hip, the mixed closure reusing factors within one call
// (PREC 4) in closure_wave_lag.hip.
#include "closure_wave_kernel.h"

namespace uqk {

// PREC 2 launches (closure_wave_reuse.hip) and PREC 4 launches
// (closure_wave_lag.hip); false when the shape has none
bool newton_wave_reuse_launch(int kind, int M, int Q, const NewtonArgs& a, int max_lds, int num_cu, hipStream_t st,
                              hipError_t* err, const LaunchShape& shape);
bool newton_wave_lag_launch(int kind, int M, int Q, const NewtonArgs& a, int max_lds, int num_cu, hipStream_t st,
                            hipError_t* err, const LaunchShape& shape);

template <int KIND, int M, int Q>
static bool wave_try(int kind, int m, int q, const NewtonArgs& a, int max_lds, int num_cu, hipStream_t st,
                     hipError_t* err, const LaunchShape& shape) {
    if (!wave_fits<KIND, M, Q>(kind, m, q, max_lds)) return false;
    if (a.prec == UQK_FP64) {
        wave_go<KIND, M, Q, 0>(a, max_lds, num_cu, st, err, shape);
        return true;
    }
    // the time loop's mixed closure with kept inverses
    if (a.pinv_mode == 1 && newton_wave_reuse_launch(kind, m, q, a, max_lds, num_cu, st, err, shape)) return true;
    // one closure call's mixed closure reusing each cell's factor within the call
    if (a.pinv_mode == 2 && newton_wave_lag_launch(kind, m, q, a, max_lds, num_cu, st, err, shape)) return true;
    wave_go<KIND, M, Q, 1>(a, max_lds, num_cu, st, err, shape);
    return true;
}

template <int KIND, int M, int Q>
static bool wave_limits(int kind, int m, int q, int prec, int max_lds, int waves, int* max_waves, int* occupancy) {
    if (!wave_fits<KIND, M, Q>(kind, m, q, max_lds)) return false;
    if (prec == UQK_FP64) {
        *max_waves = wave_max_waves<KIND, M, Q, 0>(max_lds);
        *occupancy = wave_occupancy<KIND, M, Q, 0>(waves < *max_waves ? waves : *max_waves);
    } else {
        *max_waves = wave_max_waves<KIND, M, Q, 1>(max_lds);
        *occupancy = wave_occupancy<KIND, M, Q, 1>(waves < *max_waves ? waves : *max_waves);
    }
    return true;
}

bool newton_wave_launch(int kind, int M, int Q, const NewtonArgs& a, int max_lds, int num_cu, hipStream_t st,
                        hipError_t* err, const LaunchShape& shape) {
    if (a.nt_cell) return false;  // per-cell orders use the generic kernel
    return wave_try<UQK_BOUNDED_BARRIER, 6, 12>(kind, M, Q, a, max_lds, num_cu, st, err, shape) ||
           wave_try<UQK_EULER1D, 15, 49>(kind, M, Q, a, max_lds, num_cu, st, err, shape) ||
           wave_try<UQK_EULER2D, 10, 25>(kind, M, Q, a, max_lds, num_cu, st, err, shape) ||
           wave_try<UQK_SHALLOW_WATER2D, 10, 25>(kind, M, Q, a, max_lds, num_cu, st, err, shape) ||
           wave_try<UQK_SHALLOW_WATER2D_BARRIER, 10, 25>(kind, M, Q, a, max_lds, num_cu, st, err, shape);
}

bool newton_wave_limits(int kind, int M, int Q, int prec, int max_lds, int waves, int* max_waves, int* occupancy) {
    return wave_limits<UQK_BOUNDED_BARRIER, 6, 12>(kind, M, Q, prec, max_lds, waves, max_waves, occupancy) ||
           wave_limits<UQK_EULER1D, 15, 49>(kind, M, Q, prec, max_lds, waves, max_waves, occupancy) ||
           wave_limits<UQK_EULER2D, 10, 25>(kind, M, Q, prec, max_lds, waves, max_waves, occupancy) ||
           wave_limits<UQK_SHALLOW_WATER2D, 10, 25>(kind, M, Q, prec, max_lds, waves, max_waves, occupancy) ||
           wave_limits<UQK_SHALLOW_WATER2D_BARRIER, 10, 25>(kind, M, Q, prec, max_lds, waves, max_waves, occupancy);
}

}  // namespace uqk
