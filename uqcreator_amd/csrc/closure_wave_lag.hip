// closure_wave_lag.hip -- the mixed one-wavefront Newton closure of a single
// closure call with the reuse of each cell's factorisation within the call
// (closure_wave_kernel.h, PREC 4).
#include "closure_wave_kernel.h"

namespace uqk {

template <int KIND, int M, int Q>
static bool lag_try(int kind, int m, int q, const NewtonArgs& a, int max_lds, int num_cu, hipStream_t st,
                    hipError_t* err, const LaunchShape& shape) {
    if (!wave_fits<KIND, M, Q>(kind, m, q, max_lds)) return false;
    wave_go<KIND, M, Q, 4>(a, max_lds, num_cu, st, err, shape);
    return true;
}

bool newton_wave_lag_launch(int kind, int M, int Q, const NewtonArgs& a, int max_lds, int num_cu, hipStream_t st,
                            hipError_t* err, const LaunchShape& shape) {
    return lag_try<UQK_BOUNDED_BARRIER, 6, 12>(kind, M, Q, a, max_lds, num_cu, st, err, shape) ||
           lag_try<UQK_EULER1D, 15, 49>(kind, M, Q, a, max_lds, num_cu, st, err, shape) ||
           lag_try<UQK_EULER2D, 10, 25>(kind, M, Q, a, max_lds, num_cu, st, err, shape) ||
           lag_try<UQK_SHALLOW_WATER2D, 10, 25>(kind, M, Q, a, max_lds, num_cu, st, err, shape) ||
           lag_try<UQK_SHALLOW_WATER2D_BARRIER, 10, 25>(kind, M, Q, a, max_lds, num_cu, st, err, shape);
}

}  // namespace uqk
