// fused_f32.hip -- the viscous sweeps of the fused right-hand side with the
// intermediates between sweeps stored in FP32 (cs_solver_set_precision 1):
// the velocity and c^2 gradients of the first two gradient sweeps, the stress
// tensor and the heat flux.  The conserved variables, the partial RHS, the
// closure and seam rows, the characteristic BCs and the RK accumulation stay
// FP64, and every sweep computes in FP64.  A separate translation unit so
// that the two sets of kernels compile in parallel.
#include "fused_kernels.hpp"

namespace cs {
namespace fused {

int sweep_at_f32(cs_solver* sv, int pos, const SweepCommon& cm, const double* q, double* out, const double* q0,
                 int div, hipStream_t s) {
    if (pos < 3) return s1_sweep<S1F>(sv, pos, cm, q, s);
    return s2_sweep<true, S2F>(sv, pos - 3, cm, q, out, q0, div, s);
}

}  // namespace fused
}  // namespace cs
