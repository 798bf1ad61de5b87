// fused_curv.hip -- the sweeps of the fused right-hand side on a curvilinear
// grid (cs_solver_set_metrics): every sweep reads the metric row of its axis
// (and the gradient closure all nine metric components and the Jacobian) at
// each point instead of the solver's constant diagonal metrics, the flux
// sweeps take the whole stress tensor and heat flux, the characteristic BCs
// take the caller's face normals and the point's Jacobian; with
// cs_solver_set_precision 1 the viscous intermediates are stored in FP32 as
// on Cartesian grids.  A separate
// translation unit so that these kernels compile in parallel with the others.
#include "fused_kernels.hpp"

namespace cs {
namespace fused {

int sweep_at_curv(cs_solver* sv, int pos, const SweepCommon& cm, const double* q, double* out, const double* q0,
                  int div, hipStream_t s) {
    if (sv->cfg.viscous && sv->prec == 1) {
        // the viscous intermediates stored in FP32 (cs_solver_set_precision 1)
        if (pos < 3) return s1_sweep<S1CF>(sv, pos, cm, q, s);
        return s2_sweep<true, S2CF>(sv, pos - 3, cm, q, out, q0, div, s);
    }
    if (pos < 3) return s1_sweep<S1C>(sv, pos, cm, q, s);
    if (sv->cfg.viscous) return s2_sweep<true, S2C>(sv, pos - 3, cm, q, out, q0, div, s);
    return s2_sweep<false, S2C>(sv, pos - 3, cm, q, out, q0, div, s);
}

}  // namespace fused
}  // namespace cs
