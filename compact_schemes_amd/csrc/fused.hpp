// fused.hpp -- state of the fused right-hand-side path (cs_solver).
#pragma once

#include <type_traits>

#include "physics_kernels.hpp"

namespace cs {
namespace fused {

constexpr int kYLines = 16;   // lines per workgroup of a y/z sweep
constexpr int kMaxP = 32;     // chunks per line

// one axis: derivative operator with balanced chunks
struct AxisDesc {
    const double* tab;   // [5][N]: nu, mu, rho, Mf, Rb
    const double* ftab;  // filter operator (build_line_table layout, 8-point chunks), or null
    int N, P, K, GS;     // GS: lanes per line in the x layout (pow2 >= P)
    long nlines;
    long st;
    int seam;            // both faces seam (periodic wrap on one subdomain)
    int gcbc;            // both faces non-reflective characteristic BCs
    // kernel variant (chosen per axis, see cs_solver_tune):
    int variant;         // 0: direct (ragged when N != P*K); 1: LDS row-block tiles; 2: direct, 16-point chunks
    int tl;              // lines per workgroup (y/z lines of variant 0, all of variant 1): 8 or 16
    int xcd;             // 1: neighbouring line groups go to workgroups on the same XCD
    // filter sweep along this axis: 1 = one field per workgroup (grid.y over
    // the five fields), 5 = all five fields per workgroup; lines per workgroup
    // of the y/z filter sweep (8 or 16)
    int ffields;
    int ftl;
};

// per-point geometry of a curvilinear solver (cs_solver_set_metrics; caller's
// device arrays): m[a][k] = component k of the metric row of axis a (xim, etm,
// zem of the reference, [3][lmx] each), J = yaco [lmx], cm[a] = the face
// normals of axis a, [2 face][3][face] as gcbc_go takes them
struct Metric {
    const double* m[3][3];
    const double* J;
    const double* cm[3];
};

struct SweepCommon {
    int lxi;
    long plane;
    const double* pb;    // pbco[16] | pbci[16]
    const double* sbcc;
};

}  // namespace fused
}  // namespace cs

struct cs_solver {
    cs_solver_config cfg;
    cs::Geo g;
    cs_numerics* num = nullptr;         // interface table, filter operators
    cs::fused::AxisDesc ax[3];     // 8-point chunks up to 256 points: filter sweeps, adjoint sweeps
    // the right-hand-side sweeps: ax, or the same axis cut into 16-point
    // chunks (cs_solver_tune variant 2, with tabs16)
    cs::fused::AxisDesc rax[3];
    // default (variant, lines, xcd) of the right-hand-side sweeps (rax), then of ax
    int shape0[3][6];
    // the filter sweeps: ax, or (y/z, cs_solver_tune_filter) 16-point chunks with ftabs16
    cs::fused::AxisDesc fax[3];
    double* tabs[3] = {nullptr, nullptr, nullptr};
    double* tabs16[3] = {nullptr, nullptr, nullptr};
    double* ftabs16[3] = {nullptr, nullptr, nullptr};
    double* ftabs[3] = {nullptr, nullptr, nullptr};
    int order = 1;             // sweep order (cs_solver_set_order)
    double* ga = nullptr;      // [4][lmx] gradients of the first S1 sweep
    double* gb = nullptr;      // [4][lmx] gradients of the second S1 sweep
    double* tau = nullptr;     // [6][lmx]
    double* hf = nullptr;      // [3][lmx]
    double* dpart = nullptr;   // [5][lmx]
    double* qb = nullptr;      // [5][lmx] stage buffers
    double* qc = nullptr;
    double* red = nullptr;     // time-step partial maxima
    int red_blocks = 0;
    // time-step maxima per workgroup of the gradient sweep that reduces them
    // inside an RK4 step (dt_fold = its position + 1, 0 = a pass of its own)
    double* dtpart = nullptr;
    int dt_fold = 0;
    // viscous solvers: the second gradient sweep of each RHS runs on aux,
    // concurrently with the first (both read q and write their own gradients)
    hipStream_t aux = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    double* dt = nullptr;      // [0] dt, [1..2] maxima
    double* dt_host = nullptr; // pinned
    double* vjp = nullptr;     // [41][lmx] adjoint scratch (allocated on the first adjoint call)
    // adjoint sweeps (tape.hip, set up on the first call that needs them)
    double* ttabs[3] = {nullptr, nullptr, nullptr};   // LU of A^T of the derivative, 8-point chunks
    double* fttabs[3] = {nullptr, nullptr, nullptr};  // LU of A^T of the closure-line filter
    double* adj = nullptr;     // seed, accumulator (5 fields each), face buffers, internal tape
    double* rec = nullptr;     // faces of the RHS being evaluated go here when set ([3 axes][2][5][face])
    int adj_mode = 1;          // 1: fused transposed sweeps where the lines allow; 0: per-operator products
    int prec = 0;              // 1: the viscous intermediates (gradients, stress, heat flux) stored in FP32
    int curv = 0;              // 1: per-point metrics, Jacobian and face normals (met)
    cs::fused::Metric met{};
};

namespace cs {
namespace fused {
// axes of the S1 sweeps (the third one closes) and of the S2 sweeps (the
// third one finishes the RHS) for each sweep order
constexpr int kS1Axes[2][3] = {{0, 1, 2}, {1, 2, 0}};
constexpr int kS2Axes[2][3] = {{0, 1, 2}, {2, 1, 0}};

int rhs_sweeps(cs_solver* sv, const double* q, const double* sbcc, double* out, const double* q0, int div,
               hipStream_t s);
// sweep `pos` of a viscous RHS with the intermediates in FP32 (fused_f32.hip)
int sweep_at_f32(cs_solver* sv, int pos, const SweepCommon& cm, const double* q, double* out, const double* q0,
                 int div, hipStream_t s);
// sweep `pos` of a curvilinear solver (fused_curv.hip)
int sweep_at_curv(cs_solver* sv, int pos, const SweepCommon& cm, const double* q, double* out, const double* q0,
                  int div, hipStream_t s);
int time_step(cs_solver* sv, const double* q, double dt_fixed, hipStream_t s);
// the position of the gradient sweep that can reduce the time step of an RK4
// step (viscous solvers in the default order: 0 = the y sweep on the direct
// kernel with whole waves, constant metrics; 2 = the closing x sweep on the
// transposing kernel, per-point metrics), -1 = none (a pass of its own)
int dt_fold_pos(const cs_solver* sv);
int one_sweep(cs_solver* sv, int which, const double* q, const double* sbcc, double* out, hipStream_t s);
int build_axis_table(int n, int seam, int K, int P, std::vector<double>& h);
// every axis filters with the 8-point fused sweep of ax (the adjoint's assumption)
bool filter_fusable(const cs_solver* sv);
// the filter along axis a runs a fused sweep (fax), else the line solver
bool filter_axis_fused(const cs_solver* sv, int a);
int filter_sweeps(cs_solver* sv, double* q, hipStream_t s, int axis = -1);
int step_filter(cs_solver* sv, double* qa, hipStream_t s);
// adjoint (vjp.hip: per-operator products; tape.hip: fused transposed sweeps)
bool adjoint_fused(const cs_solver* sv);
long tape_doubles(const cs_solver* sv);
int rk4_vjp_lines(cs_solver* sv, const double* q0, const double* Q, double* lam, double dt, const double* sbcc,
                  hipStream_t s);
int rk4_step_tape(cs_solver* sv, double* qa, const double* sbcc, double dt, double* tape, hipStream_t s);
int rk4_vjp_tape(cs_solver* sv, const double* tape, double* lam, double dt, const double* sbcc, hipStream_t s);
int rk4_vjp_fused(cs_solver* sv, const double* q0, double* lam, double dt, const double* sbcc, hipStream_t s);
int rhs_vjp_fused(cs_solver* sv, const double* q, const double* lam, double* g, const double* sbcc, hipStream_t s);
}  // namespace fused
}  // namespace cs
