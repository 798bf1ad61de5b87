/*
 * compact_schemes.h -- C-ABI of the MI355X compact-scheme hot path.
 *
 * Every entry point replaces one operator of the reference's L5 API
 * (EnricoDeg/compact_schemes, paths relative to that repository); the cited
 * file:line is the interface it stands in for.  All field pointers are device
 * pointers to FP64 data in the reference layout: SoA, variable-major, x
 * fastest, idx = (k*let + j)*lxi + i, lmx = lxi*let*lze.
 *
 * Conventions
 *   - every call returns CS_OK (0) or a CS_ERR_* code; cs_last_error() gives
 *     the message of the last failing call on the calling thread;
 *   - launches are asynchronous on the given stream (a hipStream_t, NULL for
 *     the null stream); only cs_calc_time_step synchronises (it returns a
 *     host value, as physics::calc_time_step does);
 *   - contexts are thread-compatible, not thread-safe;
 *   - halo exchange does not exist on one GPU: faces with a seam code
 *     (BC_INTER_STRAIGHT 35, BC_INTER_SUBDOMAINS 40, BC_PERIODIC 45) exchange
 *     with the opposite face of the same subdomain, exactly what the reference
 *     does when the neighbour rank is the rank itself.
 */
#ifndef COMPACT_SCHEMES_H
#define COMPACT_SCHEMES_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CS_OK              0
#define CS_ERR_ARG         1   /* null pointer, bad axis/variable/BC code   */
#define CS_ERR_SHAPE       2   /* line length or grid shape not supported   */
#define CS_ERR_HIP         3   /* a HIP runtime call failed                 */
#define CS_ERR_UNSUPPORTED 4   /* needs something one GPU does not have     */
#define CS_ERR_STATE       5   /* call order (e.g. no deriv_setup yet)      */

/* boundary-condition codes, src/common/parameters.hpp:49-55 */
#define CS_BC_NON_REFLECTIVE   10
#define CS_BC_WALL_INVISCID    20
#define CS_BC_WALL_VISCOUS     25
#define CS_BC_INTER_CURV       30
#define CS_BC_INTER_STRAIGHT   35
#define CS_BC_INTER_SUBDOMAINS 40
#define CS_BC_PERIODIC         45

typedef void* cs_stream;                  /* hipStream_t                        */
typedef struct { int lxi, let, lze; } cs_dims;   /* t_dcomp, data_types.hpp:33-39 */
typedef struct { double x, y, z; } cs_point;     /* t_point, data_types.hpp:41-47 */

typedef struct cs_numerics cs_numerics;   /* numerics_base / numerics_pc state  */
typedef struct cs_physics  cs_physics;    /* physics_base / physics state       */

const char* cs_last_error(void);
const char* cs_version(void);

/* ---- numerics ---------------------------------------------------------- */

/* numerics_base(t_dcomp, int nbc[2][3]), src/numerics_base.hpp:43-94.
 * nbc is [2 faces][3 axes] row-major.  Allocates drva, send, recv, pbco, pbci
 * on the current device. */
int cs_numerics_create(const cs_dims* dims, const int* nbc, cs_numerics** out);
/* ~numerics_base, src/numerics_base.hpp:96-111 */
int cs_numerics_destroy(cs_numerics* num);
/* numerics_base::deriv_setup, src/numerics_base.hpp:182-297: the as-shipped
 * quasi-disjoint interface table (computed in FP64). */
int cs_deriv_setup(cs_numerics* num);
/* replaces the pbco/pbci table (e.g. the D1-corrected one); no reference
 * counterpart -- the reference only has deriv_setup. */
int cs_set_interface_table(cs_numerics* num, const double* pbco16, const double* pbci16);
int cs_get_interface_table(const cs_numerics* num, double* pbco16, double* pbci16);
/* numerics_base members drva_buffer/send_buffer/recv_buffer,
 * src/numerics_base.hpp:299-305: device pointers of axis `axis`.
 * drva = [2 face][5 var][face], send/recv = [5 var][2 face][2 row][face]. */
int cs_numerics_buffers(cs_numerics* num, int axis, double** drva, double** send, double** recv);
/* numerics_base::ndf, src/numerics_base.hpp:73-93: ndf[face*3+axis]. */
int cs_numerics_ndf(const cs_numerics* num, int* ndf6);

/* numerics_pc::fill_buffers<lo,hi>, src/numerics_pc.hpp:231-287: fills send
 * for variables lo..hi-1 of a [5][3][lmx] field; component nzk=(1-nrt)*axis. */
int cs_fill_buffers(cs_numerics* num, const double* infield, int var_lo, int var_hi, int nrt,
                    cs_stream stream);
/* numerics_base::mpigo (2D), src/numerics_base.hpp:114-180, one subdomain:
 * recv[var][face] = send[var][1-face] on every seam face. */
int cs_mpigo(cs_numerics* num, int var, cs_stream stream);
/* numerics_base::mpigo between two subdomains on the same GPU (the neighbour
 * rank mcd[face][axis] of the reference, src/numerics_base.hpp:114-180, is
 * another context here): for the adjoining faces face_a of `a` and face_b of
 * `b` along `axis`, recv_b[var][face_b] = send_a[var][face_a] and
 * recv_a[var][face_a] = send_b[var][face_b] (device copies, asynchronous on
 * `stream`).  Both faces must be seam faces with the same points; a == b with
 * face_a != face_b is the periodic self-wrap of cs_mpigo. */
int cs_exchange(cs_numerics* a, int face_a, cs_numerics* b, int face_b, int axis, int var, cs_stream stream);
/* numerics_pc::deriv1d<Axis>, src/numerics_pc.hpp:74-119: out = d(in)/d(xi_axis),
 * closure rows where nstart/nend == 0, seam rows (reading recv) where 1. */
int cs_deriv1d(cs_numerics* num, int axis, const double* in, double* out, int nstart, int nend,
               int var, cs_stream stream);
/* numerics_pc::deriv2d<Axis>, src/numerics_pc.hpp:121-168: reads component
 * `comp` of in ([3][lmx] per variable), writes component `axis` of out. */
int cs_deriv2d(cs_numerics* num, int axis, const double* in, double* out, int nstart, int nend,
               int var, int comp, cs_stream stream);
/* grid metrics from point coordinates (no reference counterpart: the
 * reference's grid allocates xim/etm/zem and never fills them,
 * src/grid.hpp:66-72): xyz = [3][lmx] coordinates x, y, z; period9 = [axis][c]
 * the coordinate jump over one period of each seam axis (NULL = none).
 * Writes xim, etm, zem ([3][lmx] each: grad(xi)/J, grad(eta)/J, grad(zeta)/J
 * as cofactors of d x_c / d xi_a, each from compact derivatives with the
 * context's BCs) and yaco = J = 1 / det(d x_c / d xi_a). */
int cs_metrics(cs_numerics* num, const double* xyz, const double* period9, double* xim, double* etm, double* zem,
               double* yaco, cs_stream stream);
/* fill + mpigo + deriv1d in one launch: the seam sums are formed from the
 * opposite end of each line inside the kernel (no send/recv traffic).  Same
 * result as cs_fill_buffers + cs_mpigo + cs_deriv1d with the context's ndf. */
int cs_deriv(cs_numerics* num, int axis, const double* in, double* out, int var, cs_stream stream);
/* compact filter along `axis` (absent upstream; 10th-order tridiagonal
 * filter, alpha_f from cs_set_filter_alpha, default 0.49).  Cyclic when both
 * faces of the axis are seam codes; otherwise the end points are left as they
 * are and the rows at distance d = 1..4 from an end run the order-2d centred
 * filter with alpha = min(alpha_f, 0.2, 0.3, 0.4, 0.45 for d = 1..4).
 * in == out is allowed. */
int cs_filter(cs_numerics* num, int axis, const double* in, double* out, cs_stream stream);
int cs_set_filter_alpha(cs_numerics* num, double alpha_f);
/* arithmetic of the line solves of this context's derivatives and filters
 * (no reference counterpart; the reference templates everything on one
 * Type): 0 = FP64 (default); k = 1..4: the solve in FP32 followed by k
 * refinement sweeps -- FP64 residual of the FP64 system, FP32 solve of it,
 * FP64 update.  Right-hand sides, boundary rows and results stay FP64. */
int cs_set_precision(cs_numerics* num, int mode);
/* transposed line operators (no reference counterpart): out = D^T in and
 * out = F^T in for the derivative D = A^-1 B (cs_deriv1d/cs_deriv with this
 * context's closure or seam rows) and the filter F of cs_filter along
 * `axis`: a solve with A^T, then the transposed explicit stencil B^T.  Lines
 * of at least 16 points; out != in; the first call allocates one field of
 * scratch in the context. */
int cs_deriv_adjoint(cs_numerics* num, int axis, const double* in, double* out, cs_stream stream);
int cs_filter_adjoint(cs_numerics* num, int axis, const double* in, double* out, cs_stream stream);

/* ---- physics ----------------------------------------------------------- */

/* physics_base<Visc,T>(t_dcomp), src/physics_base.hpp:45-130, with umf set as
 * read_config does (src/physics_base.hpp:135-145) and dudtmf = 0. */
int cs_physics_create(const cs_dims* dims, int viscous, cs_point umf, cs_physics** out);
int cs_physics_destroy(cs_physics* phys);
int cs_physics_set_frame(cs_physics* phys, cs_point umf, cs_point dudtmf);
/* stress = [6][lmx] (xx,yy,zz,xy,yz,zx), heat = [3][lmx] (xx,yy,zz):
 * the device storage behind physics_base::stress_tensor / heat_fluxes,
 * src/physics_base.hpp:57-86. */
int cs_physics_buffers(cs_physics* phys, double** stress, double** heat, double** buffer,
                       double** buffer_deriv);

/* primitive variables and pressure (absent upstream; the layout the
 * reference kernels read, src/cuda/kernels/physics.hpp:106-108,395):
 * de0 = 1/rho, de1..3 = u,v,w, de4 = c^2 = gam*p/rho, p = (gam-1)(E-rho|u|^2/2). */
int cs_primitives(const cs_dims* dims, const double* qa, double* de, double* pressure,
                  cs_stream stream);
/* physics::calc_fluxes, src/physics_pc.hpp:83-149: de = div of the fluxes. */
int cs_calc_fluxes(cs_physics* phys, cs_numerics* num, const double* qa, const double* pressure,
                   double* de, const double* xim, const double* etm, const double* zem,
                   cs_stream stream);
/* physics::calc_viscous_shear_stress, src/physics_pc.hpp:151-227 (D6-D8
 * corrected: seam fill from component 0, axis-matched BC rows, chain rule on
 * the derivatives). */
int cs_calc_viscous_shear_stress(cs_physics* phys, cs_numerics* num, double* de, const double* ssk,
                                 const double* xim, const double* etm, const double* zem,
                                 const double* yaco, cs_stream stream);
/* physics::calc_time_step, src/physics_pc.hpp:229-256 (D9 corrected: FP64
 * max, zeroed accumulators).  max2 (host, may be NULL) receives the two
 * maxima.  Synchronises `stream`. */
int cs_calc_time_step(cs_physics* phys, const double* xim, const double* etm, const double* zem,
                      const double* de, const double* yaco, const double* ssk, double cfl,
                      double* dte, double* max2, cs_stream stream);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory or wait on the host (code under HIP graph capture, tight serving
 * loops): the time step of cs_calc_time_step without a physics context,
 * without any allocation and without synchronising.  out3 (device, 3
 * doubles) receives dte = min(cfl / max0, 0.5 / max1), max0, max1 -- the
 * values cs_calc_time_step returns; dte is not checked (it is +inf when both
 * maxima are 0).  Partial maxima go to a static per-GPU scratch array of
 * 2*2048+2 doubles (32.8 KB, held until the process ends); any grid fits,
 * since the reduction strides over the points with at most 2048 blocks, and
 * a larger demand is rejected with CS_ERR_SHAPE.  Every call on one GPU
 * shares that scratch: calls that overlap on one GPU (different streams or
 * host threads) race. */
int cs_calc_time_step_noalloc(const cs_dims* dims, cs_point umf, const double* xim, const double* etm,
                              const double* zem, const double* de, const double* yaco, const double* ssk,
                              double cfl, double* out3, cs_stream stream);
/* gcbc_go, src/gcbc.hpp:86-186 (D10-D13 corrected).  cm[a] = [2][3][face]
 * unit face normals; drvb[a] = neighbour characteristics, only read on
 * BC_INTER_CURV faces (the exchange itself is the caller's). */
int cs_gcbc_go(cs_numerics* num, const double* const* cm, const double* const* drvb,
               const double* qa, double* de, const double* pressure, const double* yaco,
               const double* sbcc, cs_point umf, cs_point dudtmf, double dtwi, cs_stream stream);
/* wall_condition_update, src/gcbc.hpp:188-228. */
int cs_wall_condition_update(cs_numerics* num, double* qa, const int* npex, cs_point umf,
                             cs_stream stream);
/* initial state (no reference counterpart: the driver's physics init,
 * driver/canard.cu:129, is not defined upstream): uniform flow carried by the
 * frame velocity umf (stored velocity 0), rho0 = 1, p0 = 1/gam (c0 = 1), plus
 * a Gaussian pressure pulse p' = amp exp(-ln2 r^2 / width^2) around `center`
 * (index units) with rho' = p' / c0^2: qa = (rho, 0, 0, 0, p / (gam - 1)). */
int cs_init_pulse(const cs_dims* dims, cs_point center, double amp, double width, double* qa, cs_stream stream);
/* characteristic-update weights sbcc (allocated and never set upstream,
 * src/gcbc.hpp:51-67): 1 on every boundary point of the subdomain, 0 inside. */
int cs_boundary_weights(const cs_dims* dims, double* sbcc, cs_stream stream);
/* RK stage update (absent upstream; canard.cu:196-200 names the weights):
 * q = q0 - dtk * yaco * de for the 5 conserved variables. */
int cs_rk_stage(const cs_dims* dims, double* q, const double* q0, const double* de,
                const double* yaco, double dtk, cs_stream stream);

/* ---- fused right-hand side and RK4 step (one call per step) -------------
 * The reference evaluates the RHS as calc_viscous_shear_stress + calc_fluxes
 * + gcbc_go (src/physics_pc.hpp:83-227, src/gcbc.hpp:86-186) and leaves the
 * stage update as a stub (driver/canard.cu:132-205, physics_base::movef).
 * cs_solver computes the same quantities with six fused line sweeps per RHS.
 * By default the grid is uniform and Cartesian: the metrics are the constant diagonals
 * xim = (metric[0],0,0), etm = (0,metric[1],0), zem = (0,0,metric[2]), yaco
 * and ssk (= mu) are constants, cm = sign(yaco * metric[a]) e_a (the unit
 * normal of grad xi_a = yaco * metric[a] e_a; it is -e_a on an axis whose
 * coordinate decreases with the index, as z of src/grid.hpp:103);
 * cs_solver_set_metrics makes them per-point arrays (curvilinear grids).
 * Per axis both faces must be
 * non-reflective (10) or both seam codes (35/40/45); walls and curved
 * interfaces use the per-operator API above. */
typedef struct {
    cs_dims  dims;
    int      nbc[6];      /* [2 faces][3 axes] */
    int      viscous;
    cs_point umf;         /* frame velocity (amach1..3, physics_base.hpp:135-145) */
    double   mu;          /* ssk: constant dynamic viscosity, 1/Re */
    double   metric[3];   /* xi_x, eta_y, zeta_z */
    double   yaco;        /* Jacobian */
    double   cfl;         /* canard.cu:120 uses 0.95 */
    double   alpha_f;     /* compact filter parameter (0.49) */
    int      filter;      /* filter the conserved variables after each RK4 step */
} cs_solver_config;

typedef struct cs_solver cs_solver;

/* a viscous solver owns one more stream (default priority) on which the
 * second gradient sweep of every RHS runs concurrently with the first; events
 * fork it from and join it back to the caller's stream, so the caller sees
 * one stream's ordering and a capture of that stream records both.  Calls on
 * one solver from several streams at once race, as for every solver call. */
int cs_solver_create(const cs_solver_config* cfg, cs_solver** out);
int cs_solver_destroy(cs_solver* sv);
/* de = the reference's de after calc_fluxes + gcbc_go (dq/dt = -yaco*de).
 * sbcc: [lmx] characteristic-update weights (read on the 4 rows next to each
 * non-reflective face). */
int cs_solver_rhs(cs_solver* sv, const double* qa, double* de, const double* sbcc, cs_stream stream);
/* one of the six sweeps of cs_solver_rhs (0..5 = the six sweeps in launch
 * order of cs_solver_set_order, three gradient sweeps then three flux
 * divergence sweeps; 6 = the last sweep fused with the last RK4 stage update
 * de = qa - dt*yaco*RHS, dt from the last time step, 0 before any; 7 = the
 * compact filter of the five variables of qa along x, y, z into de, three
 * launches, when every axis has 8-point chunks; 8, 9, 10 = that filter along
 * x, y or z only)
 * on the solver's scratch state; no reference counterpart -- it exists to
 * time and profile single kernels.  Sweeps depend on the outputs of the
 * earlier ones of the same RHS. */
int cs_solver_sweep(cs_solver* sv, int which, const double* qa, double* de, const double* sbcc,
                    cs_stream stream);
/* kernel variant of the sweeps along `axis` (no reference counterpart; the
 * reference specialises its NVRTC kernels with -D macros instead,
 * src/numerics_rtc.hpp:120-126): variant 0 = each chunk thread evaluates its
 * own rows (halo rows exchanged between chunks), 1 = LDS row-block tiles
 * with coalesced loads, 2 = variant 0 on 16-point chunks (y/z lines of 32 to
 * 256 points, a multiple of 16; the default along z); lines = 8 or 16 lines
 * per workgroup (y/z lines of variants 0 and 2, all lines of variant 1);
 * xcd = 1 keeps neighbouring line groups on one XCD.  Results agree to
 * rounding whatever the choice. */
int cs_solver_tune(cs_solver* sv, int axis, int variant, int lines, int xcd);
/* launch shape of the filter sweep along axis (no reference counterpart):
 * fields = 1 filters one conserved variable per workgroup (the five variables
 * spread over the grid; default along y and z), 5 all five per workgroup
 * (default along x); lines = 8 or 16 lines per workgroup of the y/z sweeps
 * (default 16); chunk = 8 or 16 points per chunk thread (16: y/z lines of 32
 * to 256 points, a multiple of 16; the default there).  Results agree to
 * rounding whatever the choice. */
int cs_solver_tune_filter(cs_solver* sv, int axis, int fields, int lines, int chunk);
/* launch shapes in use along axis (no reference counterpart): shape6 =
 * {variant, lines, xcd} of the right-hand-side sweeps, then {fields, lines,
 * chunk} of the fused filter sweep ({0, 0, 0} when that filter runs the line
 * solver). */
int cs_solver_get_tune(const cs_solver* sv, int axis, int* shape6);
/* autotuner of the fused sweeps (no reference counterpart; the reference
 * fixes its block shapes per build, src/numerics_rtc.hpp:120-126): times
 * every admissible shape of cs_solver_tune for the gradient and flux sweeps
 * of each axis, and of cs_solver_tune_filter for each fused filter sweep, on
 * qa (reps timed launches each, 1..1000), applies the fastest and records it
 * in the process-wide cache (cs_tune_set_sweep, cs_tune_set_filter_sweep),
 * which every solver created afterwards in the process applies.  A shape
 * replaces the one in use only when it is 2% faster.  Overwrites the solver's
 * stage buffers; synchronises the stream. */
int cs_solver_autotune(cs_solver* sv, const double* qa, const double* sbcc, int reps, cs_stream stream);
/* storage precision of the viscous intermediates between the sweeps (no
 * reference counterpart; the reference templates everything on one Type):
 * 0 = FP64 (default); 1 = the velocity and c^2 gradients of the first two
 * gradient sweeps, the stress tensor and the heat flux stored in FP32 (588
 * instead of 736 HBM bytes per point and RHS).  The conserved variables, the
 * partial RHS, closure and seam rows, characteristic BCs and the RK
 * accumulation stay FP64, and every sweep computes in FP64.  No effect on
 * inviscid solvers. */
int cs_solver_set_precision(cs_solver* sv, int mode);
/* curvilinear grid (the fused form of calc_fluxes + calc_viscous_shear_stress
 * + gcbc_go with full metrics, src/cuda/kernels/physics.hpp:98-191,228-333,
 * src/cuda/kernels/gcbc.hpp:251-454): device arrays the caller keeps alive
 * while the solver uses them -- xim, etm, zem [3][lmx] (the metric rows of
 * xi, eta, zeta as the per-operator calls take them), yaco [lmx], and the face
 * normals cm0, cm1, cm2 of the x, y and z faces ([2 face][3][face], as
 * cs_gcbc_go takes them).  Every sweep, the time step and the RK stage then
 * read them per point instead of cfg.metric / cfg.yaco (ssk stays cfg.mu).
 * All seven NULL: back to the constants.  cs_solver_set_precision applies as
 * on Cartesian grids, and so do the adjoint calls (vjp and tape,
 * inviscid: the fused transposed sweeps or the per-operator products, with
 * the same per-point arrays). */
int cs_solver_set_metrics(cs_solver* sv, const double* xim, const double* etm, const double* zem,
                          const double* yaco, const double* cm0, const double* cm1, const double* cm2);
/* order of the six sweeps (no reference counterpart): 0 = gradients along
 * x, y, then z with the stress closure, fluxes along x, y, then z with the
 * RHS/stage write; 1 (default) = gradients y, z, x (closure in the x sweep),
 * fluxes z, y, x (RHS/stage write in the x sweep).  Results agree to
 * rounding (the three flux divergences are summed in the other order). */
int cs_solver_set_order(cs_solver* sv, int order);
/* time step of physics::calc_time_step on qa (D9 corrected); dte may be NULL
 * (then nothing synchronises and the value stays on the device). */
int cs_solver_time_step(cs_solver* sv, const double* qa, double* dte, cs_stream stream);
/* one RK4 step in place on qa: dt from cs_solver_time_step when dt <= 0
 * (the same value bit for bit; a viscous solver in the default sweep order
 * reduces it inside a gradient sweep of the first stage: with constant
 * metrics the first one where it runs whole waves of y lines, with per-point
 * metrics the closing x sweep where x lines split into 8-point chunks), stages q_k = q0 - dt/(5-k) * yaco * de(q_{k-1}),
 * k = 1..4, then the compact filter along x, y, z when cfg.filter.  dt_out
 * (host, may be NULL) receives the step used; passing it synchronises the
 * stream. */
int cs_solver_rk4_step(cs_solver* sv, double* qa, const double* sbcc, double dt, double* dt_out,
                       cs_stream stream);
/* adjoint of the inviscid right-hand side (no reference counterpart):
 * g = (dR/dq at qa)^T lam, R = cs_solver_rhs, characteristic corrections
 * included.  The first adjoint call allocates 41 fields of scratch. */
int cs_solver_rhs_vjp(cs_solver* sv, const double* qa, const double* lam, double* g, const double* sbcc,
                      cs_stream stream);
/* adjoint of one RK4 step of cs_solver_rk4_step with the given dt (held
 * fixed; the filter included when the solver filters): lam (the gradient of
 * a loss with respect to the state after the step) is replaced by the
 * gradient with respect to q0, the state before it.  q0 is not modified;
 * the three inner stages are recomputed from it. */
int cs_solver_rk4_vjp(cs_solver* sv, const double* q0, double* lam, double dt, const double* sbcc,
                      cs_stream stream);
/* size, in doubles, of the tape of one RK4 step (no reference counterpart):
 * the stage states q0..q3 (20 fields) and, per stage, the boundary
 * derivatives of the flux divergence along each axis ([3][2][5][face]). */
int cs_solver_tape_size(const cs_solver* sv, size_t* doubles);
/* cs_solver_rk4_step that also records the tape of the step into `tape`
 * (device, cs_solver_tape_size doubles) for cs_solver_rk4_vjp_tape; inviscid
 * solvers.  Same state and dt semantics as cs_solver_rk4_step. */
int cs_solver_rk4_step_tape(cs_solver* sv, double* qa, const double* sbcc, double dt, double* dt_out, double* tape,
                            cs_stream stream);
/* lam <- (d step / d q0)^T lam for the step that recorded `tape`, with the
 * step's dt: the adjoint of cs_solver_rk4_vjp without recomputing the stages. */
int cs_solver_rk4_vjp_tape(cs_solver* sv, const double* tape, double* lam, double dt, const double* sbcc,
                           cs_stream stream);
/* adjoint implementation (no reference counterpart): 1 (default) = transposed
 * line sweeps fused with the pointwise Jacobian products, used when every
 * axis splits into 8-point chunks (N % 8 == 0, N <= 256); 0 = per-operator
 * transposed line solves (cs_deriv_adjoint, cs_filter_adjoint).  Results
 * agree to rounding. */
int cs_solver_adjoint_mode(cs_solver* sv, int mode);

/* ---- batched subdomains (no reference counterpart) ----------------------
 * The plain method -- derivative of a field along x, y and z
 * (numerics_pc::deriv1d, src/numerics_pc.hpp:74-119, seam faces wrapped on
 * the subdomain itself) and the compact filter along x, y, z -- on many
 * independent subdomains of unrelated shapes: one launch per kernel variant
 * (chunk length, axis, chunk count) serves the items of all subdomains.
 * Each subdomain's results are bit-identical to cs_deriv / cs_filter on it
 * alone. */
typedef struct cs_batch cs_batch;
/* count subdomains; dims[count]; nbc6 = [count][2 face][3 axis] */
int cs_batch_create(int count, const cs_dims* dims, const int* nbc6, cs_batch** out);
int cs_batch_destroy(cs_batch* b);
/* binds the device fields: f[count] (input, filtered in place by the filter
 * phase) and d[3*count] (d[3*s + axis] = derivative of f[s] along axis).
 * Builds the work lists; the only call of the batch that uploads. */
int cs_batch_bind(cs_batch* b, double* const* f, double* const* d);
int cs_batch_deriv(cs_batch* b, cs_stream stream);      /* derivatives along x, y, z */
int cs_batch_filter(cs_batch* b, cs_stream stream);     /* filter x, then y, then z, in place */
int cs_batch_step(cs_batch* b, cs_stream stream);       /* deriv, then filter */
/* grid of the batched launches: 0 (default) = one workgroup per work item,
 * the hardware dispatcher balances; 1 = persistent, as many workgroups as
 * fit on the device at once, items dealt round-robin; 2 = one launch per
 * subdomain and operator of the single-subdomain kernels (cs_deriv /
 * cs_filter), cheap to launch once captured; a whole step (cs_batch_step,
 * cs_batch_capture) deals the subdomains over eight streams that fork from
 * and join back into `stream`, so that small launches overlap.  Same
 * results. */
int cs_batch_set_persistent(cs_batch* b, int persistent);
/* the grid mode in use (set, or picked by cs_batch_autotune) */
int cs_batch_get_persistent(const cs_batch* b, int* mode);
/* records cs_batch_step into a hipGraph on `stream` (not the null stream);
 * cs_batch_replay launches it */
int cs_batch_capture(cs_batch* b, cs_stream stream);
int cs_batch_replay(cs_batch* b, cs_stream stream);
/* work items and kernel launches of the four phases (deriv, filter x, y, z) */
int cs_batch_info(const cs_batch* b, int* items4, int* launches4);
/* autotuner: times the admissible chunk lengths (8, 16, 32 points) of every
 * (operator, line length, start/end rows, precision) of the batch, records
 * the fastest (cs_tune_set), rebuilds the
 * batch and picks the fastest grid mode (cs_batch_set_persistent).  The
 * choices hold for every operator table built afterwards in
 * the process (the numerics contexts created later included), so batched and
 * single-subdomain results stay identical. */
int cs_batch_autotune(cs_batch* b, int reps, cs_stream stream);
/* the process-wide cache of launch shapes (no reference counterpart; the
 * reference specialises its NVRTC kernels per line length, boundary rows and
 * block shape with -D macros, src/numerics_rtc.hpp:105-126).  Line entries:
 * chunk length K (8, 16, 32; 0 clears the entry) of the per-operator and
 * batched line kernels for op (0 = derivative, 1 = filter), line length n,
 * start/end = the first/last rows of the line (0 = closure rows, 1 = seam
 * halo rows) and precision (cs_set_precision's mode); operator tables built
 * afterwards use it. */
int cs_tune_set(int op, int n, int start, int end, int precision, int K);
int cs_tune_get(int op, int n, int start, int end, int precision, int* K);   /* K = 0: none */
/* fused right-hand-side sweeps: shape of cs_solver_tune for (axis, line
 * length n, seam = 1 for a periodic or inter-subdomain axis, precision =
 * cs_solver_set_precision's mode, metrics = 0 for constant metrics, 1 for
 * per-point metrics (cs_solver_set_metrics)); solvers apply it at creation
 * and when their precision or metrics are set */
int cs_tune_set_sweep(int axis, int n, int seam, int precision, int metrics, int variant, int lines, int xcd);
int cs_tune_get_sweep(int axis, int n, int seam, int precision, int metrics, int* found, int* shape3);
/* fused filter sweeps: shape of cs_solver_tune_filter for (axis, n, seam);
 * solvers apply it at creation */
int cs_tune_set_filter_sweep(int axis, int n, int seam, int fields, int lines, int chunk);
int cs_tune_get_filter_sweep(int axis, int n, int seam, int* found, int* shape3);
int cs_tune_clear(void);   /* all three kinds */
/* persisted as text, one entry per line ('#' starts a comment):
 *   line <op> <n> <start> <end> <precision> <K>
 *   sweep <axis> <n> <seam> <precision> <metrics> <variant> <lines> <xcd>
 *   filter <axis> <n> <seam> <fields> <lines> <chunk>
 * cs_tune_load adds the entries of a file to the cache; a line of three
 * integers "<op> <n> <K>" sets K for FP64 lines with closure rows at both ends
 * and with seam rows at both ends, a sweep line without the metrics field is
 * an entry for constant metrics.  A malformed or inadmissible entry stops
 * the load with CS_ERR_ARG (or CS_ERR_SHAPE) and a message naming file:line;
 * the entries before it stay. */
int cs_tune_save(const char* path);
int cs_tune_load(const char* path);

/* ---- run configuration, block grid and output (host code, no device) ---
 * The reference's driver reads config.yaml with yaml-cpp
 * (driver/canard.cu:37-98), generates the block's coordinates
 * (src/grid.hpp:89-110) and writes legacy VTK (src/IO.hpp:108-159).  The
 * sections driver, domdcomp, grid and physics follow config.yaml:1-18 (values
 * of block 0; one block runs as one subdomain on one GPU, whatever its nbpc);
 * the optional section run holds what driver/canard.cu:117-127 fixes in code
 * plus this build's solver settings and initial pulse. */
typedef struct {
    int      nblocks;               /* driver.nblocks                          */
    int      nbpc[3];               /* domdcomp.nbpc of block 0                */
    int      lximb, letmb, lzemb;   /* last point index per axis (points - 1)  */
    int      nbbc[6];               /* face BC codes, [2 faces][3 axes]        */
    int      mbcd[6];               /* neighbour blocks, [2 faces][3 axes]     */
    double   doml0, doml1, domh, span;   /* grid extents (grid.hpp:74-87)     */
    cs_point amach;                 /* amach1..3 -> umf (physics_base.hpp:135) */
    double   cfl, tmax;             /* run.cfl (0.95), run.tmax (1.0)          */
    int      ndata;                 /* run.ndata: outputs over tmax (2)        */
    int      nsteps;                /* run.nsteps: step cap, 0 = until tmax    */
    int      viscous;               /* run.viscous (1)                         */
    double   re, alpha_f;           /* run.re (1e4), run.alpha_f (0.49)        */
    int      filter;                /* run.filter (1)                          */
    double   pulse_amp, pulse_width;   /* initial acoustic pulse (1e-3, 8 cells) */
} cs_run_config;

/* the defaults of every field (the run section's values above, one block of
 * non-reflective faces) */
int cs_config_defaults(cs_run_config* cfg);
/* domdcomp::read_config + grid::read_config + physics_base::read_config
 * (src/domdcomp.hpp:67-107, src/grid.hpp:74-87, src/physics_base.hpp:135-145)
 * of a file in the schema of config.yaml:1-18; errors name the file and line. */
int cs_config_load(const char* path, cs_run_config* cfg);
/* t_dcomp of block 0 (points = last index + 1) and its face BC codes nbc6
 * ([2][3], may be NULL) */
int cs_config_dims(const cs_run_config* cfg, cs_dims* dims, int* nbc6);
/* grid::generate, src/grid.hpp:89-110: host array xyz = [3][lmx] of point
 * coordinates x = -doml0 + i (doml0+doml1)/lximb, y = -domh + j 2 domh/letmb,
 * z = span ((lzemb - k)/lzemb - 1/2). */
int cs_config_grid(const cs_run_config* cfg, double* xyz);
/* IOwriter::go_vtk, src/IO.hpp:108-159: legacy VTK structured grid, BINARY
 * (big-endian), points then one SCALARS block per variable.  Host arrays:
 * xyz = [3][lmx], vars[v] = [lmx]; precision 32 writes float (as the
 * reference), 64 double. */
int cs_vtk_write(const char* path, const cs_dims* dims, const double* xyz, int nvars, const char* const* names,
                 const double* const* vars, int precision);

#ifdef __cplusplus
}
#endif

#endif /* COMPACT_SCHEMES_H */
