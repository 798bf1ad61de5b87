// vjp.hip -- adjoint (vector-Jacobian products) of the fused solver's Euler
// right-hand side and RK4 step (no reference counterpart).
//
// The inviscid right-hand side of cs_solver (metric rows m_a and Jacobian J,
// constant or per point, frame velocity umf) is
//   R(q) = sum_a [ D_a F^a(q) + C^a(q) ],
// F^a the flux row of axis a (calc_fluxes_pre_compute_kernel,
// src/cuda/kernels/physics.hpp:98-191), D_a the compact derivative along a,
// and C^a the non-reflective characteristic correction of the two faces of a
// when both are BC_NON_REFLECTIVE (gcbc_update_non_reflective_kernel,
// src/cuda/kernels/gcbc.hpp:251-454): at a face point f with boundary
// derivative g = (D_a F^a)(f), c = R_e P L_e (J g), added with weight sbcc
// to the 4 rows next to the face.  Its transpose applied to lambda is
//   R'(q)^T lambda = sum_a [ (dF^a/dq)^T D_a^T (lambda + S_a^T mu_a) + G_a ],
// with mu_a = (dc/dg)^T lam_f, G_a = (dc/dq)^T lam_f at the face points and
// lam_f = sum of sbcc*lambda over the 4 rows.  The pointwise Jacobian
// products come from forward-mode dual numbers on the same formulas; D_a^T
// is cs_deriv_adjoint.  The RK4 step adjoint recomputes the three inner
// stages from q0 and runs the stages backwards (see cs_solver_rk4_vjp).
// These per-operator products serve lines that do not split into 8-point
// chunks and adjoint mode 0; tape.hip holds the fused transposed sweeps.
#include "vjp_parts.hpp"

namespace cs {
namespace fused {

namespace {

// metric row of axis ax at point t: the constant one, or the point's (mt.J
// non-null: curvilinear solvers)
__device__ __forceinline__ void metric_row(const EulerPhys& ph, const Metric& mt, int ax, long t, double* m) {
    if (mt.J) {
#pragma unroll
        for (int k = 0; k < 3; k++) m[k] = mt.m[ax][k][t];
    } else {
#pragma unroll
        for (int k = 0; k < 3; k++) m[k] = k == ax ? ph.m[ax] : 0.0;
    }
}

// F[v] = flux row of axis ax at every point
__global__ void __launch_bounds__(256) flux_fields_kernel(const double* q, double* F, int ax, EulerPhys ph,
                                                          Metric mt, long lmx) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= lmx) return;
    double qq[5], f[5];
    load_q(q, t, lmx, qq);
    if (mt.J) {
        double m[3];
        metric_row(ph, mt, ax, t, m);
        flux_row_m(qq, m, ph.umf, f);
    } else {
        flux_row(ax, qq, ph.m[ax], ph.umf, f);
    }
#pragma unroll
    for (int v = 0; v < 5; v++) F[t + v * lmx] = f[v];
}

// g += (dF^ax/dq)^T nu
__global__ void __launch_bounds__(256) flux_vjp_kernel(const double* q, const double* nu, double* g, int ax,
                                                       EulerPhys ph, Metric mt, long lmx) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= lmx) return;
    typedef Dual<5> D;
    D qd[5], f[5];
#pragma unroll
    for (int v = 0; v < 5; v++) qd[v] = D::seed(q[t + v * lmx], v);
    if (mt.J) {
        double m[3];
        metric_row(ph, mt, ax, t, m);
        flux_row_m(qd, m, ph.umf, f);
    } else {
        flux_row(ax, qd, ph.m[ax], ph.umf, f);
    }
    D phi(0.0);
#pragma unroll
    for (int v = 0; v < 5; v++) phi = phi + f[v] * nu[t + v * lmx];
#pragma unroll
    for (int v = 0; v < 5; v++) g[t + v * lmx] += phi.d[v];
}

// face points of axis ax: lam_f = sum_ii sbcc*lambda over the 4 rows next to
// the face; seed[face row] += (dc/dg)^T lam_f; g[face point] += (dc/dq)^T lam_f.
// drva = [2 face][5 var][face] boundary derivatives (the forward solve's)
__global__ void __launch_bounds__(256) gcbc_vjp_kernel(const double* q, const double* lam, const double* sbcc,
                                                       const double* drva, double* seed, double* g, int ax, int N,
                                                       long nface, long lxi, long plane, long st, EulerPhys ph,
                                                       Metric mt, long lmx) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 2 * nface) return;
    const int face = (int)(t / nface);
    const long f = t % nface;
    long base;
    if (ax == 0) base = f * lxi;
    else if (ax == 1) base = (f / lxi) * plane + (f % lxi);
    else base = f;
    const long pt = base + (long)(face ? N - 1 : 0) * st;
    const long dir = face ? -st : st;
    double lf[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    for (int ii = 0; ii < kMbci; ii++) {
        const long r = pt + ii * dir;
        const double w = sbcc[r];
#pragma unroll
        for (int v = 0; v < 5; v++) lf[v] += w * lam[r + v * lmx];
    }
    double qq[5], gd[5];
#pragma unroll
    for (int v = 0; v < 5; v++) {
        qq[v] = q[pt + v * lmx];
        gd[v] = drva[((long)face * 5 + v) * nface + f];
    }
    // face normal and Jacobian: the caller's arrays (curvilinear solvers) or the constants
    double cm[3], J = ph.J;
    if (mt.J) {
        for (int k = 0; k < 3; k++) cm[k] = mt.cm[ax][((long)face * 3 + k) * nface + f];
        J = mt.J[pt];
    } else {
        const double sg = ph.m[ax] * ph.J < 0.0 ? -1.0 : 1.0;
        for (int k = 0; k < 3; k++) cm[k] = k == ax ? sg : 0.0;
    }
    typedef Dual<5> D;
    // with respect to the boundary derivative (q fixed)
    {
        D gdd[5], c[5];
#pragma unroll
        for (int v = 0; v < 5; v++) gdd[v] = D::seed(gd[v], v);
        face_corr_at<double, D>(face, qq, gdd, cm, J, ph.umf, c);
        D phi(0.0);
#pragma unroll
        for (int v = 0; v < 5; v++) phi = phi + c[v] * lf[v];
#pragma unroll
        for (int v = 0; v < 5; v++) seed[pt + v * lmx] += phi.d[v];
    }
    // with respect to the state at the face point (g fixed)
    {
        D qd[5], gdd[5], c[5];
#pragma unroll
        for (int v = 0; v < 5; v++) {
            qd[v] = D::seed(qq[v], v);
            gdd[v] = D(gd[v]);
        }
        face_corr_at<D, D>(face, qd, gdd, cm, J, ph.umf, c);
        D phi(0.0);
#pragma unroll
        for (int v = 0; v < 5; v++) phi = phi + c[v] * lf[v];
#pragma unroll
        for (int v = 0; v < 5; v++) g[pt + v * lmx] += phi.d[v];
    }
}

__global__ void __launch_bounds__(256) zero_kernel(double* y, long n) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) y[t] = 0.0;
}

__global__ void __launch_bounds__(256) axpby_kernel(double* y, double a, const double* x, double b, long n) {
    // y = a*x + b*y
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) y[t] = a * x[t] + b * y[t];
}

__global__ void __launch_bounds__(256) stage_kernel(double* q, const double* q0, const double* de, double c, long n) {
    // q = q0 - c*de
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) q[t] = q0[t] - c * de[t];
}

// the same with the per-point Jacobian: q = q0 - c*J*de, five fields lmx apart
__global__ void __launch_bounds__(256) stage_j_kernel(double* q, const double* q0, const double* de, double c,
                                                      const double* J, long lmx) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 5 * lmx) return;
    q[t] = q0[t] - c * J[t % lmx] * de[t];
}

// y = J x, five fields lmx apart (the seed of a stage's R^T on a curvilinear grid)
__global__ void __launch_bounds__(256) scale_j_kernel(double* y, const double* x, const double* J, long lmx) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < 5 * lmx) y[t] = J[t % lmx] * x[t];
}

dim3 grid_of(long n) { return dim3((unsigned)((n + 255) / 256)); }

int axpby(double* y, double a, const double* x, double b, long n, hipStream_t s) {
    hipLaunchKernelGGL(axpby_kernel, grid_of(n), dim3(256), 0, s, y, a, x, b, n);
    CS_HIP(hipGetLastError());
    return CS_OK;
}

}  // namespace

static int vjp_buffers(cs_solver* sv) {
    if (sv->vjp) return CS_OK;
    // F, seed, nu (5 fields each), the throwaway derivative (1), stage states
    // q1..q3 (15), the stage adjoint and the RHS adjoint (5 each)
    const size_t n = (size_t)sv->g.lmx * 41;
    CS_HIP(hipMalloc(&sv->vjp, n * sizeof(double)));
    return CS_OK;
}

static EulerPhys euler_phys(const cs_solver* sv) {
    EulerPhys ph;
    ph.umf = sv->cfg.umf;
    ph.J = sv->cfg.yaco;
    for (int a = 0; a < 3; a++) ph.m[a] = sv->cfg.metric[a];
    return ph;
}

// the per-point geometry of a curvilinear solver, or all null (J null: constants)
static Metric metric_of(const cs_solver* sv) { return sv->curv ? sv->met : Metric{}; }

// g = R'(q)^T lam
static int rhs_vjp(cs_solver* sv, const double* q, const double* lam, double* g, const double* sbcc, hipStream_t s) {
    const long lmx = sv->g.lmx;
    const EulerPhys ph = euler_phys(sv);
    double* F = sv->vjp;
    double* seed = F + 5 * lmx;
    double* nu = seed + 5 * lmx;
    double* junk = nu + 5 * lmx;
    hipLaunchKernelGGL(zero_kernel, grid_of(5 * lmx), dim3(256), 0, s, g, 5 * lmx);
    CS_HIP(hipGetLastError());
    int rc;
    for (int ax = 0; ax < 3; ax++) {
        const AxisDesc& A = sv->ax[ax];
        CS_HIP(hipMemcpyAsync(seed, lam, sizeof(double) * 5 * lmx, hipMemcpyDeviceToDevice, s));
        if (A.gcbc) {
            hipLaunchKernelGGL(flux_fields_kernel, grid_of(lmx), dim3(256), 0, s, q, F, ax, ph, metric_of(sv), lmx);
            CS_HIP(hipGetLastError());
            // boundary derivatives of the flux row into drva (the forward solve)
            for (int v = 0; v < 5; v++)
                if ((rc = line_apply(sv->num, OP_DERIV, ax, F + v * lmx, junk, 0, 0, nullptr, v, true, s))) return rc;
            const long nface = sv->g.face_size(ax);
            hipLaunchKernelGGL(gcbc_vjp_kernel, grid_of(2 * nface), dim3(256), 0, s, q, lam, sbcc, sv->num->drva[ax],
                               seed, g, ax, A.N, nface, (long)sv->g.lxi, (long)sv->g.lxi * sv->g.let, (long)A.st, ph,
                               metric_of(sv), lmx);
            CS_HIP(hipGetLastError());
        }
        for (int v = 0; v < 5; v++)
            if ((rc = deriv_adjoint(sv->num, ax, seed + v * lmx, nu + v * lmx, s))) return rc;
        hipLaunchKernelGGL(flux_vjp_kernel, grid_of(lmx), dim3(256), 0, s, q, nu, g, ax, ph, metric_of(sv), lmx);
        CS_HIP(hipGetLastError());
    }
    return CS_OK;
}

int solver_rhs_vjp(cs_solver* sv, const double* q, const double* lam, double* g, const double* sbcc, hipStream_t s) {
    if (sv->cfg.viscous) return fail(CS_ERR_UNSUPPORTED, "the adjoint covers the inviscid right-hand side");
    int rc = vjp_buffers(sv);
    if (rc) return rc;
    return rhs_vjp(sv, q, lam, g, sbcc, s);
}

// lam <- (d q_next / d q0)^T lam for the RK4 step of cs_solver_rk4_step with
// the time step dt (held fixed), including the filter when enabled, from the
// stage states q0 and Q = q1, q2, q3 (per-operator line solves)
int rk4_vjp_lines(cs_solver* sv, const double* q0, const double* Q, double* lam, double dt, const double* sbcc,
                  hipStream_t s) {
    int rc = vjp_buffers(sv);
    if (rc) return rc;
    const long lmx = sv->g.lmx, n5 = 5 * lmx;
    double* t = sv->vjp + 31 * lmx;          // adjoint of the current stage
    double* gr = t + 5 * lmx;                // R'(q)^T t
    // a stage adds -c_k J(x) R(q) at each point x: with a constant Jacobian it
    // is folded into c_k, a per-point one scales the seed of R^T (in qc)
    const double J = sv->curv ? 1.0 : sv->cfg.yaco;
    const double ck[5] = {0.0, dt / 4.0 * J, dt / 3.0 * J, dt / 2.0 * J, dt * J};
    // through the filter (applied x, then y, then z after the last stage)
    if (sv->cfg.filter) {
        double* w1 = gr;
        double* w2 = sv->qb;
        for (int v = 0; v < 5; v++) {
            double* l = lam + v * lmx;
            if ((rc = filter_adjoint(sv->num, 2, l, w1 + v * lmx, s))) return rc;
            if ((rc = filter_adjoint(sv->num, 1, w1 + v * lmx, w2 + v * lmx, s))) return rc;
            if ((rc = filter_adjoint(sv->num, 0, w2 + v * lmx, l, s))) return rc;
        }
    }
    // a0 = a4 + a3 + a2 + a1 - c1 R'(q0)^T a1, a_{k-1} = -c_k R'(q_{k-1})^T a_k
    CS_HIP(hipMemcpyAsync(t, lam, sizeof(double) * n5, hipMemcpyDeviceToDevice, s));
    for (int k = 4; k >= 1; k--) {
        const double* qk1 = k == 1 ? q0 : Q + (long)(k - 2) * n5;
        const double* seed = t;
        if (sv->curv) {
            hipLaunchKernelGGL(scale_j_kernel, grid_of(n5), dim3(256), 0, s, sv->qc, t, sv->met.J, lmx);
            CS_HIP(hipGetLastError());
            seed = sv->qc;
        }
        if ((rc = rhs_vjp(sv, qk1, seed, gr, sbcc, s))) return rc;
        if (k > 1) {
            if ((rc = axpby(t, -ck[k], gr, 0.0, n5, s))) return rc;      // t = a_{k-1}
            if ((rc = axpby(lam, 1.0, t, 1.0, n5, s))) return rc;        // a0 += a_{k-1}
        } else {
            if ((rc = axpby(lam, -ck[1], gr, 1.0, n5, s))) return rc;    // a0 -= c1 R'(q0)^T a1
        }
    }
    return CS_OK;
}

// the per-operator form of cs_solver_rk4_vjp: the three inner stage states
// recomputed from q0, then rk4_vjp_lines
int solver_rk4_vjp(cs_solver* sv, const double* q0, double* lam, double dt, const double* sbcc, hipStream_t s) {
    int rc = vjp_buffers(sv);
    if (rc) return rc;
    const long lmx = sv->g.lmx, n5 = 5 * lmx;
    double* Q = sv->vjp + 16 * lmx;          // q1, q2, q3
    double* de = sv->dpart;                  // forward RHS (scratch of the solver)
    const double J = sv->curv ? 1.0 : sv->cfg.yaco;
    const double ck[4] = {0.0, dt / 4.0 * J, dt / 3.0 * J, dt / 2.0 * J};
    // stage states: q_k = q0 - c_k J R(q_{k-1})
    const double* prev = q0;
    for (int k = 1; k <= 3; k++) {
        if ((rc = rhs_sweeps(sv, prev, sbcc, de, nullptr, 0, s))) return rc;
        double* qk = Q + (long)(k - 1) * n5;
        if (sv->curv)
            hipLaunchKernelGGL(stage_j_kernel, grid_of(n5), dim3(256), 0, s, qk, q0, de, ck[k], sv->met.J, lmx);
        else
            hipLaunchKernelGGL(stage_kernel, grid_of(n5), dim3(256), 0, s, qk, q0, de, ck[k], n5);
        CS_HIP(hipGetLastError());
        prev = qk;
    }
    return rk4_vjp_lines(sv, q0, Q, lam, dt, sbcc, s);
}

}  // namespace fused
}  // namespace cs

extern "C" {

int cs_solver_rhs_vjp(cs_solver* sv, const double* qa, const double* lam, double* g, const double* sbcc,
                      cs_stream stream) {
    if (!sv || !qa || !lam || !g) return cs::fail(CS_ERR_ARG, "null argument");
    if (!sbcc && (sv->ax[0].gcbc || sv->ax[1].gcbc || sv->ax[2].gcbc))
        return cs::fail(CS_ERR_ARG, "non-reflective faces need sbcc");
    if (sv->cfg.viscous) return cs::fail(CS_ERR_UNSUPPORTED, "the adjoint covers the inviscid right-hand side");
    if (cs::fused::adjoint_fused(sv)) return cs::fused::rhs_vjp_fused(sv, qa, lam, g, sbcc, (hipStream_t)stream);
    return cs::fused::solver_rhs_vjp(sv, qa, lam, g, sbcc, (hipStream_t)stream);
}

int cs_solver_rk4_vjp(cs_solver* sv, const double* q0, double* lam, double dt, const double* sbcc, cs_stream stream) {
    if (!sv || !q0 || !lam) return cs::fail(CS_ERR_ARG, "null argument");
    if (!sbcc && (sv->ax[0].gcbc || sv->ax[1].gcbc || sv->ax[2].gcbc))
        return cs::fail(CS_ERR_ARG, "non-reflective faces need sbcc");
    if (sv->cfg.viscous) return cs::fail(CS_ERR_UNSUPPORTED, "the adjoint covers the inviscid right-hand side");
    if (!(dt > 0.0)) return cs::fail(CS_ERR_ARG, "the adjoint needs the step's dt > 0");
    if (cs::fused::adjoint_fused(sv)) return cs::fused::rk4_vjp_fused(sv, q0, lam, dt, sbcc, (hipStream_t)stream);
    return cs::fused::solver_rk4_vjp(sv, q0, lam, dt, sbcc, (hipStream_t)stream);
}

}  // extern "C"
