// fused.hip -- the right-hand side as six fused line sweeps (uniform
// Cartesian metrics), with the characteristic BCs and the RK stage update
// folded into the sweeps.
//
// The reference evaluates one RHS as calc_viscous_shear_stress + calc_fluxes
// (src/physics_pc.hpp:83-227) -- 27 line solves and 12 pointwise kernels, each
// reading and writing whole fields -- followed by gcbc_go (src/gcbc.hpp:86-186).
// Here each sweep reads its inputs once, differentiates several fields along
// one axis and writes only what the next sweep needs:
//   S1x, S1y: q -> (u, v, w, c^2) -> d/dx, d/dy                      (4 fields out)
//   S1z:      q, x/y gradients -> d/dz -> stress (6) + heat flux (3)
//   S2x:      q, stress row x, heat x -> F^x (5) -> dF/dx            (partial RHS)
//   S2y, S2z: ... + previous partial RHS; S2z adds the RK stage update
// and every S2 sweep applies the non-reflective characteristic update of its
// two faces from the boundary values of its own solve (the reference's drva).
//
// Line solve: chunked LU recurrences as in line_solver.hip, chunks balanced
// (lengths differ by at most one, every chunk >= 4 rows).  Kernel variants,
// chosen per axis (cs_solver_tune):
//   direct: a thread evaluates the pointwise inputs of its own chunk rows and
//     gets the two halo rows on each side from the neighbouring chunks (wave
//     shuffles for x lines, LDS for y/z lines); carries by wave scans (x) or
//     LDS folds (y/z).  Needs equal chunks (N = P*K); other lengths run the
//   ragged kernel, where every chunk also evaluates its halo rows;
//   tile: the inputs of TL lines pass through an LDS tile one row block at a
//     time with coalesced loads (x lines load along the line).
// Pointwise arithmetic (primitives, fluxes, closure and seam rows, stress,
// characteristic BCs, stage update) is not contracted, so it rounds as the
// CPU oracle does; the recurrences of the solve use FMA.
//
// Addressing: a field is at most 4 GiB, so point offsets are 32-bit byte
// offsets added to the per-field base pointers the kernel arguments hold.
#include "fused_kernels.hpp"

namespace cs {
namespace fused {

// ---------------------------------------------------------------------------
// time step with uniform Cartesian metrics (calc_time_step_kernel,
// src/cuda/kernels/physics.hpp:335-427, D9 corrected), primitives inline
// ---------------------------------------------------------------------------
constexpr int kTB = 256;

__global__ void __launch_bounds__(kTB) dt_partial_kernel(const double* q, long lmx, Phys ph, double* part) {
    __shared__ double s0[kTB], s1[kTB];
    double m0 = 0.0, m1 = 0.0;
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < lmx; t += (long)gridDim.x * blockDim.x) {
        double qq[5];
        load_q(q, t, lmx, qq);
        const Prim p = primitive(qq);
        m0 = fmax(m0, dt_conv(ph, p.u, p.v, p.w, p.c2));
        m1 = fmax(m1, dt_visc(ph, p.rhoi));
    }
    s0[threadIdx.x] = m0;
    s1[threadIdx.x] = m1;
    __syncthreads();
    for (int w = kTB / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            s0[threadIdx.x] = fmax(s0[threadIdx.x], s0[threadIdx.x + w]);
            s1[threadIdx.x] = fmax(s1[threadIdx.x], s1[threadIdx.x + w]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) { part[2 * blockIdx.x] = s0[0]; part[2 * blockIdx.x + 1] = s1[0]; }
}

// the same with per-point metric rows and Jacobian (curvilinear solvers;
// calc_time_step_kernel's full form)
__global__ void __launch_bounds__(kTB) dt_partial_curv_kernel(const double* q, long lmx, Phys ph, Metric mt,
                                                              double* part) {
    __shared__ double s0[kTB], s1[kTB];
    double m0 = 0.0, m1 = 0.0;
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < lmx; t += (long)gridDim.x * blockDim.x) {
        double qq[5];
        load_q(q, t, lmx, qq);
        double M[3][3];
        for (int a = 0; a < 3; a++)
            for (int k = 0; k < 3; k++) M[a][k] = mt.m[a][k][t];
        dt_curv_point(M, mt.J[t], ph, primitive(qq), m0, m1);
    }
    s0[threadIdx.x] = m0;
    s1[threadIdx.x] = m1;
    __syncthreads();
    for (int w = kTB / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            s0[threadIdx.x] = fmax(s0[threadIdx.x], s0[threadIdx.x + w]);
            s1[threadIdx.x] = fmax(s1[threadIdx.x], s1[threadIdx.x + w]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) { part[2 * blockIdx.x] = s0[0]; part[2 * blockIdx.x + 1] = s1[0]; }
}

// out[0] = dt = min(cfl/max0, 0.5/max1) (or dt_fixed when > 0), out[1..2] = maxima
__global__ void __launch_bounds__(kTB) dt_final_kernel(const double* part, int n, double cfl, double dt_fixed,
                                                       double* out) {
    __shared__ double s0[kTB], s1[kTB];
    double m0 = 0.0, m1 = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) { m0 = fmax(m0, part[2 * i]); m1 = fmax(m1, part[2 * i + 1]); }
    s0[threadIdx.x] = m0;
    s1[threadIdx.x] = m1;
    __syncthreads();
    for (int w = kTB / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            s0[threadIdx.x] = fmax(s0[threadIdx.x], s0[threadIdx.x + w]);
            s1[threadIdx.x] = fmax(s1[threadIdx.x], s1[threadIdx.x + w]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[1] = s0[0];
        out[2] = s1[0];
        out[0] = dt_fixed > 0.0 ? dt_fixed : fmin(cfl / s0[0], 0.5 / s1[0]);
    }
}

__global__ void set_dt_kernel(double dt, double* out) {
    if (threadIdx.x == 0) { out[0] = dt; out[1] = 0.0; out[2] = 0.0; }
}

static int sweep_at(cs_solver* sv, int pos, const SweepCommon& cm, const double* q, double* out, const double* q0,
                    int div, hipStream_t s) {
    if (sv->curv) return sweep_at_curv(sv, pos, cm, q, out, q0, div, s);
    if (sv->cfg.viscous && sv->prec == 1) return sweep_at_f32(sv, pos, cm, q, out, q0, div, s);
    if (pos < 3) return s1_sweep(sv, pos, cm, q, s);
    if (sv->cfg.viscous) return s2_sweep<true>(sv, pos - 3, cm, q, out, q0, div, s);
    return s2_sweep<false>(sv, pos - 3, cm, q, out, q0, div, s);
}

int dt_fold_pos(const cs_solver* sv) {
    if (!sv->cfg.viscous || !sv->dtpart || sv->order != 1) return -1;
    if (sv->curv) {
        // the closing x sweep on the transposing kernel (8-point chunks)
        const AxisDesc& r = sv->rax[kS1Axes[1][2]];
        return r.variant != 1 && r.K == 8 && r.N == r.P * 8 ? 2 : -1;
    }
    // the first (y) sweep on the direct kernel with whole waves
    const AxisDesc& r = sv->rax[kS1Axes[1][0]];
    const int TL = r.tl == 8 ? 8 : 16;
    return r.variant != 1 && r.N == r.P * r.K && (TL * r.P) % 64 == 0 ? 0 : -1;
}

// workgroups of the sweep at position pos (0: the direct y/z kernel, 2: the
// transposing x kernel), one pair of maxima each
static int fold_blocks(const cs_solver* sv, int pos) {
    const AxisDesc& r = sv->rax[kS1Axes[sv->order][pos]];
    const int lines = pos == 2 ? 64 / r.GS : (r.tl == 8 ? 8 : 16);
    return (int)((r.nlines + lines - 1) / lines);
}

int rhs_sweeps(cs_solver* sv, const double* q, const double* sbcc, double* out, const double* q0, int div,
               hipStream_t s) {
    const SweepCommon cm = common_of(sv, sbcc);
    for (int pos = sv->cfg.viscous ? 0 : 3; pos < 6; pos++) {
        if (pos == 0) {
            // the second gradient sweep on the aux stream, concurrently with
            // the first; the closing sweep waits for both
            CS_HIP(hipEventRecord(sv->ev_fork, s));
            CS_HIP(hipStreamWaitEvent(sv->aux, sv->ev_fork, 0));
            const int rc = sweep_at(sv, 1, cm, q, out, nullptr, div, sv->aux);
            if (rc) return rc;
            CS_HIP(hipEventRecord(sv->ev_join, sv->aux));
        }
        if (pos == 1) continue;
        if (pos == 2) CS_HIP(hipStreamWaitEvent(s, sv->ev_join, 0));
        const int rc = sweep_at(sv, pos, cm, q, out, pos == 5 ? q0 : nullptr, div, s);
        if (rc) return rc;
        if (sv->dt_fold == pos + 1) {
            // the sweep wrote one pair of maxima per workgroup
            hipLaunchKernelGGL(dt_final_kernel, dim3(1), dim3(kTB), 0, s, sv->dtpart, fold_blocks(sv, pos),
                               sv->cfg.cfl, 0.0, sv->dt);
            CS_HIP(hipGetLastError());
        }
    }
    return CS_OK;
}

// One sweep of the right-hand side on the solver's scratch buffers (0..5 =
// the six sweeps in launch order, the last in RHS mode; 6 = the last sweep with
// the RK stage update out = q - dt*J*RHS of the last RK4 stage, dt from the
// device time-step slot; 7 = the filter of the five variables along x, y, z
// into out, 8/9/10 = that filter along x/y/z only); for timing single kernels.
int one_sweep(cs_solver* sv, int which, const double* q, const double* sbcc, double* out, hipStream_t s) {
    const SweepCommon cm = common_of(sv, sbcc);
    if (which < 0 || which > 10) return fail(CS_ERR_ARG, "sweep index out of 0..10");
    if (which >= 7) {
        for (int a = 0; a < 3; a++)
            if ((which == 7 || which - 8 == a) && !filter_axis_fused(sv, a))
                return fail(CS_ERR_UNSUPPORTED, "the filter of this axis has no fused sweep");
        if (out != q) CS_HIP(hipMemcpyAsync(out, q, sizeof(double) * 5 * sv->g.lmx, hipMemcpyDeviceToDevice, s));
        return filter_sweeps(sv, out, s, which == 7 ? -1 : which - 8);
    }
    if (which < 3 && !sv->cfg.viscous) return fail(CS_ERR_STATE, "gradient sweeps exist only for the viscous RHS");
    if (which == 6) return sweep_at(sv, 5, cm, q, out, q, 1, s);
    return sweep_at(sv, which, cm, q, out, nullptr, 0, s);
}

bool filter_fusable(const cs_solver* sv) {
    for (int a = 0; a < 3; a++)
        if (sv->ax[a].K != 8 || sv->ax[a].N != sv->ax[a].P * 8 || sv->ax[a].P > kMaxP || !sv->ax[a].ftab)
            return false;
    return true;
}

bool filter_axis_fused(const cs_solver* sv, int a) { return sv->fax[a].ftab != nullptr; }

template <int AX>
static int filter_axis(cs_solver* sv, double* q, const SweepCommon& cm, hipStream_t s) {
    const AxisDesc& ax = sv->fax[AX];
    if (ax.ffields == 1) {
        SF1<AX> f{q, sv->g.lmx};
        if constexpr (AX != 0)
            if (ax.K == 16) return launch_filter<16, AX>(ax, cm, f, s);
        return launch_filter<8, AX>(ax, cm, f, s);
    }
    SF<AX> f{};
    fields(q, sv->g.lmx, f.q);
    if constexpr (AX != 0)
        if (ax.K == 16) return launch_filter<16, AX>(ax, cm, f, s);
    return launch_filter<8, AX>(ax, cm, f, s);
}

// the compact filter of the five conserved variables in place: along x, y, z
// (axis < 0) or along one axis, one launch per axis
int filter_sweeps(cs_solver* sv, double* q, hipStream_t s, int axis) {
    const SweepCommon cm = common_of(sv, nullptr);
    int rc = CS_OK;
    if ((axis < 0 || axis == 0) && (rc = filter_axis<0>(sv, q, cm, s))) return rc;
    if ((axis < 0 || axis == 1) && (rc = filter_axis<1>(sv, q, cm, s))) return rc;
    if (axis < 0 || axis == 2) rc = filter_axis<2>(sv, q, cm, s);
    return rc;
}

int time_step(cs_solver* sv, const double* q, double dt_fixed, hipStream_t s) {
    const long lmx = sv->g.lmx;
    if (dt_fixed > 0.0) {
        hipLaunchKernelGGL(set_dt_kernel, dim3(1), dim3(64), 0, s, dt_fixed, sv->dt);
        CS_HIP(hipGetLastError());
        return CS_OK;
    }
    if (sv->curv)
        hipLaunchKernelGGL(dt_partial_curv_kernel, dim3(sv->red_blocks), dim3(kTB), 0, s, q, lmx, phys_of(sv), sv->met,
                           sv->red);
    else
        hipLaunchKernelGGL(dt_partial_kernel, dim3(sv->red_blocks), dim3(kTB), 0, s, q, lmx, phys_of(sv), sv->red);
    CS_HIP(hipGetLastError());
    hipLaunchKernelGGL(dt_final_kernel, dim3(1), dim3(kTB), 0, s, sv->red, sv->red_blocks, sv->cfg.cfl, 0.0, sv->dt);
    CS_HIP(hipGetLastError());
    return CS_OK;
}

}  // namespace fused
}  // namespace cs
