// physics_kernels.hpp -- pointwise physics shared by the per-stage kernels and
// the fused sweeps.  The arithmetic follows src/cuda/kernels/physics.hpp of
// the reference operation for operation, so that both paths give the same
// values as the reference kernels up to rounding.
#pragma once

#include "cs_internal.hpp"

namespace cs {

__device__ __forceinline__ void load_q(const double* qa, long t, long lmx, double* q) {
#pragma unroll
    for (int v = 0; v < 5; v++) q[v] = qa[t + (long)v * lmx];
}

struct Prim {
    double rhoi, u, v, w, p, c2;
};

// primitive variables from the conserved ones (absent upstream)
__device__ __forceinline__ Prim primitive(const double* q) {
    Prim r;
    r.rhoi = 1.0 / q[0];
    r.u = q[1] * r.rhoi;
    r.v = q[2] * r.rhoi;
    r.w = q[3] * r.rhoi;
    r.p = kGamm1 * (q[4] - 0.5 * (q[1] * r.u + q[2] * r.v + q[3] * r.w));
    r.c2 = kGam * r.p * r.rhoi;
    return r;
}

// Flux of variable m along direction j (metric row mt[j]),
// calc_fluxes_pre_compute_kernel, src/cuda/kernels/physics.hpp:98-191.
// tau = (xx, yy, zz, xy, yz, zx), hf = (x, y, z).
template <bool Visc>
__device__ __forceinline__ void flux_all(const double* q, double p, const double* u, const double (*mt)[3],
                                         const double* tau, const double* hf, cs_point umf, double (*F)[3]) {
    double rr[3] = {u[0] + umf.x, u[1] + umf.y, u[2] + umf.z};
    double ss[3];
#pragma unroll
    for (int j = 0; j < 3; j++) ss[j] = mt[j][0] * rr[0] + mt[j][1] * rr[1] + mt[j][2] * rr[2];
    const double ep = q[4] + p;
#pragma unroll
    for (int j = 0; j < 3; j++) {
        F[0][j] = q[0] * ss[j];
        double f1 = q[1] * ss[j] + mt[j][0] * p;
        double f2 = q[2] * ss[j] + mt[j][1] * p;
        double f3 = q[3] * ss[j] + mt[j][2] * p;
        double f4 = ep * ss[j] - p * (umf.x * mt[j][0] + umf.y * mt[j][1] + umf.z * mt[j][2]);
        if (Visc) {
            f1 -= (mt[j][0] * tau[0] + mt[j][1] * tau[3] + mt[j][2] * tau[5]);
            f2 -= (mt[j][0] * tau[3] + mt[j][1] * tau[1] + mt[j][2] * tau[4]);
            f3 -= (mt[j][0] * tau[5] + mt[j][1] * tau[4] + mt[j][2] * tau[2]);
            f4 -= (mt[j][0] * hf[0] + mt[j][1] * hf[1] + mt[j][2] * hf[2]);
        }
        F[1][j] = f1;
        F[2][j] = f2;
        F[3][j] = f3;
        F[4][j] = f4;
    }
}

// Stress and heat flux from the velocity and c^2 gradients,
// calc_viscous_shear_stress_final_kernel, src/cuda/kernels/physics.hpp:285-333.
// g = (u_x, u_y, u_z, v_x, v_y, v_z, w_x, w_y, w_z) in physical coordinates.
__device__ __forceinline__ void viscous_closure(double mu, double yaco, const double* g, const double* c2g,
                                                const double* u, double* tau, double* hf, double* d4) {
    const double b0 = mu * yaco;
    const double b1 = kGamm1Prndtli * b0;
    const double dv = 2.0 / 3.0 * (g[0] + g[4] + g[8]);
    *d4 = dv;
    tau[0] = b0 * (2.0 * g[0] - dv);
    tau[1] = b0 * (2.0 * g[4] - dv);
    tau[2] = b0 * (2.0 * g[8] - dv);
    tau[3] = b0 * (g[3] + g[1]);
    tau[4] = b0 * (g[7] + g[5]);
    tau[5] = b0 * (g[2] + g[6]);
    hf[0] = b1 * c2g[0] + u[0] * tau[0] + u[1] * tau[3] + u[2] * tau[5];
    hf[1] = b1 * c2g[1] + u[0] * tau[3] + u[1] * tau[1] + u[2] * tau[4];
    hf[2] = b1 * c2g[2] + u[0] * tau[5] + u[1] * tau[4] + u[2] * tau[2];
}

int launch_primitives(const double* qa, double* de, double* p, long lmx, hipStream_t s);
int launch_fluxes_pre(bool visc, double* buffer, const double* qa, const double* p, double* de, const double* xim,
                      const double* etm, const double* zem, const double* stress, const double* heat,
                      cs_point umf, long lmx, hipStream_t s);
int launch_fluxes_post(double* de, const double* bd, int m, long lmx, hipStream_t s);
int launch_visc_init(double* buffer, const double* de, long lmx, hipStream_t s);
int launch_visc_post(double* stress, double* heat, double* ss, const double* bd, const double* xim,
                     const double* etm, const double* zem, int m, long lmx, hipStream_t s);
int launch_visc_final(double* de, const double* ssk, const double* yaco, const double* ss, double* stress,
                      double* heat, long lmx, hipStream_t s);
int time_step_blocks(long lmx);
int launch_time_step(const double* xim, const double* etm, const double* zem, const double* de, const double* yaco,
                     const double* ssk, cs_point umf, double* part, int nblocks, long lmx, hipStream_t s);
int launch_rk_stage(double* q, const double* q0, const double* de, const double* yaco, double dtk, long lmx,
                    hipStream_t s);
int launch_fill(const double* in, double* send, const double* pbco, int axis, int face, int N, long nlines, int lxi,
                long plane, long st, hipStream_t s);

// characteristic boundary conditions (gcbc.hip)
int launch_gcbc_setup(cs_numerics* num, int axis, int face, int flag, const double* cm, const double* qa,
                      const double* de, const double* p, const double* yaco, cs_point umf, hipStream_t s);
int launch_gcbc_update(cs_numerics* num, int axis, int face, int code, const double* cm, const double* drvb,
                       const double* qa, double* de, const double* p, const double* sbcc, cs_point umf,
                       cs_point dudtmf, double dtwi, hipStream_t s);
int launch_wall(cs_numerics* num, int axis, int face, int code, double* qa, const int* npex, cs_point umf,
                hipStream_t s);

}  // namespace cs
