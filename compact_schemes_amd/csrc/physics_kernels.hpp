// physics_kernels.hpp -- pointwise physics shared by the per-stage kernels and
// the fused sweeps.  The arithmetic follows src/cuda/kernels/physics.hpp of
// the reference operation for operation, so that both paths give the same
// values as the reference kernels up to rounding.
#pragma once

#include "cs_internal.hpp"
#include "dual.hpp"

namespace cs {

__device__ __forceinline__ void load_q(const double* qa, long t, long lmx, double* q) {
#pragma unroll
    for (int v = 0; v < 5; v++) q[v] = qa[t + (long)v * lmx];
}

template <class T>
struct PrimT {
    T rhoi, u, v, w, p, c2;
};
using Prim = PrimT<double>;

// primitive variables from the conserved ones (absent upstream); T = double,
// or a dual number for Jacobian products
template <class T>
__device__ __forceinline__ PrimT<T> primitive(const T* q) {
    PrimT<T> r;
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

// Eigen-decomposition of the normal-flux Jacobian at a face point
// (gcbc_vgpr, src/cuda/kernels/gcbc.hpp:38-158; left vectors D10-corrected).
// T = double, or a dual number for Jacobian products; the vectors the
// matrices apply to may be dual numbers of their own (type X).
template <class T>
struct EigT {
    T ao, aoi, hv2, vn, vs, ve[3];
    // gcbc_vgpr::eleme, src/cuda/kernels/gcbc.hpp:44-61
    __device__ void eleme(const double* cm, T qa1, const T* qa24, T pp, cs_point umf) {
        T rhoi = 1.0 / qa1;
        ao = sqrt(kGam * rhoi * pp);
        aoi = 1.0 / ao;
        ve[0] = rhoi * qa24[0];
        ve[1] = rhoi * qa24[1];
        ve[2] = rhoi * qa24[2];
        hv2 = 0.5 * (ve[0] * ve[0] + ve[1] * ve[1] + ve[2] * ve[2]);
        vn = cm[0] * ve[0] + cm[1] * ve[1] + cm[2] * ve[2];
        vs = cm[0] * umf.x + cm[1] * umf.y + cm[2] * umf.z;
    }
    // characteristic -> conservative, gcbc_vgpr::xtr2q, src/cuda/kernels/gcbc.hpp:63-105
    template <class X>
    __device__ void r_apply(const double* cm, const X* x, X* y) const {
        const T bo = hv2 + kHamm1 * ao * ao, co = ao * vn;
        const T dm0 = ao * cm[0], dm1 = ao * cm[1], dm2 = ao * cm[2];
        const T a13 = 0.5 * (ve[0] + dm0), a23 = 0.5 * (ve[1] + dm1), a33 = 0.5 * (ve[2] + dm2);
        const T a43 = 0.5 * (bo + co);
        y[0] = cm[0] * x[0] + cm[1] * x[1] + cm[2] * x[2] + 0.5 * x[3] + 0.5 * x[4];
        y[1] = cm[0] * ve[0] * x[0] + (cm[1] * ve[0] - dm2) * x[1] + (cm[2] * ve[0] + dm1) * x[2] + a13 * x[3] +
               (a13 - dm0) * x[4];
        y[2] = (cm[0] * ve[1] + dm2) * x[0] + cm[1] * ve[1] * x[1] + (cm[2] * ve[1] - dm0) * x[2] + a23 * x[3] +
               (a23 - dm1) * x[4];
        y[3] = (cm[0] * ve[2] - dm1) * x[0] + (cm[1] * ve[2] + dm0) * x[1] + cm[2] * ve[2] * x[2] + a33 * x[3] +
               (a33 - dm2) * x[4];
        y[4] = (hv2 * cm[0] + dm2 * ve[1] - dm1 * ve[2]) * x[0] + (hv2 * cm[1] + dm0 * ve[2] - dm2 * ve[0]) * x[1] +
               (hv2 * cm[2] + dm1 * ve[0] - dm0 * ve[1]) * x[2] + a43 * x[3] + (a43 - co) * x[4];
    }
    // conservative -> characteristic, gcbc_vgpr::xtq2r, src/cuda/kernels/gcbc.hpp:107-150 (D10 corrected)
    template <class X>
    __device__ void l_apply(const double* cm, const X* x, X* y) const {
        const T ho = kGamm1 * aoi * aoi, bo = 1.0 - ho * hv2, co = aoi * vn;
        const T dm0 = aoi * cm[0], dm1 = aoi * cm[1], dm2 = aoi * cm[2];
        const T rv0 = ho * ve[0], rv1 = ho * ve[1], rv2 = ho * ve[2];
        y[0] = (bo * cm[0] + dm1 * ve[2] - dm2 * ve[1]) * x[0] + cm[0] * rv0 * x[1] + (cm[0] * rv1 + dm2) * x[2] +
               (cm[0] * rv2 - dm1) * x[3] - ho * cm[0] * x[4];
        y[1] = (bo * cm[1] + dm2 * ve[0] - dm0 * ve[2]) * x[0] + (cm[1] * rv0 - dm2) * x[1] + cm[1] * rv1 * x[2] +
               (cm[1] * rv2 + dm0) * x[3] - ho * cm[1] * x[4];
        y[2] = (bo * cm[2] + dm0 * ve[1] - dm1 * ve[0]) * x[0] + (cm[2] * rv0 + dm1) * x[1] +
               (cm[2] * rv1 - dm0) * x[2] + cm[2] * rv2 * x[3] - ho * cm[2] * x[4];
        y[3] = (1.0 - bo - co) * x[0] + (dm0 - rv0) * x[1] + (dm1 - rv1) * x[2] + (dm2 - rv2) * x[3] + ho * x[4];
        y[4] = (1.0 - bo + co) * x[0] + (-dm0 - rv0) * x[1] + (-dm1 - rv1) * x[2] + (-dm2 - rv2) * x[3] + ho * x[4];
    }
};
using Eig = EigT<double>;

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
                       const double* qa, double* de, const double* p, const double* sbcc, const double* yaco,
                       cs_point umf, cs_point dudtmf, double dtwi, hipStream_t s);
int launch_wall(cs_numerics* num, int axis, int face, int code, double* qa, const int* npex, cs_point umf,
                hipStream_t s);

}  // namespace cs
