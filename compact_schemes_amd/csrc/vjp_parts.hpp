// vjp_parts.hpp -- pointwise pieces of the Euler right-hand side shared by
// its adjoints (vjp.hip: per-operator products; tape.hip: transposed sweeps),
// templated on the number type so that dual numbers give Jacobian products.
#pragma once

#include "dual.hpp"
#include "fused.hpp"

namespace cs {
namespace fused {

struct EulerPhys {
    cs_point umf;
    double J;
    double m[3];
};

// flux row of axis ax (the S2 sweep's point(), Euler part)
template <class T>
__device__ __forceinline__ void flux_row(int ax, const T* q, double m, cs_point umf, T* F) {
    PrimT<T> p = primitive(q);
    const T vel = ax == 0 ? p.u : (ax == 1 ? p.v : p.w);
    const double uf = ax == 0 ? umf.x : (ax == 1 ? umf.y : umf.z);
    const T ss = m * (vel + uf);
    F[0] = q[0] * ss;
    F[1] = q[1] * ss;
    F[2] = q[2] * ss;
    F[3] = q[3] * ss;
    F[1 + ax] = F[1 + ax] + m * p.p;
    F[4] = (q[4] + p.p) * ss - p.p * (uf * m);
}

// flux row of axis a with a general metric row m (curvilinear solvers; the
// S2C sweep's point(), Euler part)
template <class T>
__device__ __forceinline__ void flux_row_m(const T* q, const double* m, cs_point umf, T* F) {
    PrimT<T> p = primitive(q);
    const T ss = m[0] * (p.u + umf.x) + m[1] * (p.v + umf.y) + m[2] * (p.w + umf.z);
    F[0] = q[0] * ss;
    F[1] = q[1] * ss + m[0] * p.p;
    F[2] = q[2] * ss + m[1] * p.p;
    F[3] = q[3] * ss + m[2] * p.p;
    F[4] = (q[4] + p.p) * ss - p.p * (umf.x * m[0] + umf.y * m[1] + umf.z * m[2]);
}

// characteristic correction of a face point (the S2 sweep's gcbc()) with
// the face normal cm and the point's Jacobian J
template <class T, class X>
__device__ __forceinline__ void face_corr_at(int face, const T* q, const X* drv, const double* cm, double J,
                                             cs_point umf, X* corr) {
    PrimT<T> p = primitive(q);
    EigT<T> e;
    e.eleme(cm, q[0], q + 1, p.p, umf);
    X cha[5], ch[5];
#pragma unroll
    for (int v = 0; v < 5; v++) cha[v] = J * drv[v];
    e.l_apply(cm, cha, ch);
    const double ra0 = 1 - 2 * face;
    X dl[5] = {X(0.0), X(0.0), X(0.0), X(0.0), X(0.0)};
    if (ra0 * value_of(e.vn + e.vs + e.ao) > 0) dl[3] = -ch[3];
    if (ra0 * value_of(e.vn + e.vs - e.ao) > 0) dl[4] = -ch[4];
    e.r_apply(cm, dl, corr);
#pragma unroll
    for (int v = 0; v < 5; v++) corr[v] = corr[v] / J;   // D25: de changes by R*cha / J
}

// the same with the constant metrics of a Cartesian solver
template <class T, class X>
__device__ __forceinline__ void face_corr(int ax, int face, const T* q, const X* drv, const EulerPhys& ph, X* corr) {
    const double sg = ph.m[ax] * ph.J < 0.0 ? -1.0 : 1.0;   // unit normal of grad(xi_ax) = J * metric * e_ax
    const double cm[3] = {ax == 0 ? sg : 0.0, ax == 1 ? sg : 0.0, ax == 2 ? sg : 0.0};
    face_corr_at<T, X>(face, q, drv, cm, ph.J, ph.umf, corr);
}

}  // namespace fused
}  // namespace cs
