// gcbc.hip -- generalized characteristic boundary conditions on one subdomain.
//
// Restates src/cuda/kernels/gcbc.hpp of the reference with its defects
// corrected: the left-eigenvector matrix is the inverse of the right one
// (D10), the update index is base + v*lmx with sbcc read at the point (D11),
// each thread owns one face entry (D12), and the far face sits at dim-1
// (D13).  One thread per line ending on the face; face arrays are
// [face][component][face_idx] like the reference's drva and cm.
#include "physics_kernels.hpp"

namespace cs {

constexpr int kGBlock = 256;

struct FaceGeo {
    long fs, lmx, plane, st;
    int lxi, axis, face, N;
    __device__ long point(long f) const {
        return line_base(axis, f, lxi, plane) + (long)face * (N - 1) * st;
    }
};

// gcbc_setup_kernel<Axis>, src/cuda/kernels/gcbc.hpp:251-342:
// drva <- L * (yaco * (flag ? drva : de)) at the face points.
__global__ void __launch_bounds__(kGBlock) gcbc_setup_kernel(FaceGeo fg, int flag, const double* cm, double* drva,
                                                             const double* qa, const double* de, const double* pr,
                                                             const double* yaco, cs_point umf) {
    long f = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= fg.fs) return;
    const long idx = fg.point(f);
    double cmv[3], cha[5], res[5];
    for (int d = 0; d < 3; d++) cmv[d] = cm[(long)fg.face * 3 * fg.fs + d * fg.fs + f];
    double* dv = drva + (long)fg.face * 5 * fg.fs + f;
    const double J = yaco[idx];
    for (int v = 0; v < 5; v++) cha[v] = J * (flag ? dv[v * fg.fs] : de[idx + v * fg.lmx]);
    double mom[3] = {qa[idx + fg.lmx], qa[idx + 2 * fg.lmx], qa[idx + 3 * fg.lmx]};
    Eig e;
    e.eleme(cmv, qa[idx], mom, pr[idx], umf);
    e.l_apply(cmv, cha, res);
    for (int v = 0; v < 5; v++) dv[v * fg.fs] = res[v];
}

// gcbc_update_{non_reflective,wall,inter_curv}_kernel<Axis>,
// src/cuda/kernels/gcbc.hpp:344-695
__global__ void __launch_bounds__(kGBlock) gcbc_update_kernel(FaceGeo fg, int code, const double* cm,
                                                              const double* drva, const double* drvb,
                                                              const double* qa, double* de, const double* pr,
                                                              const double* sbcc, const double* yaco, cs_point umf,
                                                              cs_point dudtmf, double dtwi) {
    long f = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= fg.fs) return;
    const long idx = fg.point(f);
    double cmv[3], dv[5], cha[5], res[5];
    for (int d = 0; d < 3; d++) cmv[d] = cm[(long)fg.face * 3 * fg.fs + d * fg.fs + f];
    for (int v = 0; v < 5; v++) dv[v] = drva[(long)fg.face * 5 * fg.fs + v * fg.fs + f];
    double mom[3] = {qa[idx + fg.lmx], qa[idx + 2 * fg.lmx], qa[idx + 3 * fg.lmx]};
    Eig e;
    e.eleme(cmv, qa[idx], mom, pr[idx], umf);
    const double ra0 = 1 - 2 * fg.face;
    for (int v = 0; v < 5; v++) cha[v] = dv[v];
    if (code == CS_BC_NON_REFLECTIVE) {
        if (ra0 * (e.vn + e.vs + e.ao) > 0) cha[3] = 0;
        if (ra0 * (e.vn + e.vs - e.ao) > 0) cha[4] = 0;
        for (int v = 0; v < 5; v++) cha[v] -= dv[v];
    } else if (code == CS_BC_INTER_CURV) {
        double dh[5];
        for (int v = 0; v < 5; v++) dh[v] = drvb[(long)fg.face * 5 * fg.fs + v * fg.fs + f];
        if (ra0 * (e.vn + e.vs) > 0) { cha[0] = dh[0]; cha[1] = dh[1]; cha[2] = dh[2]; }
        if (ra0 * (e.vn + e.vs + e.ao) > 0) cha[3] = dh[3];
        if (ra0 * (e.vn + e.vs - e.ao) > 0) cha[4] = dh[4];
        for (int v = 0; v < 5; v++) cha[v] -= dv[v];
    } else {
        const double accum = cmv[0] * dudtmf.x + cmv[1] * dudtmf.y + cmv[2] * dudtmf.z;
        cha[3 + fg.face] = cha[4 - fg.face] + 2 * ra0 * e.aoi * qa[idx] * (accum + dtwi * (e.vn + e.vs));
        // D22: the change of the amplitudes, as in the other branches (the
        // reference adds R*cha, i.e. J*de once more, at the wall rows)
        for (int v = 0; v < 5; v++) cha[v] -= dv[v];
    }
    e.r_apply(cmv, cha, res);
    // D25: R*cha is a change of yaco*de (the setup multiplied by yaco), so de
    // changes by R*cha / yaco (the reference adds R*cha to de; the same when yaco = 1)
    for (int v = 0; v < 5; v++) res[v] /= yaco[idx];
    const int iq = 1 - 2 * fg.face;
    for (int ii = 0; ii < kMbci; ii++) {
        const long o = idx + (long)ii * iq * fg.st;
        const double w = sbcc[o];
        for (int v = 0; v < 5; v++) de[o + v * fg.lmx] += w * res[v];
    }
}

// wall_inviscid_kernel / wall_viscous_kernel, src/cuda/kernels/gcbc.hpp:160-249
__global__ void __launch_bounds__(kGBlock) wall_kernel(FaceGeo fg, int code, double* qa, const int* npex,
                                                       cs_point umf) {
    long f = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= fg.fs) return;
    const long idx = fg.point(f);
    const long lmx = fg.lmx;
    double q0 = qa[idx], q1 = qa[idx + lmx], q2 = qa[idx + 2 * lmx], q3 = qa[idx + 3 * lmx], q4 = qa[idx + 4 * lmx];
    double n = npex[idx];
    if (code == CS_BC_WALL_INVISCID) {
        double v2 = q1 * q1 + q2 * q2 + q3 * q3;
        qa[idx + 4 * lmx] = n * q4 + (1.0 - n) * (kHamHamm1 * pow(q0, kGam) + 0.5 * v2 / q0);
    } else {
        // D23: the kinetic energy of the momentum just set, so that an
        // isothermal point (npex = 0) ends at p = rho / gamma
        double fctr = (1 - n) * q0;
        double m1 = n * q1 - fctr * umf.x, m2 = n * q2 - fctr * umf.y, m3 = n * q3 - fctr * umf.z;
        double v2 = m1 * m1 + m2 * m2 + m3 * m3;
        qa[idx + lmx] = m1;
        qa[idx + 2 * lmx] = m2;
        qa[idx + 3 * lmx] = m3;
        qa[idx + 4 * lmx] = n * q4 + (1 - n) * (kHamHamm1 * q0 + 0.5 * v2 / q0);
    }
}

static FaceGeo face_geo(const cs_numerics* num, int axis, int face) {
    FaceGeo fg;
    fg.fs = num->g.face_size(axis);
    fg.lmx = num->g.lmx;
    fg.plane = (long)num->g.lxi * num->g.let;
    fg.st = num->g.stride(axis);
    fg.lxi = num->g.lxi;
    fg.axis = axis;
    fg.face = face;
    fg.N = num->g.n(axis);
    return fg;
}

int launch_gcbc_setup(cs_numerics* num, int axis, int face, int flag, const double* cm, const double* qa,
                      const double* de, const double* p, const double* yaco, cs_point umf, hipStream_t s) {
    FaceGeo fg = face_geo(num, axis, face);
    dim3 grid((unsigned)((fg.fs + kGBlock - 1) / kGBlock));
    hipLaunchKernelGGL(gcbc_setup_kernel, grid, dim3(kGBlock), 0, s, fg, flag, cm, num->drva[axis], qa, de, p, yaco,
                       umf);
    CS_HIP(hipGetLastError());
    return CS_OK;
}

int launch_gcbc_update(cs_numerics* num, int axis, int face, int code, const double* cm, const double* drvb,
                       const double* qa, double* de, const double* p, const double* sbcc, const double* yaco,
                       cs_point umf, cs_point dudtmf, double dtwi, hipStream_t s) {
    FaceGeo fg = face_geo(num, axis, face);
    dim3 grid((unsigned)((fg.fs + kGBlock - 1) / kGBlock));
    hipLaunchKernelGGL(gcbc_update_kernel, grid, dim3(kGBlock), 0, s, fg, code, cm, num->drva[axis], drvb, qa, de, p,
                       sbcc, yaco, umf, dudtmf, dtwi);
    CS_HIP(hipGetLastError());
    return CS_OK;
}

int launch_wall(cs_numerics* num, int axis, int face, int code, double* qa, const int* npex, cs_point umf,
                hipStream_t s) {
    FaceGeo fg = face_geo(num, axis, face);
    dim3 grid((unsigned)((fg.fs + kGBlock - 1) / kGBlock));
    hipLaunchKernelGGL(wall_kernel, grid, dim3(kGBlock), 0, s, fg, code, qa, npex, umf);
    CS_HIP(hipGetLastError());
    return CS_OK;
}

}  // namespace cs
