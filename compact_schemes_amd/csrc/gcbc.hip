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

struct Eig {
    double ao, aoi, hv2, vn, vs, ve[3];
    // gcbc_vgpr::eleme, src/cuda/kernels/gcbc.hpp:44-61
    __device__ void eleme(const double* cm, double qa1, const double* qa24, double pp, cs_point umf) {
        double rhoi = 1.0 / qa1;
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
    __device__ void r_apply(const double* cm, const double* x, double* y) const {
        const double bo = hv2 + kHamm1 * ao * ao, co = ao * vn;
        const double dm0 = ao * cm[0], dm1 = ao * cm[1], dm2 = ao * cm[2];
        const double a13 = 0.5 * (ve[0] + dm0), a23 = 0.5 * (ve[1] + dm1), a33 = 0.5 * (ve[2] + dm2);
        const double a43 = 0.5 * (bo + co);
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
    __device__ void l_apply(const double* cm, const double* x, double* y) const {
        const double ho = kGamm1 * aoi * aoi, bo = 1.0 - ho * hv2, co = aoi * vn;
        const double dm0 = aoi * cm[0], dm1 = aoi * cm[1], dm2 = aoi * cm[2];
        const double rv0 = ho * ve[0], rv1 = ho * ve[1], rv2 = ho * ve[2];
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
                                                              const double* sbcc, cs_point umf, cs_point dudtmf,
                                                              double dtwi) {
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
    }
    e.r_apply(cmv, cha, res);
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
    double v2 = q1 * q1 + q2 * q2 + q3 * q3;
    double n = npex[idx];
    if (code == CS_BC_WALL_INVISCID) {
        qa[idx + 4 * lmx] = n * q4 + (1.0 - n) * (kHamHamm1 * pow(q0, kGam) + 0.5 * v2 / q0);
    } else {
        double fctr = (1 - n) * q0;
        qa[idx + lmx] = n * q1 - fctr * umf.x;
        qa[idx + 2 * lmx] = n * q2 - fctr * umf.y;
        qa[idx + 3 * lmx] = n * q3 - fctr * umf.z;
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
                       const double* qa, double* de, const double* p, const double* sbcc, cs_point umf,
                       cs_point dudtmf, double dtwi, hipStream_t s) {
    FaceGeo fg = face_geo(num, axis, face);
    dim3 grid((unsigned)((fg.fs + kGBlock - 1) / kGBlock));
    hipLaunchKernelGGL(gcbc_update_kernel, grid, dim3(kGBlock), 0, s, fg, code, cm, num->drva[axis], drvb, qa, de, p,
                       sbcc, umf, dudtmf, dtwi);
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
