// metrics.hip -- grid metrics of a curvilinear subdomain from its point
// coordinates (absent upstream: src/grid.hpp allocates xim/etm/zem and fills
// only the coordinates, src/grid.hpp:66-72,89-213).
//
// With x(xi, eta, zeta) the coordinates and the compact derivative D_a along
// each index direction, the metric rows the flux and stress kernels read are
// the cofactors of the coordinate Jacobian M[c][a] = d x_c / d xi_a,
//   xim = (grad xi)/J, etm = (grad eta)/J, zem = (grad zeta)/J,
// i.e. xim = M_eta x M_zeta, etm = M_zeta x M_xi, zem = M_xi x M_eta (cross
// products of the columns of M), and yaco = J = 1 / det M -- the convention
// of the RHS (flux = metric row . F, stage update q0 - dt*yaco*de,
// stress = mu*J*(...)).  Along a seam (periodic) axis the coordinates jump by
// one period vector; the derivative is taken of x - period*i/N and the
// constant period/N is added back.
#include "cs_internal.hpp"

namespace cs {

namespace {

// tmp = x - jump * (i_axis / N)
__global__ void __launch_bounds__(256) deramp_kernel(const double* x, double* tmp, double jump, int axis, int lxi,
                                                     int let, int N, long lmx) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= lmx) return;
    const long i = t % lxi, jk = t / lxi;
    const long idx = axis == 0 ? i : (axis == 1 ? jk % let : jk / let);
    tmp[t] = x[t] - jump * ((double)idx / (double)N);
}

__global__ void __launch_bounds__(256) add_kernel(double* y, double c, long n) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) y[t] += c;
}

// in place: m[a][c] = d x_c / d xi_a (a = 0: xim, 1: etm, 2: zem) -> cofactors
__global__ void __launch_bounds__(256) cofactor_kernel(double* xim, double* etm, double* zem, double* yaco,
                                                       long lmx) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= lmx) return;
    double M[3][3];   // M[c][a]
#pragma unroll
    for (int c = 0; c < 3; c++) {
        M[c][0] = xim[t + c * lmx];
        M[c][1] = etm[t + c * lmx];
        M[c][2] = zem[t + c * lmx];
    }
    // columns u = M[.][0], v = M[.][1], w = M[.][2]
    const double cx[3] = {M[1][1] * M[2][2] - M[2][1] * M[1][2], M[2][1] * M[0][2] - M[0][1] * M[2][2],
                          M[0][1] * M[1][2] - M[1][1] * M[0][2]};   // v x w
    const double ce[3] = {M[1][2] * M[2][0] - M[2][2] * M[1][0], M[2][2] * M[0][0] - M[0][2] * M[2][0],
                          M[0][2] * M[1][0] - M[1][2] * M[0][0]};   // w x u
    const double cz[3] = {M[1][0] * M[2][1] - M[2][0] * M[1][1], M[2][0] * M[0][1] - M[0][0] * M[2][1],
                          M[0][0] * M[1][1] - M[1][0] * M[0][1]};   // u x v
    const double det = M[0][0] * cx[0] + M[1][0] * cx[1] + M[2][0] * cx[2];
#pragma unroll
    for (int c = 0; c < 3; c++) {
        xim[t + c * lmx] = cx[c];
        etm[t + c * lmx] = ce[c];
        zem[t + c * lmx] = cz[c];
    }
    yaco[t] = 1.0 / det;
}

dim3 grid_of(long n) { return dim3((unsigned)((n + 255) / 256)); }

}  // namespace

}  // namespace cs

using namespace cs;

extern "C" {

int cs_metrics(cs_numerics* num, const double* xyz, const double* period9, double* xim, double* etm, double* zem,
               double* yaco, cs_stream stream) {
    if (!num || !xyz || !xim || !etm || !zem || !yaco) return fail(CS_ERR_ARG, "null argument");
    if (!num->table_set) return fail(CS_ERR_STATE, "metrics need deriv_setup (seam rows) first");
    const Geo& g = num->g;
    const long lmx = g.lmx;
    hipStream_t s = (hipStream_t)stream;
    double* out[3] = {xim, etm, zem};
    for (int a = 0; a < 3; a++) {
        const int st = num->ndf[0][a], en = num->ndf[1][a];
        if (st != en) return fail(CS_ERR_UNSUPPORTED, "an axis with one seam face needs its neighbour's coordinates");
        for (int c = 0; c < 3; c++) {
            const double* x = xyz + (long)c * lmx;
            double* d = out[a] + (long)c * lmx;
            const double jump = (st == 1 && period9) ? period9[3 * a + c] : 0.0;
            if (jump != 0.0) {
                // yaco is free until the cofactors: the de-ramped coordinate
                hipLaunchKernelGGL(deramp_kernel, grid_of(lmx), dim3(256), 0, s, x, yaco, jump, a, g.lxi, g.let,
                                   g.n(a), lmx);
                CS_HIP(hipGetLastError());
                x = yaco;
            }
            int rc = line_apply(num, OP_DERIV, a, x, d, st, en, nullptr, 0, false, s);
            if (rc) return rc;
            if (jump != 0.0) {
                hipLaunchKernelGGL(add_kernel, grid_of(lmx), dim3(256), 0, s, d, jump / g.n(a), lmx);
                CS_HIP(hipGetLastError());
            }
        }
    }
    hipLaunchKernelGGL(cofactor_kernel, grid_of(lmx), dim3(256), 0, s, xim, etm, zem, yaco, lmx);
    CS_HIP(hipGetLastError());
    return CS_OK;
}

}  // extern "C"
