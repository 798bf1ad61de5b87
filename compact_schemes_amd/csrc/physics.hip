// physics.hip -- pointwise kernels of the compressible right-hand side.
//
// Each kernel computes what the reference kernel named in its comment
// computes (src/cuda/kernels/physics.hpp of EnricoDeg/compact_schemes) on the
// same SoA layout; corrections of reference defects are named where applied.
// All are one-point-per-thread streaming kernels: HBM-bound, 256 threads.
#include "physics_kernels.hpp"

#include <cmath>
#include <mutex>
#include <string>

namespace cs {

constexpr int kBlock = 256;

static dim3 grid_for(long n) { return dim3((unsigned)((n + kBlock - 1) / kBlock)); }

// primitive variables (absent upstream): de0 = 1/rho, de1..3 = u, v, w,
// de4 = c^2 = gam p / rho, p = (gam-1)(E - rho |u|^2 / 2)
__global__ void __launch_bounds__(kBlock) primitives_kernel(const double* qa, double* de, double* p, long lmx) {
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= lmx) return;
    double q[5];
    load_q(qa, t, lmx, q);
    Prim w = primitive(q);
    de[t] = w.rhoi;
    de[t + lmx] = w.u;
    de[t + 2 * lmx] = w.v;
    de[t + 3 * lmx] = w.w;
    de[t + 4 * lmx] = w.c2;
    p[t] = w.p;
}

// calc_fluxes_pre_compute_kernel, src/cuda/kernels/physics.hpp:42-193
// (D21 intended: stress/heat are not read when Visc is false).
template <bool Visc>
__global__ void __launch_bounds__(kBlock) fluxes_pre_kernel(double* buffer, const double* qa, const double* pr,
                                                            double* de, const double* xim, const double* etm,
                                                            const double* zem, const double* stress,
                                                            const double* heat, cs_point umf, long lmx) {
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= lmx) return;
    double q[5], u[3], mt[3][3], tau[6] = {0, 0, 0, 0, 0, 0}, hf[3] = {0, 0, 0};
    load_q(qa, t, lmx, q);
    u[0] = de[t + lmx]; u[1] = de[t + 2 * lmx]; u[2] = de[t + 3 * lmx];
    for (int d = 0; d < 3; d++) {
        mt[0][d] = xim[t + d * lmx]; mt[1][d] = etm[t + d * lmx]; mt[2][d] = zem[t + d * lmx];
    }
    if (Visc) {
        for (int k = 0; k < 6; k++) tau[k] = stress[t + k * lmx];
        for (int k = 0; k < 3; k++) hf[k] = heat[t + k * lmx];
    }
    double F[5][3];
    flux_all<Visc>(q, pr[t], u, mt, tau, hf, umf, F);
    for (int m = 0; m < 5; m++)
        for (int j = 0; j < 3; j++) buffer[t + (long)j * lmx + (long)m * 3 * lmx] = F[m][j];
    de[t + 4 * lmx] = q[4] + pr[t];
}

// calc_fluxes_post_compute_kernel<m>, src/cuda/kernels/physics.hpp:195-209
__global__ void __launch_bounds__(kBlock) fluxes_post_kernel(double* de, const double* bd, int m, long lmx) {
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= lmx) return;
    const double* d = bd + (long)m * 3 * lmx;
    de[t + (long)m * lmx] = d[t] + d[t + lmx] + d[t + 2 * lmx];
}

// calc_viscous_shear_stress_init_kernel, src/cuda/kernels/physics.hpp:211-225
__global__ void __launch_bounds__(kBlock) visc_init_kernel(double* buffer, const double* de, long lmx) {
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= lmx) return;
    for (int m = 1; m < 5; m++) buffer[t + (long)m * 3 * lmx] = de[t + (long)m * lmx];
}

// calc_viscous_shear_stress_post_compute_kernel<m>, src/cuda/kernels/physics.hpp:228-283,
// D8 intended (reads the derivatives).  Slot map as the reference: m=1 ->
// (S.xx, H.zz, S.zx), m=2 -> (S.xy, S.yy, H.xx), m=3 -> (H.yy, S.yz, S.zz), m=4 -> ss.
__global__ void __launch_bounds__(kBlock) visc_post_kernel(double* stress, double* heat, double* ss,
                                                           const double* bd, const double* xim,
                                                           const double* etm, const double* zem, int m,
                                                           long lmx) {
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= lmx) return;
    double *o0, *o1, *o2;
    if (m == 1)      { o0 = stress;           o1 = heat + 2 * lmx;   o2 = stress + 5 * lmx; }
    else if (m == 2) { o0 = stress + 3 * lmx; o1 = stress + lmx;     o2 = heat; }
    else if (m == 3) { o0 = heat + lmx;       o1 = stress + 4 * lmx; o2 = stress + 2 * lmx; }
    else             { o0 = ss;               o1 = ss + lmx;         o2 = ss + 2 * lmx; }
    const double* d = bd + (long)m * 3 * lmx;
    double f0 = d[t], f1 = d[t + lmx], f2 = d[t + 2 * lmx];
    o0[t] = xim[t] * f0 + etm[t] * f1 + zem[t] * f2;
    o1[t] = xim[t + lmx] * f0 + etm[t + lmx] * f1 + zem[t + lmx] * f2;
    o2[t] = xim[t + 2 * lmx] * f0 + etm[t + 2 * lmx] * f1 + zem[t + 2 * lmx] * f2;
}

// calc_viscous_shear_stress_final_kernel, src/cuda/kernels/physics.hpp:285-333
__global__ void __launch_bounds__(kBlock) visc_final_kernel(double* de, const double* ssk, const double* yaco,
                                                            const double* ss, double* stress, double* heat,
                                                            long lmx) {
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= lmx) return;
    double g[9];   // (u_x, u_y, u_z, v_x, v_y, v_z, w_x, w_y, w_z) from the slot map
    g[0] = stress[t];           g[1] = heat[t + 2 * lmx];   g[2] = stress[t + 5 * lmx];
    g[3] = stress[t + 3 * lmx]; g[4] = stress[t + lmx];     g[5] = heat[t];
    g[6] = heat[t + lmx];       g[7] = stress[t + 4 * lmx]; g[8] = stress[t + 2 * lmx];
    double c2g[3] = {ss[t], ss[t + lmx], ss[t + 2 * lmx]};
    double u[3] = {de[t + lmx], de[t + 2 * lmx], de[t + 3 * lmx]};
    double tau[6], hf[3], d4;
    viscous_closure(ssk[t], yaco[t], g, c2g, u, tau, hf, &d4);
    de[t] = ssk[t];
    de[t + 4 * lmx] = d4;
    for (int k = 0; k < 6; k++) stress[t + k * lmx] = tau[k];
    for (int k = 0; k < 3; k++) heat[t + k * lmx] = hf[k];
}

// calc_time_step_kernel, src/cuda/kernels/physics.hpp:335-427 (D9 intended:
// FP64 max, lanes past the end and the accumulators start at 0).  Stage 1
// writes one pair of partial maxima per block; stage 2 reduces them.
__global__ void __launch_bounds__(kBlock) time_step_kernel(const double* xim, const double* etm, const double* zem,
                                                           const double* de, const double* yaco, const double* ssk,
                                                           cs_point umf, double* part, long lmx) {
    __shared__ double s0[kBlock], s1[kBlock];
    double m0 = 0.0, m1 = 0.0;
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < lmx; t += (long)gridDim.x * blockDim.x) {
        double rr1 = 0.0;
        for (int d = 0; d < 3; d++) rr1 += xim[t + d * lmx] * xim[t + d * lmx];
        for (int d = 0; d < 3; d++) rr1 += etm[t + d * lmx] * etm[t + d * lmx];
        for (int d = 0; d < 3; d++) rr1 += zem[t + d * lmx] * zem[t + d * lmx];
        double u = de[t + lmx] + umf.x, v = de[t + 2 * lmx] + umf.y, w = de[t + 3 * lmx] + umf.z;
        double rr2 = fabs(xim[t] * u + xim[t + lmx] * v + xim[t + 2 * lmx] * w) +
                     fabs(etm[t] * u + etm[t + lmx] * v + etm[t + 2 * lmx] * w) +
                     fabs(zem[t] * u + zem[t + lmx] * v + zem[t + 2 * lmx] * w);
        double ssi = fabs(yaco[t]);
        m0 = fmax(m0, (sqrt(de[t + 4 * lmx] * rr1) + rr2) * ssi);
        m1 = fmax(m1, de[t] * ssk[t] * rr1 * ssi * ssi);
    }
    s0[threadIdx.x] = m0;
    s1[threadIdx.x] = m1;
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            s0[threadIdx.x] = fmax(s0[threadIdx.x], s0[threadIdx.x + w]);
            s1[threadIdx.x] = fmax(s1[threadIdx.x], s1[threadIdx.x + w]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) { part[2 * blockIdx.x] = s0[0]; part[2 * blockIdx.x + 1] = s1[0]; }
}

// out3 (may be null): also writes dte = min(cfl / max0, 0.5 / max1) and the
// two maxima, so a caller that does not synchronise gets the time step on the device
__global__ void __launch_bounds__(kBlock) max_pairs_kernel(double* part, int n, double cfl, double* out3) {
    __shared__ double s0[kBlock], s1[kBlock];
    double m0 = 0.0, m1 = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) { m0 = fmax(m0, part[2 * i]); m1 = fmax(m1, part[2 * i + 1]); }
    s0[threadIdx.x] = m0;
    s1[threadIdx.x] = m1;
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            s0[threadIdx.x] = fmax(s0[threadIdx.x], s0[threadIdx.x + w]);
            s1[threadIdx.x] = fmax(s1[threadIdx.x], s1[threadIdx.x + w]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[2 * n] = s0[0];
        part[2 * n + 1] = s1[0];
        if (out3) {
            out3[0] = fmin(cfl / s0[0], 0.5 / s1[0]);
            out3[1] = s0[0];
            out3[2] = s1[0];
        }
    }
}

// RK stage update (absent upstream): q = q0 - dtk * yaco * de
__global__ void __launch_bounds__(kBlock) rk_stage_kernel(double* q, const double* q0, const double* de,
                                                          const double* yaco, double dtk, long lmx) {
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= lmx) return;
    double f = dtk * yaco[t];
    for (int v = 0; v < 5; v++) q[t + v * lmx] = q0[t + v * lmx] - f * de[t + v * lmx];
}

// initial state: uniform flow at rest in the frame (stored velocity 0),
// rho0 = 1, p0 = 1/gam, and a Gaussian pressure pulse with rho' = p' (c0 = 1)
__global__ void __launch_bounds__(kBlock) init_pulse_kernel(double* qa, int lxi, int let, long lmx, cs_point cen,
                                                            double amp, double width) {
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= lmx) return;
    const long plane = (long)lxi * let;
    const double z = (double)(t / plane), y = (double)((t % plane) / lxi), x = (double)(t % lxi);
    const double r2 = (x - cen.x) * (x - cen.x) + (y - cen.y) * (y - cen.y) + (z - cen.z) * (z - cen.z);
    const double pp = amp * exp(-M_LN2 * r2 / (width * width));
    qa[t] = 1.0 + pp;
    qa[t + lmx] = 0.0;
    qa[t + 2 * lmx] = 0.0;
    qa[t + 3 * lmx] = 0.0;
    qa[t + 4 * lmx] = (kHam + pp) / kGamm1;
}

// sbcc: 1 on the boundary points of the subdomain, 0 inside
__global__ void __launch_bounds__(kBlock) boundary_weights_kernel(double* sbcc, int lxi, int let, int lze, long lmx) {
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= lmx) return;
    const long plane = (long)lxi * let;
    const long z = t / plane, y = (t % plane) / lxi, x = t % lxi;
    const bool on = x == 0 || x == lxi - 1 || y == 0 || y == let - 1 || z == 0 || z == lze - 1;
    sbcc[t] = on ? 1.0 : 0.0;
}

// fill_gpu_buffer_kernel<Axis>, src/cuda/kernels/numerics_pc.hpp:42-122: for
// every line of the face, the two pbco-weighted 8-point sums from the face
// inward (istart/increment of numerics_pc.hpp:275-280).
__global__ void __launch_bounds__(kBlock) fill_kernel(const double* in, double* send, const double* pbco, int axis,
                                                      int face, int N, long nlines, int lxi, long plane, long st) {
    long f = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nlines) return;
    long b = line_base(axis, f, lxi, plane);
    int istart = face * (N - 1), inc = 1 - 2 * face;
    double v0 = 0.0, v1 = 0.0;
    for (int i = 0; i < kLmd; i++) {
        double x = in[b + (long)(istart + inc * i) * st];
        v0 += pbco[i] * x;
        v1 += pbco[kLmd + i] * x;
    }
    send[f] = v0;
    send[nlines + f] = v1;
}

// ---------------------------------------------------------------------------
int launch_primitives(const double* qa, double* de, double* p, long lmx, hipStream_t s) {
    hipLaunchKernelGGL(primitives_kernel, grid_for(lmx), dim3(kBlock), 0, s, qa, de, p, lmx);
    CS_HIP(hipGetLastError());
    return CS_OK;
}

int launch_fluxes_pre(bool visc, double* buffer, const double* qa, const double* p, double* de, const double* xim,
                      const double* etm, const double* zem, const double* stress, const double* heat,
                      cs_point umf, long lmx, hipStream_t s) {
    if (visc)
        hipLaunchKernelGGL(fluxes_pre_kernel<true>, grid_for(lmx), dim3(kBlock), 0, s, buffer, qa, p, de, xim, etm,
                           zem, stress, heat, umf, lmx);
    else
        hipLaunchKernelGGL(fluxes_pre_kernel<false>, grid_for(lmx), dim3(kBlock), 0, s, buffer, qa, p, de, xim,
                           etm, zem, stress, heat, umf, lmx);
    CS_HIP(hipGetLastError());
    return CS_OK;
}

int launch_fluxes_post(double* de, const double* bd, int m, long lmx, hipStream_t s) {
    hipLaunchKernelGGL(fluxes_post_kernel, grid_for(lmx), dim3(kBlock), 0, s, de, bd, m, lmx);
    CS_HIP(hipGetLastError());
    return CS_OK;
}

int launch_visc_init(double* buffer, const double* de, long lmx, hipStream_t s) {
    hipLaunchKernelGGL(visc_init_kernel, grid_for(lmx), dim3(kBlock), 0, s, buffer, de, lmx);
    CS_HIP(hipGetLastError());
    return CS_OK;
}

int launch_visc_post(double* stress, double* heat, double* ss, const double* bd, const double* xim,
                     const double* etm, const double* zem, int m, long lmx, hipStream_t s) {
    hipLaunchKernelGGL(visc_post_kernel, grid_for(lmx), dim3(kBlock), 0, s, stress, heat, ss, bd, xim, etm, zem, m,
                       lmx);
    CS_HIP(hipGetLastError());
    return CS_OK;
}

int launch_visc_final(double* de, const double* ssk, const double* yaco, const double* ss, double* stress,
                      double* heat, long lmx, hipStream_t s) {
    hipLaunchKernelGGL(visc_final_kernel, grid_for(lmx), dim3(kBlock), 0, s, de, ssk, yaco, ss, stress, heat, lmx);
    CS_HIP(hipGetLastError());
    return CS_OK;
}

int time_step_blocks(long lmx) {
    long b = (lmx + kBlock - 1) / kBlock;
    return (int)(b < 2048 ? b : 2048);
}

int launch_time_step(const double* xim, const double* etm, const double* zem, const double* de, const double* yaco,
                     const double* ssk, cs_point umf, double* part, int nblocks, long lmx, hipStream_t s) {
    hipLaunchKernelGGL(time_step_kernel, dim3(nblocks), dim3(kBlock), 0, s, xim, etm, zem, de, yaco, ssk, umf, part,
                       lmx);
    CS_HIP(hipGetLastError());
    hipLaunchKernelGGL(max_pairs_kernel, dim3(1), dim3(kBlock), 0, s, part, nblocks, 0.0, (double*)nullptr);
    CS_HIP(hipGetLastError());
    return CS_OK;
}

int launch_rk_stage(double* q, const double* q0, const double* de, const double* yaco, double dtk, long lmx,
                    hipStream_t s) {
    hipLaunchKernelGGL(rk_stage_kernel, grid_for(lmx), dim3(kBlock), 0, s, q, q0, de, yaco, dtk, lmx);
    CS_HIP(hipGetLastError());
    return CS_OK;
}

int launch_fill(const double* in, double* send, const double* pbco, int axis, int face, int N, long nlines, int lxi,
                long plane, long st, hipStream_t s) {
    hipLaunchKernelGGL(fill_kernel, grid_for(nlines), dim3(kBlock), 0, s, in, send, pbco, axis, face, N, nlines, lxi,
                       plane, st);
    CS_HIP(hipGetLastError());
    return CS_OK;
}

int launch_init_pulse(const cs_dims* d, cs_point cen, double amp, double width, double* qa, hipStream_t s) {
    const long lmx = (long)d->lxi * d->let * d->lze;
    hipLaunchKernelGGL(init_pulse_kernel, grid_for(lmx), dim3(kBlock), 0, s, qa, d->lxi, d->let, lmx, cen, amp, width);
    CS_HIP(hipGetLastError());
    return CS_OK;
}

int launch_boundary_weights(const cs_dims* d, double* sbcc, hipStream_t s) {
    const long lmx = (long)d->lxi * d->let * d->lze;
    hipLaunchKernelGGL(boundary_weights_kernel, grid_for(lmx), dim3(kBlock), 0, s, sbcc, d->lxi, d->let, d->lze, lmx);
    CS_HIP(hipGetLastError());
    return CS_OK;
}

// ---------------------------------------------------------------------------
// allocation-free time step
// ---------------------------------------------------------------------------
// Scratch of cs_calc_time_step_noalloc: one pair of partial maxima per block
// of time_step_kernel and the reduced pair (kTsMaxBlocks pairs + 1, 32.8 KB).
// Each GPU holds its own copy from the library's load until the process ends,
// shared by every call on that GPU.  Nothing in it needs clearing: every block
// writes its pair before max_pairs_kernel reads it.  In host code the name is
// a stand-in, never a device address: kernels get the per-device pointer
// from hipGetSymbolAddress.
constexpr int kTsMaxBlocks = 2048;
static __device__ __align__(256) double scratch[2 * kTsMaxBlocks + 2];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch[kMaxDevices] = {};
}  // namespace

// device address of `scratch` on the current device, looked up on the first
// call on each device
static int scratch_on_device(double** out) {
    int dev = 0;
    CS_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return fail(CS_ERR_UNSUPPORTED, "device ordinal above 63");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch[dev]) {
        void* p = nullptr;
        CS_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return fail(CS_ERR_HIP, "hipGetSymbolAddress returned a null address for the time-step scratch");
        g_scratch[dev] = static_cast<double*>(p);
    }
    *out = g_scratch[dev];
    return CS_OK;
}

}  // namespace cs

using namespace cs;

extern "C" int cs_calc_time_step_noalloc(const cs_dims* dims, cs_point umf, const double* xim, const double* etm,
                                         const double* zem, const double* de, const double* yaco, const double* ssk,
                                         double cfl, double* out3, cs_stream stream) {
    if (!dims || !out3) return fail(CS_ERR_ARG, "null argument");
    if (!xim || !etm || !zem || !de || !yaco || !ssk) return fail(CS_ERR_ARG, "null field pointer");
    if (dims->lxi < 1 || dims->let < 1 || dims->lze < 1) return fail(CS_ERR_SHAPE, "grid dimensions must be positive");
    const long lmx = (long)dims->lxi * dims->let * dims->lze;
    const int nblocks = time_step_blocks(lmx);
    if (nblocks > kTsMaxBlocks)
        return fail(CS_ERR_SHAPE, "time-step reduction needs " + std::to_string(nblocks) +
                                      " partial pairs; the static scratch holds " + std::to_string(kTsMaxBlocks));
    double* part = nullptr;
    int rc = scratch_on_device(&part);
    if (rc) return rc;
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(time_step_kernel, dim3(nblocks), dim3(kBlock), 0, s, xim, etm, zem, de, yaco, ssk, umf, part,
                       lmx);
    CS_HIP(hipGetLastError());
    hipLaunchKernelGGL(max_pairs_kernel, dim3(1), dim3(kBlock), 0, s, part, nblocks, cfl, out3);
    CS_HIP(hipGetLastError());
    return CS_OK;
}

extern "C" int cs_init_pulse(const cs_dims* dims, cs_point center, double amp, double width, double* qa,
                             cs_stream stream) {
    if (!dims || !qa) return fail(CS_ERR_ARG, "null argument");
    if (dims->lxi < 1 || dims->let < 1 || dims->lze < 1) return fail(CS_ERR_SHAPE, "grid dimensions must be positive");
    if (!(width > 0.0) || !std::isfinite(amp)) return fail(CS_ERR_ARG, "pulse width must be > 0 and amp finite");
    return launch_init_pulse(dims, center, amp, width, qa, (hipStream_t)stream);
}

extern "C" int cs_boundary_weights(const cs_dims* dims, double* sbcc, cs_stream stream) {
    if (!dims || !sbcc) return fail(CS_ERR_ARG, "null argument");
    if (dims->lxi < 1 || dims->let < 1 || dims->lze < 1) return fail(CS_ERR_SHAPE, "grid dimensions must be positive");
    return launch_boundary_weights(dims, sbcc, (hipStream_t)stream);
}
