// fused_kernels.hpp -- the sweep functors, kernels and launchers of the
// fused right-hand side (see fused.hip for the design), shared by the
// translation units that instantiate them: fused.hip (FP64 intermediates) and
// fused_f32.hip (the viscous intermediates stored in FP32).
#pragma once

#include <type_traits>

#include "sweep_parts.hpp"

namespace cs {
namespace fused {

// loads and stores of a field of storage type TS at the byte offset of the
// point in an FP64 field (an FP32 field holds the same points at half the offset)
__device__ __forceinline__ double ldv(const double* p, uoff bo) { return ldg(p, bo); }
__device__ __forceinline__ double ldv(const float* p, uoff bo) {
    return (double)*reinterpret_cast<const float*>(reinterpret_cast<const char*>(p) + (bo >> 1));
}
__device__ __forceinline__ void stv(double* p, uoff bo, double v) { stg(p, bo, v); }
__device__ __forceinline__ void stv(float* p, uoff bo, double v) {
    *reinterpret_cast<float*>(reinterpret_cast<char*>(p) + (bo >> 1)) = (float)v;
}


// ---------------------------------------------------------------------------
// sweep functors
// ---------------------------------------------------------------------------
// S1: velocity and c^2 gradients along AX.  The first two S1 sweeps of an
// RHS store their index-space gradients; the third (CLOSE) combines its own
// gradients with the stored ones into the stress tensor and heat flux.
template <int AX, bool CLOSE, class TS, bool CURV = false, bool DT = false>
struct S1T {
    typedef TS Store;                      // storage type of the gradients, stress and heat flux
    static constexpr bool kCurv = CURV;    // per-point metrics (curvilinear solvers)
    static constexpr int NF = 4;
    static constexpr bool kGcbc = false;   // gradients only; BC updates act on the flux sweeps
    static constexpr bool kPv = false;     // kernels without an LDS stash re-derive u, v, w from q at the store
    // the transposing x kernel keeps u, v, w of its points in LDS for the
    // closure (store_pv) instead of reading q again
    static constexpr int kStash = CLOSE ? 3 : 0;
    static constexpr int kH = 2;           // derivative stencil half-width
    static constexpr int kWpe = 3;         // waves per SIMD of the y/z kernels (168 VGPRs)
    static constexpr bool kFilter = false;
    // DT: a sweep that also reduces the time step's maxima over the points it
    // evaluates, one pair per workgroup: a y/z gradient sweep
    // (constant metrics), or the closing x sweep with per-point metrics, which
    // loads the metric rows and J of every point anyway
    static constexpr bool kDt = DT;
    static_assert(!DT || (!CLOSE && !CURV && AX != 0) || (CLOSE && CURV && AX == 0),
                  "time-step maxima: y/z gradient sweeps of constant metrics, closing x sweep of per-point metrics");
    const double* q[5];
    TS* gout[4];              // !CLOSE: d/d(axis AX) of u, v, w, c^2
    const TS* gin[3][4];      // CLOSE: stored gradients along the other two axes
    TS* tau[6];               // xx, yy, zz, xy, yz, zx
    TS* hf[3];
    Phys ph;
    Metric mt;                // CURV: the point's metric rows and Jacobian
    double* dtpart;           // kDt: the workgroup's two time-step maxima [2*blockIdx.x]

    __device__ __forceinline__ void point(uoff bo, double* v) const {
        double qq[5];
        Prim p = prim_at(q, bo, qq);
        v[0] = p.u;
        v[1] = p.v;
        v[2] = p.w;
        v[3] = p.c2;
    }
    // point() and the largest 1/rho so far (kDt)
    __device__ __forceinline__ void point_dt(uoff bo, double* v, double& rhoi_max) const {
        double qq[5];
        Prim p = prim_at(q, bo, qq);
        v[0] = p.u;
        v[1] = p.v;
        v[2] = p.w;
        v[3] = p.c2;
        rhoi_max = fmax(rhoi_max, p.rhoi);
    }
    __device__ __forceinline__ void fetch(uoff, double*) const {}
    __device__ __forceinline__ void store(uoff bo, const double* d, const double* pv) const {
        if constexpr (!CLOSE) {
#pragma unroll
            for (int f = 0; f < 4; f++) stv(gout[f], bo, d[f]);
        } else {
            double qq[5];
            const Prim p = prim_at(q, bo, qq);      // q of this point: just read by this sweep
            const double uvw[3] = {p.u, p.v, p.w};
            close(bo, d, uvw);
        }
    }
    // the store of the closing sweep with u, v, w of the point supplied
    __device__ __forceinline__ void store_pv(uoff bo, const double* d, const double* uvw) const {
        if constexpr (CLOSE) close(bo, d, uvw);
        else store(bo, d, nullptr);
    }
    // the same, updating the time step's maxima mm[0..1] at the point (kDt, CURV)
    __device__ __forceinline__ void store_pv_dt(uoff bo, const double* d, const double* uvw, double* mm) const {
        close_dt<true>(bo, d, uvw, mm);
    }
    __device__ __forceinline__ void close(uoff bo, const double* d, const double* uvw) const {
        close_dt<false>(bo, d, uvw, nullptr);
    }
    template <bool WITH_DT>
    __device__ __forceinline__ void close_dt(uoff bo, const double* d, const double* uvw, double* mm) const {
        double gr[3][4];
#pragma unroll
        for (int a = 0; a < 3; a++)
#pragma unroll
            for (int f = 0; f < 4; f++) gr[a][f] = (a == AX) ? d[f] : ldv(gin[a][f], bo);
        // chain rule of calc_viscous_shear_stress_post_compute_kernel
        // (src/cuda/kernels/physics.hpp:267-281): d/dx_k = sum_a m_a[k] d/dxi_a
        double g[9], c2g[3], J = ph.J;
        double M[3][3];
        if constexpr (CURV) {
#pragma unroll
            for (int a = 0; a < 3; a++)
#pragma unroll
                for (int k = 0; k < 3; k++) M[a][k] = ldg(mt.m[a][k], bo);
#pragma unroll
            for (int f = 0; f < 3; f++)
#pragma unroll
                for (int k = 0; k < 3; k++) g[3 * f + k] = M[0][k] * gr[0][f] + M[1][k] * gr[1][f] + M[2][k] * gr[2][f];
#pragma unroll
            for (int k = 0; k < 3; k++) c2g[k] = M[0][k] * gr[0][3] + M[1][k] * gr[1][3] + M[2][k] * gr[2][3];
            J = ldg(mt.J, bo);
        } else {
            // diagonal metrics: m_a[k] = m[a] for k = a
#pragma unroll
            for (int f = 0; f < 3; f++)
#pragma unroll
                for (int a = 0; a < 3; a++) g[3 * f + a] = ph.m[a] * gr[a][f];
#pragma unroll
            for (int a = 0; a < 3; a++) c2g[a] = ph.m[a] * gr[a][3];
        }
        double t6[6], h3[3], d4;
        viscous_closure(ph.mu, J, g, c2g, uvw, t6, h3, &d4);
#pragma unroll
        for (int k = 0; k < 6; k++) stv(tau[k], bo, t6[k]);
#pragma unroll
        for (int k = 0; k < 3; k++) stv(hf[k], bo, h3[k]);
        if constexpr (WITH_DT && CURV) {
            // the point's q again (this sweep read it for the gradients)
            double qq[5];
            dt_curv_point(M, J, ph, prim_at(q, bo, qq), mm[0], mm[1]);
        }
    }
    __device__ __forceinline__ void gcbc(uoff, int, const double*, double*) const {}
};

// the gradient sweeps with FP64 intermediates, and with FP32 ones
// (cs_solver_set_precision 1); the names the kernels carry
template <int AX, bool CLOSE>
struct S1 : S1T<AX, CLOSE, double> {};
template <int AX, bool CLOSE>
struct S1F : S1T<AX, CLOSE, float> {};
// curvilinear solvers (cs_solver_set_metrics): FP64 intermediates
template <int AX, bool CLOSE>
struct S1C : S1T<AX, CLOSE, double, true> {};
template <int AX, bool CLOSE>
struct S1CF : S1T<AX, CLOSE, float, true> {};
// the first gradient sweep of an RK4 step, reducing the time step as well
template <int AX, bool CLOSE>
struct S1D : S1T<AX, CLOSE, double, false, true> {};
template <int AX, bool CLOSE>
struct S1FD : S1T<AX, CLOSE, float, false, true> {};
// the closing sweep of an RK4 step's first stage with per-point metrics,
// reducing the time step as well
template <int AX, bool CLOSE>
struct S1CD : S1T<AX, CLOSE, double, true, true> {};
template <int AX, bool CLOSE>
struct S1CFD : S1T<AX, CLOSE, float, true, true> {};

// S2: flux divergence along AX.  MODE 0: write the partial RHS; 1: add to it;
// 2: add and write the RHS; 3: add and write q0 - dtk*J*RHS (RK stage).
template <int AX, int MODE, bool VISC, class TS, bool CURV = false>
struct S2T {
    typedef TS Store;
    static constexpr bool kCurv = CURV;    // per-point metrics (curvilinear solvers)
    static constexpr int NF = 5;
    static constexpr bool kGcbc = true;
    static constexpr bool kPv = false;
    static constexpr int kStash = 0;
    static constexpr int kH = 2;
    static constexpr int kWpe = 3;
    static constexpr bool kFilter = false;
    const double* q[5];
    const TS* tr[3];          // row AX of the stress tensor
    const TS* hx;             // heat flux component AX
    const TS* tau6[6];        // CURV: the whole stress tensor (xx, yy, zz, xy, yz, zx)
    const TS* h3[3];          // CURV: the whole heat flux
    Metric mt;                // CURV: the point's metric rows, Jacobian, face normals
    const double* din[5];
    double* dout[5];
    const double* q0[5];
    const double* dtp;
    int div;
    Phys ph;
    double* rec;              // [2 face][5 var][face] boundary derivatives of this sweep, or null
    long nface;
    int lxi;
    long plane;

    // the AX row of calc_fluxes_pre_compute_kernel with metric row m*e_AX
    // (src/cuda/kernels/physics.hpp:98-191); zero metric terms dropped
    __device__ __forceinline__ void point(uoff bo, double* F) const {
        double qq[5];
        Prim p = prim_at(q, bo, qq);
        if constexpr (CURV) {
            // the AX row with the point's metric row (m0, m1, m2)
            const double m0 = ldg(mt.m[AX][0], bo), m1 = ldg(mt.m[AX][1], bo), m2 = ldg(mt.m[AX][2], bo);
            const double ss = m0 * (p.u + ph.umf.x) + m1 * (p.v + ph.umf.y) + m2 * (p.w + ph.umf.z);
            F[0] = qq[0] * ss;
            F[1] = qq[1] * ss + m0 * p.p;
            F[2] = qq[2] * ss + m1 * p.p;
            F[3] = qq[3] * ss + m2 * p.p;
            F[4] = (qq[4] + p.p) * ss - p.p * (ph.umf.x * m0 + ph.umf.y * m1 + ph.umf.z * m2);
            if (VISC) {
                const double txx = ldv(tau6[0], bo), tyy = ldv(tau6[1], bo), tzz = ldv(tau6[2], bo);
                const double txy = ldv(tau6[3], bo), tyz = ldv(tau6[4], bo), tzx = ldv(tau6[5], bo);
                F[1] -= m0 * txx + m1 * txy + m2 * tzx;
                F[2] -= m0 * txy + m1 * tyy + m2 * tyz;
                F[3] -= m0 * tzx + m1 * tyz + m2 * tzz;
                F[4] -= m0 * ldv(h3[0], bo) + m1 * ldv(h3[1], bo) + m2 * ldv(h3[2], bo);
            }
            return;
        }
        const double m = ph.m[AX];
        const double vel = (AX == 0) ? p.u : (AX == 1 ? p.v : p.w);
        const double uf = (AX == 0) ? ph.umf.x : (AX == 1 ? ph.umf.y : ph.umf.z);
        const double ss = m * (vel + uf);
        F[0] = qq[0] * ss;
        F[1] = qq[1] * ss;
        F[2] = qq[2] * ss;
        F[3] = qq[3] * ss;
        F[1 + AX] += m * p.p;
        F[4] = (qq[4] + p.p) * ss - p.p * (uf * m);
        if (VISC) {
            F[1] -= m * ldv(tr[0], bo);
            F[2] -= m * ldv(tr[1], bo);
            F[3] -= m * ldv(tr[2], bo);
            F[4] -= m * ldv(hx, bo);
        }
    }
    // the epilogue inputs of one row folded into d (loads only); store() then
    // only writes.  The kernels fetch all their rows before their first store:
    // dout may alias din or q0 for all the compiler knows, so a load placed
    // after a store waits for it, and interleaved they would run one memory
    // round trip per field and row.
    __device__ __forceinline__ void fetch(uoff bo, double* d) const {
        if (MODE == 1 || MODE == 2) {
#pragma unroll
            for (int f = 0; f < 5; f++) d[f] = ldg(din[f], bo) + d[f];
        } else if (MODE == 3) {
            const double dtk = dtp[0] / (double)div;
            const double fj = dtk * (CURV ? ldg(mt.J, bo) : ph.J);
#pragma unroll
            for (int f = 0; f < 5; f++) {
                const double de = ldg(din[f], bo) + d[f];
                d[f] = ldg(q0[f], bo) - fj * de;
            }
        }
    }
    __device__ __forceinline__ void store(uoff bo, const double* d, const double*) const {
#pragma unroll
        for (int f = 0; f < 5; f++) stg(dout[f], bo, d[f]);
    }
    // non-reflective characteristic correction at a face point from the
    // boundary derivatives of this sweep (gcbc_setup_kernel +
    // gcbc_update_non_reflective_kernel, src/cuda/kernels/gcbc.hpp:251-454)
    __device__ __forceinline__ void gcbc(uoff bo, int face, const double* drv, double* corr) const {
        // face index of the face point (x: k*let+j, y: k*lxi+i, z: j*lxi+i)
        const long o = bo / 8u;
        const long fi = AX == 0 ? o / lxi : (AX == 1 ? (o / plane) * lxi + o % lxi : o % plane);
        if (rec) {
#pragma unroll
            for (int v = 0; v < 5; v++) rec[(face * 5 + v) * nface + fi] = drv[v];
        }
        double qq[5];
        Prim p = prim_at(q, bo, qq);
        double cm[3], J = ph.J;
        if constexpr (CURV) {
            // the caller's face normals ([2 face][3][face], as gcbc_go takes them)
            const double* c = mt.cm[AX] + (long)face * 3 * nface + fi;
            cm[0] = c[0];
            cm[1] = c[nface];
            cm[2] = c[2 * nface];
            J = ldg(mt.J, bo);
        } else {
            // unit normal of the face: grad(xi_AX) / |grad(xi_AX)|, grad(xi_AX) = J * metric * e_AX
            const double sg = ph.m[AX] * ph.J < 0.0 ? -1.0 : 1.0;
            cm[0] = AX == 0 ? sg : 0.0;
            cm[1] = AX == 1 ? sg : 0.0;
            cm[2] = AX == 2 ? sg : 0.0;
        }
        Eig e;
        e.eleme(cm, qq[0], qq + 1, p.p, ph.umf);
        double cha[5], ch[5];
#pragma unroll
        for (int v = 0; v < 5; v++) cha[v] = J * drv[v];
        e.l_apply(cm, cha, ch);
        const double ra0 = 1 - 2 * face;
        double dl[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
        if (ra0 * (e.vn + e.vs + e.ao) > 0) dl[3] = -ch[3];
        if (ra0 * (e.vn + e.vs - e.ao) > 0) dl[4] = -ch[4];
        e.r_apply(cm, dl, corr);
        // D25: a change of J*de (the amplitudes were formed from J*drv), so de changes by corr / J
#pragma unroll
        for (int v = 0; v < 5; v++) corr[v] /= J;
    }
};

template <int AX, int MODE, bool VISC>
struct S2 : S2T<AX, MODE, VISC, double> {};
template <int AX, int MODE, bool VISC>
struct S2F : S2T<AX, MODE, VISC, float> {};
template <int AX, int MODE, bool VISC>
struct S2C : S2T<AX, MODE, VISC, double, true> {};
template <int AX, int MODE, bool VISC>
struct S2CF : S2T<AX, MODE, VISC, float, true> {};

// SF: the compact filter of the five conserved variables along AX, in place
// (the filter of cs_filter: 10th-order interior rows, reduced-order centred
// rows with graded alpha near closure ends, identity end rows; cyclic on seam
// axes)
template <int AX>
struct SF {
    static constexpr int NF = 5;
    static constexpr bool kGcbc = false;
    static constexpr bool kPv = false;
    static constexpr int kStash = 0;
    static constexpr int kH = 5;
    static constexpr int kWpe = 3;
    static constexpr bool kFilter = true;
    double* q[5];
    __device__ __forceinline__ void point(uoff bo, double* v) const {
#pragma unroll
        for (int f = 0; f < 5; f++) v[f] = ldg(q[f], bo);
    }
    __device__ __forceinline__ void fetch(uoff, double*) const {}
    __device__ __forceinline__ void store(uoff bo, const double* d, const double*) const {
#pragma unroll
        for (int f = 0; f < 5; f++) stg(q[f], bo, d[f]);
    }
    __device__ __forceinline__ void gcbc(uoff, int, const double*, double*) const {}
};

// SF1: the same filter, one conserved variable per workgroup (blockIdx.y picks
// the variable of the five stored lmx points apart); a fifth of SF's registers
template <int AX>
struct SF1 {
    static constexpr int NF = 1;
    static constexpr bool kGcbc = false;
    static constexpr bool kPv = false;
    static constexpr int kStash = 0;
    static constexpr int kH = 5;
    static constexpr int kWpe = 3;
    static constexpr bool kFilter = true;
    double* q;
    long lmx;
    __device__ __forceinline__ double* field() const { return q + (long)blockIdx.y * lmx; }
    __device__ __forceinline__ void point(uoff bo, double* v) const { v[0] = ldg(field(), bo); }
    __device__ __forceinline__ void fetch(uoff, double*) const {}
    __device__ __forceinline__ void store(uoff bo, const double* d, const double*) const { stg(field(), bo, d[0]); }
    __device__ __forceinline__ void gcbc(uoff, int, const double*, double*) const {}
};

// ---------------------------------------------------------------------------
// pieces shared by the sweep kernels
// ---------------------------------------------------------------------------
// interior rows and closure rows from a window w[f][u] holding rows s-2+u
// (build_system, numerics_device_pc.hpp:176-293)
template <int K, int NF, int W>
__device__ __forceinline__ void window_rhs(const double (&w)[NF][W], double (&r)[NF][K], int L, bool first,
                                           bool last) {
#pragma unroll
    for (int f = 0; f < NF; f++)
#pragma unroll
        for (int t = 0; t < K; t++)
            r[f][t] = kAa * (w[f][t + 3] - w[f][t + 1]) + kBb * (w[f][t + 4] - w[f][t]);
    if (first) {
#pragma unroll
        for (int f = 0; f < NF; f++) {
            r[f][0] = kAb00 * w[f][2] + kAb01 * w[f][3] + kAb02 * w[f][4];
            r[f][1] = kAb10 * (w[f][4] - w[f][2]);
        }
    }
    if (last) {
#pragma unroll
        for (int f = 0; f < NF; f++) {
            double fN1 = 0.0, fN2 = 0.0, fN3 = 0.0;
#pragma unroll
            for (int u = 0; u < W; u++) {
                if (u == L + 1) fN1 = w[f][u];
                if (u == L) fN2 = w[f][u];
                if (u == L - 1) fN3 = w[f][u];
            }
            const double r1 = -(kAb00 * fN1 + kAb01 * fN2 + kAb02 * fN3);
            const double r2 = -kAb10 * (fN3 - fN1);
#pragma unroll
            for (int t = 0; t < K; t++) {
                if (t == L - 1) r[f][t] = r1;
                if (t == L - 2) r[f][t] = r2;
            }
        }
    }
}

// filter rows from a window w[f][u] holding rows s-5+u: 10th order inside,
// order 2*dist at distance dist < 5 from a closure end, identity at the end
// (an = the order-2M coefficient sets, an[6M..6M+M], each for its row's
// alpha; the alphas enter through the LU tables); cyclic lines are all
// inside rows
template <int K, int NF, int W>
__device__ __forceinline__ void filter_rhs(const double (&w)[NF][W], double (&r)[NF][K], int s, int N, bool cyclic,
                                           const double* an) {
    constexpr int H = 5;
#pragma unroll
    for (int t = 0; t < K; t++) {
        const int j = s + t;
        const int dist = cyclic ? H : min(j, N - 1 - j);
#pragma unroll
        for (int f = 0; f < NF; f++) {
            double v = an[30] * w[f][t + H];
#pragma unroll
            for (int m = 1; m <= 5; m++) v += 0.5 * an[30 + m] * (w[f][t + H + m] + w[f][t + H - m]);
            r[f][t] = v;
        }
        if (dist < H) {
#pragma unroll
            for (int f = 0; f < NF; f++) {
                double v = w[f][t + H];
                if (dist > 0) {
                    const double* cf = an + 6 * dist;
                    v = cf[0] * w[f][t + H];
#pragma unroll
                    for (int m = 1; m < H; m++)
                        if (m <= dist) v += 0.5 * cf[m] * (w[f][t + H + m] + w[f][t + H - m]);
                }
                r[f][t] = v;
            }
        }
    }
}

// Sherman-Morrison correction of a cyclic filter solve: y0 and yN1 are the
// solution of the modified system at rows 0 and N-1 (tab layout of
// build_line_table: z at 5N, 1/(1 + z0 + kappa zN1) at 6N, kappa at 6N+1)
template <int K, int NF>
__device__ __forceinline__ void cyclic_correction(const double* tab, int N, int s, const double (&y0)[NF],
                                                  const double (&yN1)[NF], double (&r)[NF][K]) {
    const double* z = tab + 5 * N;
    const double inv = tab[6 * N], kappa = tab[6 * N + 1];
#pragma unroll
    for (int f = 0; f < NF; f++) {
        const double beta = (y0[f] + kappa * yN1[f]) * inv;
#pragma unroll
        for (int t = 0; t < K; t++) r[f][t] -= beta * z[s + t];
    }
}

// the same rows as window_rhs, streamed through a 5-row register window so
// that only the rows a right-hand side needs are live (rows from an LDS tile)
template <int K, int NF, class ROW>
__device__ __forceinline__ void stream_rhs(ROW&& row_of, double (&r)[NF][K], int L, bool first, bool last) {
    constexpr int W = K + 4;
    double w0[NF], w1[NF], w2[NF], w3[NF], w4[NF];
    double f0[NF], f1[NF], f2[NF], fN1[NF], fN2[NF], fN3[NF];
#pragma unroll
    for (int f = 0; f < NF; f++) {
        w0[f] = w1[f] = w2[f] = w3[f] = w4[f] = 0.0;
        f0[f] = f1[f] = f2[f] = fN1[f] = fN2[f] = fN3[f] = 0.0;
    }
#pragma unroll
    for (int u = 0; u < W; u++) {
        double v[NF];
        row_of(u, v);
#pragma unroll
        for (int f = 0; f < NF; f++) { w0[f] = w1[f]; w1[f] = w2[f]; w2[f] = w3[f]; w3[f] = w4[f]; w4[f] = v[f]; }
        if (u >= 4) {
            const int t = u - 4;
#pragma unroll
            for (int f = 0; f < NF; f++) r[f][t] = kAa * (w3[f] - w1[f]) + kBb * (w4[f] - w0[f]);
        }
        if (u == 4) {
#pragma unroll
            for (int f = 0; f < NF; f++) { f0[f] = w2[f]; f1[f] = w3[f]; f2[f] = w4[f]; }
        }
#pragma unroll
        for (int f = 0; f < NF; f++) {
            if (u == L + 1) fN1[f] = w4[f];
            if (u == L) fN2[f] = w4[f];
            if (u == L - 1) fN3[f] = w4[f];
        }
    }
    if (first) {
#pragma unroll
        for (int f = 0; f < NF; f++) {
            r[f][0] = kAb00 * f0[f] + kAb01 * f1[f] + kAb02 * f2[f];
            r[f][1] = kAb10 * (f2[f] - f0[f]);
        }
    }
    if (last) {
#pragma unroll
        for (int f = 0; f < NF; f++) {
            const double r1 = -(kAb00 * fN1[f] + kAb01 * fN2[f] + kAb02 * fN3[f]);
            const double r2 = -kAb10 * (fN3[f] - fN1[f]);
#pragma unroll
            for (int t = 0; t < K; t++) {
                if (t == L - 1) r[f][t] = r1;
                if (t == L - 2) r[f][t] = r2;
            }
        }
    }
}

// quasi-disjoint seam rows, wrapped onto the opposite end of the line
// (build_system seam branch, numerics_device_pc.hpp:199-282, with the
// neighbour sums formed as fill_gpu_buffer_kernel does, kernels/numerics_pc.hpp:86-121)
template <int K, int NF, class SW>
__device__ __forceinline__ void seam_rows(const SW& sw, uoff base, uoff st, int N, int L, bool first, bool last,
                                          const double* pb, double (&r)[NF][K]) {
    const double* pbco = pb;
    const double* pbci = pb + 16;
#pragma unroll
    for (int end = 0; end < 2; end++) {
        if ((end == 0 && !first) || (end == 1 && !last)) continue;
        double s0[NF], s1[NF], o0[NF], o1[NF];
#pragma unroll
        for (int f = 0; f < NF; f++) { s0[f] = 0.0; s1[f] = 0.0; o0[f] = 0.0; o1[f] = 0.0; }
        for (int i = 0; i < kLmd; i++) {
            double v[NF];
            sw.point(base + (uoff)(end == 0 ? i : N - 1 - i) * st, v);
#pragma unroll
            for (int f = 0; f < NF; f++) { s0[f] += pbci[i] * v[f]; s1[f] += pbci[kLmd + i] * v[f]; }
        }
        for (int i = 0; i < kLmd; i++) {
            double v[NF];
            sw.point(base + (uoff)(end == 0 ? N - 1 - i : i) * st, v);
#pragma unroll
            for (int f = 0; f < NF; f++) { o0[f] += pbco[i] * v[f]; o1[f] += pbco[kLmd + i] * v[f]; }
        }
#pragma unroll
        for (int f = 0; f < NF; f++) {
            if (end == 0) {
                r[f][0] = s0[f] + o0[f];
                r[f][1] = s1[f] + o1[f];
            } else {
                const double r1 = -(s0[f] + o0[f]);
                const double r2 = -(s1[f] + o1[f]);
#pragma unroll
                for (int t = 0; t < K; t++) {
                    if (t == L - 1) r[f][t] = r1;
                    if (t == L - 2) r[f][t] = r2;
                }
            }
        }
    }
}

// characteristic correction of a face (first or last chunk of a line, never
// both: lines have at least two chunks), weighted by sbcc on the kMbci rows
// next to the face; the face derivative is the uncorrected solve result
template <int K, int NF, class SW, bool GCBC, bool EX>
__device__ __forceinline__ void face_corrections(const SW& sw, uoff base, uoff st, int N, int s, int L, bool first,
                                                 bool last, double (&r)[NF][K], const double* sbcc) {
    if constexpr (GCBC) {
        if (first || last) {
            const int face = last ? 1 : 0;
            double drv[NF], corr[NF];
#pragma unroll
            for (int f = 0; f < NF; f++) {
                double d = r[f][0];
#pragma unroll
                for (int t = 1; t < K; t++)
                    if (last && ((EX && t == K - 1) || (!EX && t == L - 1))) d = r[f][t];
                drv[f] = d;
            }
            sw.gcbc(base + (uoff)(face ? N - 1 : 0) * st, face, drv, corr);
#pragma unroll
            for (int t = 0; t < K; t++) {
                const int row = s + t;
                const bool in = (EX || t < L) && (face ? row >= N - kMbci : row < kMbci);
                if (in) {
                    const double wgt = ldg(sbcc, base + (uoff)row * st);
#pragma unroll
                    for (int f = 0; f < NF; f++) r[f][t] += wgt * corr[f];
                }
            }
        }
    }
}

// ---------------------------------------------------------------------------
// variant "direct" (equal chunks, N = P*K): each thread evaluates the
// pointwise inputs of its own K rows once; the two halo rows on each side come
// from the neighbouring chunks (shuffles within a wave segment for x lines,
// LDS for y/z lines)
// ---------------------------------------------------------------------------
// functors that reduce the time step's maxima: value = in the point
// evaluation of a y/z sweep (constant metrics), close = in the closing x
// sweep's stores (per-point metrics)
template <class SW, class = void>
struct DtFold {
    static constexpr bool value = false;
    static constexpr bool close = false;
};
template <class SW>
struct DtFold<SW, std::void_t<decltype(SW::kDt)>> {
    static constexpr bool value = SW::kDt && !SW::kCurv;
    static constexpr bool close = SW::kDt && SW::kCurv;
};

template <int K, int AX, int TL, class SW, bool GCBC>
__global__ void __launch_bounds__(AX == 0 ? 256 : TL * kMaxP)
__attribute__((amdgpu_waves_per_eu(AX == 0 ? 1 : (K == 16 ? 2 : SW::kWpe))))
direct_kernel(AxisDesc ax, SweepCommon cm, SW sw) {
    constexpr int NF = SW::NF;
    constexpr int H = SW::kH;
    constexpr int W = K + 2 * H;
    static_assert(H <= K, "halo rows come from the neighbouring chunks only");
    extern __shared__ double lds[];

    const int N = ax.N, P = ax.P;
    int c, lloc = 0, GS = 1;
    long line;
    if (AX == 0) {
        GS = ax.GS;
        const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
        const int lpw = 64 / GS;
        c = lane % GS;
        line = (long)blockIdx.x * (4 * lpw) + wave * lpw + lane / GS;
    } else {
        lloc = threadIdx.x % TL;
        c = threadIdx.x / TL;
        line = group_of(ax.xcd) * TL + lloc;
    }
    const bool act = line < ax.nlines;
    const bool cv = c < P;
    const bool work = act && cv;
    const int s = c * K;
    const uoff st = (uoff)ax.st * 8u;
    const uoff base = line_off(AX, act ? (uoff)line : 0u, (uoff)cm.lxi, (uoff)cm.plane);
    const double* tab = SW::kFilter ? ax.ftab : ax.tab;
    const bool cyc = SW::kFilter && ax.seam;

    // time-step maxima (DtFold functors; every lane of the workgroup's waves
    // is a thread, the launcher checks)
    constexpr bool kDt = DtFold<SW>::value;
    double dm0 = 0.0, dm1 = 0.0;
    double w[NF][W];
#pragma unroll
    for (int t = 0; t < K; t++) {
        double x[NF];
        if (work) {
            if constexpr (kDt) sw.point_dt(base + (uoff)(s + t) * st, x, dm1);
            else sw.point(base + (uoff)(s + t) * st, x);
        } else {
#pragma unroll
            for (int f = 0; f < NF; f++) x[f] = 0.0;
        }
#pragma unroll
        for (int f = 0; f < NF; f++) w[f][t + H] = x[f];
    }
    if constexpr (kDt) {
        // the convective maximum from u, v, w, c^2 of the rows; the viscous one
        // at the largest 1/rho
        if (work) {
#pragma unroll
            for (int t = 0; t < K; t++)
                dm0 = fmax(dm0, dt_conv(sw.ph, w[0][t + H], w[1][t + H], w[2][t + H], w[3][t + H]));
            dm1 = dt_visc(sw.ph, dm1);
        }
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            dm0 = fmax(dm0, __shfl_xor(dm0, o));
            dm1 = fmax(dm1, __shfl_xor(dm1, o));
        }
    }
    if (AX == 0) {
#pragma unroll
        for (int f = 0; f < NF; f++) {
            w[f][0] = __shfl_up(w[f][K], 1, GS);
            w[f][1] = __shfl_up(w[f][K + 1], 1, GS);
            w[f][K + 2] = __shfl_down(w[f][2], 1, GS);
            w[f][K + 3] = __shfl_down(w[f][3], 1, GS);
        }
    } else if constexpr (H != 2) {
        // halo rows through LDS one field at a time: hb[k*P + c][TL], k = the
        // first H then the last H rows of chunk c (neighbours wrap on cyclic lines)
        const int cl = c > 0 ? c - 1 : (cyc ? P - 1 : 0), cr = c + 1 < P ? c + 1 : (cyc ? 0 : P - 1);
#pragma unroll
        for (int f = 0; f < NF; f++) {
            if (f > 0) __syncthreads();
#pragma unroll
            for (int k = 0; k < H; k++) {
                lds[(k * P + c) * TL + lloc] = w[f][H + k];
                lds[((H + k) * P + c) * TL + lloc] = w[f][K + k];
            }
            __syncthreads();
#pragma unroll
            for (int k = 0; k < H; k++) {
                w[f][k] = lds[((H + k) * P + cl) * TL + lloc];
                w[f][K + H + k] = lds[(k * P + cr) * TL + lloc];
            }
        }
    } else {
        // halo rows through LDS: hb[(f*4 + k)*P + c][TL], k = chunk rows 0, 1, K-2, K-1
        const int cl = c > 0 ? c - 1 : 0, cr = c + 1 < P ? c + 1 : P - 1;
#pragma unroll
        for (int f = 0; f < NF; f++) {
            lds[((f * 4 + 0) * P + c) * TL + lloc] = w[f][2];
            lds[((f * 4 + 1) * P + c) * TL + lloc] = w[f][3];
            lds[((f * 4 + 2) * P + c) * TL + lloc] = w[f][K];
            lds[((f * 4 + 3) * P + c) * TL + lloc] = w[f][K + 1];
        }
        __syncthreads();
#pragma unroll
        for (int f = 0; f < NF; f++) {
            w[f][0] = lds[((f * 4 + 2) * P + cl) * TL + lloc];
            w[f][1] = lds[((f * 4 + 3) * P + cl) * TL + lloc];
            w[f][K + 2] = lds[((f * 4 + 0) * P + cr) * TL + lloc];
            w[f][K + 3] = lds[((f * 4 + 1) * P + cr) * TL + lloc];
        }
    }
    double pv[SW::kPv ? NF : 1][K];
    if constexpr (SW::kPv) {
#pragma unroll
        for (int f = 0; f < NF; f++)
#pragma unroll
            for (int t = 0; t < K; t++) pv[f][t] = w[f][t + 2];
    }
    double r[NF][K];
    if constexpr (SW::kFilter) {
        filter_rhs<K, NF, W>(w, r, s, N, cyc, ax.ftab + 6 * N + 8);
    } else {
        window_rhs<K, NF, W>(w, r, K, work && c == 0 && !ax.seam, work && c == P - 1 && !ax.seam);
        if (work && ax.seam && (c == 0 || c == P - 1))
            seam_rows<K, NF>(sw, base, st, N, K, c == 0, c == P - 1, cm.pb, r);
    }

    double carry[NF], G[NF], X[NF];
    forward_local<K, NF, true>(tab, N, s, K, r, carry);
    if (AX == 0) {
        scan_forward<NF>(cv ? tab[3 * N + s + K - 1] : 1.0, carry, cv, c, GS, G);
    } else {
        __syncthreads();                                // halo reads done: reuse LDS
#pragma unroll
        for (int f = 0; f < NF; f++) lds[(f * P + c) * TL + lloc] = carry[f];
        // the waves' time-step maxima go past the carries of both folds
        double* dl = lds + 3 * NF * P * TL;
        if constexpr (kDt) {
            if ((threadIdx.x & 63) == 0) {
                dl[2 * (threadIdx.x >> 6)] = dm0;
                dl[2 * (threadIdx.x >> 6) + 1] = dm1;
            }
        }
        __syncthreads();
        if constexpr (kDt) {
            if (threadIdx.x == 0) {
                double a0 = 0.0, a1 = 0.0;
                for (int v = 0; v < (int)(blockDim.x >> 6); v++) {
                    a0 = fmax(a0, dl[2 * v]);
                    a1 = fmax(a1, dl[2 * v + 1]);
                }
                sw.dtpart[2 * blockIdx.x] = a0;
                sw.dtpart[2 * blockIdx.x + 1] = a1;
            }
        }
        fold_forward<NF, true, K>(lds, tab, N, P, TL, c, lloc, G);
    }
    apply_carry<K, NF, true>(tab, 3 * N, s, K, r, G);
    backward_local<K, NF, true>(tab, N, s, K, r, carry);
    if (AX == 0) {
        scan_backward<NF>(cv ? tab[4 * N + s] : 0.0, carry, cv, c, P, GS, X);
    } else {
        double* xb = lds + NF * P * TL;
#pragma unroll
        for (int f = 0; f < NF; f++) xb[(f * P + c) * TL + lloc] = carry[f];
        __syncthreads();
        fold_backward<NF, true, K>(xb, tab, N, P, TL, c, lloc, X);
    }
    apply_carry<K, NF, true>(tab, 4 * N, s, K, r, X);
    if (AX != 0 && cyc) {
        // rows 0 and N-1 of the modified system through LDS (after the folds' reads)
        double* ends = lds + 2 * NF * P * TL;
        if (c == 0) {
#pragma unroll
            for (int f = 0; f < NF; f++) ends[f * TL + lloc] = r[f][0];
        }
        if (c == P - 1) {
#pragma unroll
            for (int f = 0; f < NF; f++) ends[(NF + f) * TL + lloc] = r[f][K - 1];
        }
        __syncthreads();
        double y0[NF], yN1[NF];
#pragma unroll
        for (int f = 0; f < NF; f++) {
            y0[f] = ends[f * TL + lloc];
            yN1[f] = ends[(NF + f) * TL + lloc];
        }
        cyclic_correction<K, NF>(tab, N, s, y0, yN1, r);
    }

    if (!work) return;
    face_corrections<K, NF, SW, GCBC, true>(sw, base, st, N, s, K, c == 0, c == P - 1, r, cm.sbcc);
    // every row's epilogue loads before the first store
#pragma unroll
    for (int t = 0; t < K; t++) {
        double d[NF];
#pragma unroll
        for (int f = 0; f < NF; f++) d[f] = r[f][t];
        sw.fetch(base + (uoff)(s + t) * st, d);
#pragma unroll
        for (int f = 0; f < NF; f++) r[f][t] = d[f];
    }
#pragma unroll
    for (int t = 0; t < K; t++) {
        double d[NF], p[NF];
#pragma unroll
        for (int f = 0; f < NF; f++) { d[f] = r[f][t]; p[f] = SW::kPv ? pv[SW::kPv ? f : 0][t] : 0.0; }
        sw.store(base + (uoff)(s + t) * st, d, p);
    }
}

// ---------------------------------------------------------------------------
// variant "direct" for x lines (equal chunks): one wave holds the GS-lane
// chunk groups of 64/GS consecutive lines, which are one contiguous run of
// memory.  Lanes load and evaluate that run with fully coalesced accesses (J
// points per lane), the values are transposed field by field through a
// wave-private LDS buffer into the chunk layout (halo rows included), the
// line is solved with wave scans, and the results go back through the buffer
// for coalesced stores.  No workgroup barrier: a workgroup is one wave.
// ---------------------------------------------------------------------------
template <int K, class SW, bool GCBC>
__global__ void __launch_bounds__(64) xt_kernel(AxisDesc ax, SweepCommon cm, SW sw) {
    constexpr int NF = SW::NF;
    constexpr int J = K;                  // points per lane: (64/GS)*P*K <= 64*K
    constexpr int H = SW::kH;
    constexpr int W = K + 2 * H;
    static_assert(H <= K, "halo rows come from the neighbouring chunks only");
    // [64/GS lines][P chunks][K+1], then (kStash > 0) the stashed point
    // values [kStash][J][64]
    extern __shared__ double lds[];

    const int N = ax.N, P = ax.P, GS = ax.GS;
    const int lane = threadIdx.x;
    const int lpw = 64 / GS;
    double* stash = lds + lpw * P * (K + 1);
    const long line0 = (long)blockIdx.x * lpw;
    const int nl = (int)min((long)lpw, ax.nlines - line0);   // lines of this wave
    const int npts = nl * N;
    const uoff st = (uoff)ax.st * 8u;                        // 8: x rows are adjacent
    const uoff run = line_off(0, (uoff)line0, (uoff)cm.lxi, (uoff)cm.plane);

    // coalesced evaluation: lane handles points lane + 64*j of the run
    double xq[NF][J];
#pragma unroll
    for (int j = 0; j < J; j++) {
        const int p = lane + 64 * j;
        double x[NF];
        if (p < npts) {
            sw.point(run + (uoff)p * st, x);
        } else {
#pragma unroll
            for (int f = 0; f < NF; f++) x[f] = 0.0;
        }
#pragma unroll
        for (int f = 0; f < NF; f++) xq[f][j] = x[f];
    }
    if constexpr (SW::kStash > 0) {
        // lane-contiguous: no bank conflicts; each lane reads back only its own slots
#pragma unroll
        for (int f = 0; f < SW::kStash; f++)
#pragma unroll
            for (int j = 0; j < J; j++) stash[(f * J + j) * 64 + lane] = xq[f][j];
    }

    // LDS slot of each of the lane's points (row and line tracked without division)
    int slot[J];
    {
        int l = 0, row = lane;
        while (row >= N) { row -= N; l++; }
#pragma unroll
        for (int j = 0; j < J; j++) {
            slot[j] = lane + 64 * j < npts ? xt_slot<K>(l, P, row / K, row % K) : 0;
            row += 64;
            while (row >= N) { row -= N; l++; }
        }
    }

    const int li = lane / GS, c = lane % GS;
    const bool work = li < nl && c < P;
    const int lr = li < nl ? li : 0;
    const int cc = c < P ? c : P - 1;
    const int s = cc * K;
    const double* tab = SW::kFilter ? ax.ftab : ax.tab;
    const uoff base = run + (uoff)lr * (uoff)N * st;
    const bool cyc = SW::kFilter && ax.seam;
    // halo chunks: the neighbours, wrapped on a cyclic line (other lines clamp;
    // their values there are not used)
    const int hl = cc > 0 ? cc - 1 : (cyc ? P - 1 : 0);
    const int hr = cc + 1 < P ? cc + 1 : (cyc ? 0 : P - 1);

    double r[NF][K];
#pragma unroll
    for (int f = 0; f < NF; f++) {
        __syncthreads();
#pragma unroll
        for (int j = 0; j < J; j++)
            if (lane + 64 * j < npts) lds[slot[j]] = xq[f][j];
        __syncthreads();
        double w[1][W];
#pragma unroll
        for (int u = 0; u < H; u++) {
            w[0][u] = lds[xt_slot<K>(lr, P, hl, K - H + u)];
            w[0][K + H + u] = lds[xt_slot<K>(lr, P, hr, u)];
        }
#pragma unroll
        for (int t = 0; t < K; t++) w[0][t + H] = lds[xt_slot<K>(lr, P, cc, t)];
        double rf[1][K];
        if constexpr (SW::kFilter)
            filter_rhs<K, 1, W>(w, rf, s, N, cyc, ax.ftab + 6 * N + 8);
        else
            window_rhs<K, 1, W>(w, rf, K, work && c == 0 && !ax.seam, work && c == P - 1 && !ax.seam);
#pragma unroll
        for (int t = 0; t < K; t++) r[f][t] = rf[0][t];
    }
    if constexpr (!SW::kFilter)
        if (work && ax.seam && (c == 0 || c == P - 1)) seam_rows<K, NF>(sw, base, st, N, K, c == 0, c == P - 1, cm.pb, r);

    const bool cv = c < P;
    double carry[NF], G[NF], X[NF];
    forward_local<K, NF, true>(tab, N, s, K, r, carry);
    scan_forward<NF>(cv ? tab[3 * N + s + K - 1] : 1.0, carry, cv, c, GS, G);
    apply_carry<K, NF, true>(tab, 3 * N, s, K, r, G);
    backward_local<K, NF, true>(tab, N, s, K, r, carry);
    scan_backward<NF>(cv ? tab[4 * N + s] : 0.0, carry, cv, c, P, GS, X);
    apply_carry<K, NF, true>(tab, 4 * N, s, K, r, X);
    if (cyc) {
        double y0[NF], yN1[NF];
#pragma unroll
        for (int f = 0; f < NF; f++) {
            y0[f] = __shfl(r[f][0], 0, GS);
            yN1[f] = __shfl(r[f][K - 1], P - 1, GS);
        }
        cyclic_correction<K, NF>(tab, N, s, y0, yN1, r);
    }
    if (work) face_corrections<K, NF, SW, GCBC, true>(sw, base, st, N, s, K, c == 0, c == P - 1, r, cm.sbcc);

    // back to the coalesced layout, one field at a time (the point values
    // xq stay live only for a store that needs them)
    double xo[NF][J];
#pragma unroll
    for (int f = 0; f < NF; f++) {
        __syncthreads();
        if (work) {
#pragma unroll
            for (int t = 0; t < K; t++) lds[xt_slot<K>(li, P, c, t)] = r[f][t];
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < J; j++) xo[f][j] = lds[slot[j]];
    }
    double mm[2] = {0.0, 0.0};           // DtFold<SW>::close: the time step's maxima
    // every point's epilogue loads before the first store
#pragma unroll
    for (int j = 0; j < J; j++) {
        const int p = lane + 64 * j;
        if (p < npts) {
            double d[NF];
#pragma unroll
            for (int f = 0; f < NF; f++) d[f] = xo[f][j];
            sw.fetch(run + (uoff)p * st, d);
#pragma unroll
            for (int f = 0; f < NF; f++) xo[f][j] = d[f];
        }
    }
#pragma unroll
    for (int j = 0; j < J; j++) {
        const int p = lane + 64 * j;
        if (p < npts) {
            double d[NF];
#pragma unroll
            for (int f = 0; f < NF; f++) d[f] = xo[f][j];
            if constexpr (SW::kStash > 0) {
                double sv[SW::kStash > 0 ? SW::kStash : 1];
#pragma unroll
                for (int f = 0; f < SW::kStash; f++) sv[f] = stash[(f * J + j) * 64 + lane];
                if constexpr (DtFold<SW>::close) sw.store_pv_dt(run + (uoff)p * st, d, sv, mm);
                else sw.store_pv(run + (uoff)p * st, d, sv);
            } else {
                double pv[NF];
#pragma unroll
                for (int f = 0; f < NF; f++) pv[f] = SW::kPv ? xq[f][j] : 0.0;
                sw.store(run + (uoff)p * st, d, pv);
            }
        }
    }
    if constexpr (DtFold<SW>::close) {
        // the workgroup is one wave: reduce across it, one pair per workgroup
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            mm[0] = fmax(mm[0], __shfl_xor(mm[0], o));
            mm[1] = fmax(mm[1], __shfl_xor(mm[1], o));
        }
        if (lane == 0) {
            sw.dtpart[2 * blockIdx.x] = mm[0];
            sw.dtpart[2 * blockIdx.x + 1] = mm[1];
        }
    }
}

// ---------------------------------------------------------------------------
// variant "ragged": any line length; every chunk thread evaluates its rows and
// halo rows itself (x lines: carries by shuffles; y/z: LDS folds)
// ---------------------------------------------------------------------------
template <int K, int AX, class SW, bool GCBC>
__global__ void __launch_bounds__(AX == 0 ? 256 : kYLines * kMaxP) ragged_kernel(AxisDesc ax, SweepCommon cm, SW sw) {
    constexpr int NF = SW::NF;
    constexpr int W = K + 4;
    extern __shared__ double lds[];

    const int N = ax.N, P = ax.P;
    int c, lloc = 0, GS = 1;
    long line;
    if (AX == 0) {
        GS = ax.GS;
        const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
        const int lpw = 64 / GS;
        c = lane % GS;
        line = (long)blockIdx.x * (4 * lpw) + wave * lpw + lane / GS;
    } else {
        lloc = threadIdx.x % kYLines;
        c = threadIdx.x / kYLines;
        line = group_of(ax.xcd) * kYLines + lloc;
    }
    const bool act = line < ax.nlines;
    const bool cv = c < P;
    const bool work = act && cv;
    const int s = cv ? (int)(((long)c * N) / P) : N;
    const int L = cv ? (int)(((long)(c + 1) * N) / P) - s : 0;
    const uoff st = (uoff)ax.st * 8u;
    const uoff base = line_off(AX, act ? (uoff)line : 0u, (uoff)cm.lxi, (uoff)cm.plane);
    const double* tab = ax.tab;

    double r[NF][K];
    double pv[SW::kPv ? NF : 1][K];
    {
        double w[NF][W];
#pragma unroll
        for (int u = 0; u < W; u++) {
            const int row = s - 2 + u;
            double x[NF];
            if (work && u < L + 4 && row >= 0 && row < N) {
                sw.point(base + (uoff)row * st, x);
            } else {
#pragma unroll
                for (int f = 0; f < NF; f++) x[f] = 0.0;
            }
#pragma unroll
            for (int f = 0; f < NF; f++) w[f][u] = x[f];
        }
        window_rhs<K, NF, W>(w, r, L, work && c == 0 && !ax.seam, work && c == P - 1 && !ax.seam);
        if constexpr (SW::kPv) {
#pragma unroll
            for (int f = 0; f < NF; f++)
#pragma unroll
                for (int t = 0; t < K; t++) pv[f][t] = w[f][t + 2];
        }
    }
    if (work && ax.seam && (c == 0 || c == P - 1)) seam_rows<K, NF>(sw, base, st, N, L, c == 0, c == P - 1, cm.pb, r);

    double carry[NF], G[NF], X[NF];
    forward_local<K, NF, false>(tab, N, s, L, r, carry);
    if (AX == 0) {
        scan_forward<NF>(cv ? tab[3 * N + s + L - 1] : 1.0, carry, cv, c, GS, G);
    } else {
        if (cv) {
#pragma unroll
            for (int f = 0; f < NF; f++) lds[(f * P + c) * kYLines + lloc] = carry[f];
        }
        __syncthreads();
        fold_forward<NF, false, K>(lds, tab, N, P, kYLines, c, lloc, G);
    }
    apply_carry<K, NF, false>(tab, 3 * N, s, L, r, G);
    backward_local<K, NF, false>(tab, N, s, L, r, carry);
    if (AX == 0) {
        scan_backward<NF>(cv ? tab[4 * N + s] : 0.0, carry, cv, c, P, GS, X);
    } else {
        double* xb = lds + NF * P * kYLines;
        if (cv) {
#pragma unroll
            for (int f = 0; f < NF; f++) xb[(f * P + c) * kYLines + lloc] = carry[f];
        }
        __syncthreads();
        fold_backward<NF, false, K>(xb, tab, N, P, kYLines, c, lloc, X);
    }
    apply_carry<K, NF, false>(tab, 4 * N, s, L, r, X);

    if (!work) return;
    face_corrections<K, NF, SW, GCBC, false>(sw, base, st, N, s, L, c == 0, c == P - 1, r, cm.sbcc);
    // every row's epilogue loads before the first store
#pragma unroll
    for (int t = 0; t < K; t++) {
        if (t < L) {
            double d[NF];
#pragma unroll
            for (int f = 0; f < NF; f++) d[f] = r[f][t];
            sw.fetch(base + (uoff)(s + t) * st, d);
#pragma unroll
            for (int f = 0; f < NF; f++) r[f][t] = d[f];
        }
    }
#pragma unroll
    for (int t = 0; t < K; t++) {
        if (t < L) {
            double d[NF], p[NF];
#pragma unroll
            for (int f = 0; f < NF; f++) { d[f] = r[f][t]; p[f] = SW::kPv ? pv[SW::kPv ? f : 0][t] : 0.0; }
            sw.store(base + (uoff)(s + t) * st, d, p);
        }
    }
}

// ---------------------------------------------------------------------------
// variant "tile": the pointwise inputs of TL lines pass through an LDS tile
// one row block (kCB chunks + halo) at a time with coalesced loads -- every
// point is evaluated once -- and chunk threads stream their windows from the
// tile.  The tile memory is then reused for the carries and, for x lines, for
// a transposed coalesced store.
// ---------------------------------------------------------------------------
constexpr int kCB = 8;   // chunks per row block

template <int K>
__host__ __device__ constexpr int tile_rows() { return kCB * K + 4; }

template <int K, int AX, int TL, bool EX, class SW, bool GCBC>
__global__ void __launch_bounds__(TL * kMaxP) tile_kernel(AxisDesc ax, SweepCommon cm, SW sw) {
    constexpr int NF = SW::NF;
    constexpr int TLP = TL + 1;
    constexpr int RB = tile_rows<K>();
    extern __shared__ double lds[];    // tile [NF][RB][TLP]

    const int N = ax.N, P = ax.P;
    const int nthr = TL * P;
    const int tid = threadIdx.x;
    const int lloc = tid % TL, c = tid / TL;
    const long line0 = group_of(ax.xcd) * TL;
    const long line = line0 + lloc;
    const bool act = line < ax.nlines;
    const int s = EX ? c * K : (int)(((long)c * N) / P);
    const int L = EX ? K : (int)(((long)(c + 1) * N) / P) - s;
    const uoff st = (uoff)ax.st * 8u;
    const uoff lxi = (uoff)cm.lxi, plane = (uoff)cm.plane;
    const uoff base = line_off(AX, act ? (uoff)line : 0u, lxi, plane);
    const double* tab = ax.tab;
    const int nb = (P + kCB - 1) / kCB;

    double r[NF][K];
    double pv[SW::kPv ? NF : 1][K];
    for (int b = 0; b < nb; b++) {
        const int c0 = b * kCB, c1 = min(P, c0 + kCB);
        const int rs = max(0, (EX ? c0 * K : (int)(((long)c0 * N) / P)) - 2);
        const int re = min(N, (EX ? c1 * K : (int)(((long)c1 * N) / P)) + 2);
        const int nrows = re - rs;
        if (b > 0) __syncthreads();
        for (int p = tid; p < TL * nrows; p += nthr) {
            int l, row;
            if (AX == 0) { l = p / nrows; row = rs + p % nrows; }
            else         { l = p % TL;    row = rs + p / TL; }
            const long ln = line0 + l;
            double x[NF];
            if (ln < ax.nlines) {
                sw.point(line_off(AX, (uoff)ln, lxi, plane) + (uoff)row * st, x);
            } else {
#pragma unroll
                for (int f = 0; f < NF; f++) x[f] = 0.0;
            }
#pragma unroll
            for (int f = 0; f < NF; f++) lds[(f * RB + row - rs) * TLP + l] = x[f];
        }
        __syncthreads();
        if (c >= c0 && c < c1) {
            auto row_of = [&](int u, double* x) {
                const int row = s - 2 + u;
                const bool in = (EX || u < L + 4) && row >= 0 && row < N;
#pragma unroll
                for (int f = 0; f < NF; f++) x[f] = in ? lds[(f * RB + row - rs) * TLP + lloc] : 0.0;
            };
            stream_rhs<K, NF>(row_of, r, L, c == 0 && !ax.seam, c == P - 1 && !ax.seam);
            if constexpr (SW::kPv) {
#pragma unroll
                for (int t = 0; t < K; t++) {
                    const int row = s + t;
#pragma unroll
                    for (int f = 0; f < NF; f++)
                        pv[f][t] = (EX || t < L) ? lds[(f * RB + row - rs) * TLP + lloc] : 0.0;
                }
            }
        }
    }
    if (act && ax.seam && (c == 0 || c == P - 1)) seam_rows<K, NF>(sw, base, st, N, L, c == 0, c == P - 1, cm.pb, r);

    double carry[NF], G[NF], X[NF];
    forward_local<K, NF, EX>(tab, N, s, L, r, carry);
    __syncthreads();                                   // tile reads done: reuse for carries
#pragma unroll
    for (int f = 0; f < NF; f++) lds[(f * P + c) * TL + lloc] = carry[f];
    __syncthreads();
    fold_forward<NF, EX, K>(lds, tab, N, P, TL, c, lloc, G);
    apply_carry<K, NF, EX>(tab, 3 * N, s, L, r, G);
    backward_local<K, NF, EX>(tab, N, s, L, r, carry);
    double* xb = lds + NF * P * TL;
#pragma unroll
    for (int f = 0; f < NF; f++) xb[(f * P + c) * TL + lloc] = carry[f];
    __syncthreads();
    fold_backward<NF, EX, K>(xb, tab, N, P, TL, c, lloc, X);
    apply_carry<K, NF, EX>(tab, 4 * N, s, L, r, X);

    if (act) face_corrections<K, NF, SW, GCBC, EX>(sw, base, st, N, s, L, c == 0, c == P - 1, r, cm.sbcc);

    if (AX != 0 || SW::kPv) {             // (stores needing the point values stay per thread)
        if (!act) return;
#pragma unroll
        for (int t = 0; t < K; t++) {
            if (EX || t < L) {
                double d[NF];
#pragma unroll
                for (int f = 0; f < NF; f++) d[f] = r[f][t];
                sw.fetch(base + (uoff)(s + t) * st, d);
#pragma unroll
                for (int f = 0; f < NF; f++) r[f][t] = d[f];
            }
        }
#pragma unroll
        for (int t = 0; t < K; t++) {
            if (EX || t < L) {
                double d[NF], p[NF];
#pragma unroll
                for (int f = 0; f < NF; f++) { d[f] = r[f][t]; p[f] = SW::kPv ? pv[SW::kPv ? f : 0][t] : 0.0; }
                sw.store(base + (uoff)(s + t) * st, d, p);
            }
        }
        return;
    }
    // x lines: results back through the tile, one row block at a time, stored along the line
    for (int b = 0; b < nb; b++) {
        const int c0 = b * kCB, c1 = min(P, c0 + kCB);
        const int rs = EX ? c0 * K : (int)(((long)c0 * N) / P);
        const int re = EX ? c1 * K : (int)(((long)c1 * N) / P);
        const int nrows = re - rs;
        __syncthreads();
        if (c >= c0 && c < c1) {
#pragma unroll
            for (int t = 0; t < K; t++)
                if (EX || t < L)
#pragma unroll
                    for (int f = 0; f < NF; f++) lds[(f * RB + s + t - rs) * TLP + lloc] = r[f][t];
        }
        __syncthreads();
        for (int p = tid; p < TL * nrows; p += nthr) {
            const int l = p / nrows, row = rs + p % nrows;
            const long ln = line0 + l;
            if (ln >= ax.nlines) continue;
            double d[NF], pz[NF];
#pragma unroll
            for (int f = 0; f < NF; f++) { d[f] = lds[(f * RB + row - rs) * TLP + l]; pz[f] = 0.0; }
            const uoff bo = line_off(AX, (uoff)ln, lxi, plane) + (uoff)row * st;
            sw.fetch(bo, d);
            sw.store(bo, d, pz);
        }
    }
}

// ---------------------------------------------------------------------------
// launch
// ---------------------------------------------------------------------------
template <int K, int AX, class SW, bool GCBC>
static int launch_sweep(const AxisDesc& ax, const SweepCommon& cm, const SW& sw, hipStream_t s) {
    constexpr int NF = SW::NF;
    const bool exact = ax.N == ax.P * K;
    if constexpr (K == 8) {
        if (ax.variant == 1) {
            const size_t tile = (size_t)NF * tile_rows<K>() * (ax.tl + 1);
            const size_t carries = (size_t)2 * NF * ax.P * ax.tl;
            const size_t lds = sizeof(double) * (tile > carries ? tile : carries);
            dim3 block(ax.tl * ax.P), grid((unsigned)((ax.nlines + ax.tl - 1) / ax.tl));
            if (exact) {
                if (ax.tl == 8) hipLaunchKernelGGL((tile_kernel<K, AX, 8, true, SW, GCBC>), grid, block, lds, s, ax, cm, sw);
                else hipLaunchKernelGGL((tile_kernel<K, AX, 16, true, SW, GCBC>), grid, block, lds, s, ax, cm, sw);
            } else {
                if (ax.tl == 8) hipLaunchKernelGGL((tile_kernel<K, AX, 8, false, SW, GCBC>), grid, block, lds, s, ax, cm, sw);
                else hipLaunchKernelGGL((tile_kernel<K, AX, 16, false, SW, GCBC>), grid, block, lds, s, ax, cm, sw);
            }
            CS_HIP(hipGetLastError());
            return CS_OK;
        }
    }
    if constexpr (K == 8 && AX == 0) {
        if (exact) {
            const int lpw = 64 / ax.GS;
            const size_t lds = sizeof(double) * (lpw * ax.P * (K + 1) + SW::kStash * K * 64);
            dim3 block(64), grid((unsigned)((ax.nlines + lpw - 1) / lpw));
            hipLaunchKernelGGL((xt_kernel<K, SW, GCBC>), grid, block, lds, s, ax, cm, sw);
            CS_HIP(hipGetLastError());
            return CS_OK;
        }
    }
    if (exact) {
        dim3 block, grid;
        size_t lds = 0;
        const int TL = ax.tl == 8 ? 8 : 16;
        if (AX == 0) {
            const int lpb = 4 * (64 / ax.GS);
            block = dim3(256);
            grid = dim3((unsigned)((ax.nlines + lpb - 1) / lpb));
        } else {
            block = dim3(TL * ax.P);
            grid = dim3((unsigned)((ax.nlines + TL - 1) / TL));
            lds = sizeof(double) * 4 * NF * ax.P * TL;
            if (lds > 65536) {
                const void* k = TL == 8 ? (const void*)direct_kernel<K, AX, 8, SW, GCBC>
                                        : (const void*)direct_kernel<K, AX, 16, SW, GCBC>;
                CS_HIP(hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
            }
        }
        if (AX != 0 && TL == 8)
            hipLaunchKernelGGL((direct_kernel<K, AX, 8, SW, GCBC>), grid, block, lds, s, ax, cm, sw);
        else
            hipLaunchKernelGGL((direct_kernel<K, AX, 16, SW, GCBC>), grid, block, lds, s, ax, cm, sw);
        CS_HIP(hipGetLastError());
        return CS_OK;
    }
    dim3 block, grid;
    size_t lds = 0;
    if (AX == 0) {
        const int lpb = 4 * (64 / ax.GS);
        block = dim3(256);
        grid = dim3((unsigned)((ax.nlines + lpb - 1) / lpb));
    } else {
        block = dim3(kYLines * ax.P);
        grid = dim3((unsigned)((ax.nlines + kYLines - 1) / kYLines));
        lds = sizeof(double) * 2 * NF * ax.P * kYLines;
    }
    hipLaunchKernelGGL((ragged_kernel<K, AX, SW, GCBC>), grid, block, lds, s, ax, cm, sw);
    CS_HIP(hipGetLastError());
    return CS_OK;
}

// the filter sweep on equal chunks of K points (8 along x; 8 or 16 along y/z,
// the axis's filter descriptor); SW = SF<AX> (five variables per workgroup)
// or SF1<AX> (one variable per workgroup, grid.y = 5)
template <int K, int AX, class SW>
static int launch_filter(const AxisDesc& ax, const SweepCommon& cm, const SW& sw, hipStream_t s) {
    constexpr int NF = SW::NF, H = SW::kH;
    static_assert(AX != 0 || K == 8, "x filter sweeps run 8-point chunks");
    const unsigned gy = NF == 1 ? 5u : 1u;
    if (ax.K != K || ax.N != ax.P * K || !ax.ftab) return fail(CS_ERR_STATE, "fused filter needs equal chunks");
    if constexpr (AX == 0) {
        const int lpw = 64 / ax.GS;
        const size_t lds = sizeof(double) * lpw * ax.P * (K + 1);
        dim3 block(64), grid((unsigned)((ax.nlines + lpw - 1) / lpw), gy);
        hipLaunchKernelGGL((xt_kernel<K, SW, false>), grid, block, lds, s, ax, cm, sw);
    } else {
        const int TL = ax.ftl == 8 ? 8 : 16;
        const size_t halo = (size_t)2 * H * ax.P * TL;
        const size_t carries = (size_t)2 * NF * ax.P * TL + 2 * NF * TL;
        const size_t lds = sizeof(double) * (halo > carries ? halo : carries);
        dim3 block(TL * ax.P), grid((unsigned)((ax.nlines + TL - 1) / TL), gy);
        const void* k = TL == 8 ? (const void*)direct_kernel<K, AX, 8, SW, false>
                                : (const void*)direct_kernel<K, AX, 16, SW, false>;
        if (lds > 65536) CS_HIP(hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        if (TL == 8) hipLaunchKernelGGL((direct_kernel<K, AX, 8, SW, false>), grid, block, lds, s, ax, cm, sw);
        else hipLaunchKernelGGL((direct_kernel<K, AX, 16, SW, false>), grid, block, lds, s, ax, cm, sw);
    }
    CS_HIP(hipGetLastError());
    return CS_OK;
}

template <int AX, class SW>
static int dispatch(const AxisDesc& ax, const SweepCommon& cm, const SW& sw, hipStream_t s) {
    if constexpr (SW::kGcbc) {
        if (ax.K == 8) {
            if (ax.gcbc) return launch_sweep<8, AX, SW, true>(ax, cm, sw, s);
            return launch_sweep<8, AX, SW, false>(ax, cm, sw, s);
        }
        if (ax.gcbc) return launch_sweep<16, AX, SW, true>(ax, cm, sw, s);
        return launch_sweep<16, AX, SW, false>(ax, cm, sw, s);
    } else {
        if (ax.K == 8) return launch_sweep<8, AX, SW, false>(ax, cm, sw, s);
        return launch_sweep<16, AX, SW, false>(ax, cm, sw, s);
    }
}

// ---------------------------------------------------------------------------
// host orchestration
// ---------------------------------------------------------------------------
// the solver's FP64 buffers of the viscous intermediates hold FP32 fields of
// the same point count when the intermediates are stored in FP32
template <class TS>
inline TS* store_of(double* p) { return reinterpret_cast<TS*>(p); }

template <template <int, bool> class F, int AX, bool CLOSE>
static F<AX, CLOSE> make_s1(const cs_solver* sv, const double* q) {
    typedef typename F<AX, CLOSE>::Store TS;
    F<AX, CLOSE> a{};
    const long lmx = sv->g.lmx;
    fields(q, lmx, a.q);
    const int* axes = kS1Axes[sv->order];
    if constexpr (!CLOSE) {
        fields(store_of<TS>(AX == axes[0] ? sv->ga : sv->gb), lmx, a.gout);
    } else {
        for (int a2 = 0; a2 < 3; a2++) {
            const TS* g = a2 == axes[0] ? store_of<TS>(sv->ga) : (a2 == axes[1] ? store_of<TS>(sv->gb) : nullptr);
            fields(g, lmx, a.gin[a2]);
        }
        fields(store_of<TS>(sv->tau), lmx, a.tau);
        fields(store_of<TS>(sv->hf), lmx, a.hf);
    }
    a.ph = phys_of(sv);
    if constexpr (F<AX, CLOSE>::kCurv) a.mt = sv->met;
    return a;
}

template <template <int, int, bool> class F, int AX, int MODE, bool V>
static F<AX, MODE, V> make_s2(const cs_solver* sv, const double* q, const double* din, double* dout,
                               const double* q0, int div) {
    typedef typename F<AX, MODE, V>::Store TS;
    F<AX, MODE, V> a{};
    const long lmx = sv->g.lmx;
    fields(q, lmx, a.q);
    if (V) {
        const int t0 = AX == 0 ? 0 : (AX == 1 ? 3 : 5);
        const int t1 = AX == 0 ? 3 : (AX == 1 ? 1 : 4);
        const int t2 = AX == 0 ? 5 : (AX == 1 ? 4 : 2);
        const TS* tau = store_of<TS>(sv->tau);
        a.tr[0] = tau + (long)t0 * lmx;
        a.tr[1] = tau + (long)t1 * lmx;
        a.tr[2] = tau + (long)t2 * lmx;
        a.hx = store_of<TS>(sv->hf) + (long)AX * lmx;
        if constexpr (F<AX, MODE, V>::kCurv) {
            fields(tau, lmx, a.tau6);
            fields((const TS*)store_of<TS>(sv->hf), lmx, a.h3);
        }
    }
    if constexpr (F<AX, MODE, V>::kCurv) a.mt = sv->met;
    fields(din, lmx, a.din);
    fields(dout, lmx, a.dout);
    fields(q0, lmx, a.q0);
    a.dtp = sv->dt;
    a.div = div;
    a.ph = phys_of(sv);
    a.nface = sv->g.face_size(AX);
    a.rec = sv->rec ? sv->rec + face_offset(sv->g, AX) : nullptr;
    a.lxi = sv->g.lxi;
    a.plane = (long)sv->g.lxi * sv->g.let;
    return a;
}

// sweep `pos` of one RHS (0..2 = S1 sweeps, 3..5 = S2 sweeps, in the
// solver's order); the last S2 sweep writes the RHS to `out`, or, when q0 is
// given, the RK stage q0 - dt/div*J*RHS
template <bool V, template <int, int, bool> class F = S2>
static int s2_sweep(cs_solver* sv, int k, const SweepCommon& cm, const double* q, double* out, const double* q0,
                    int div, hipStream_t s) {
    const int ax = kS2Axes[sv->order][k];
    if (k == 0) {
        if (ax == 0) return dispatch<0>(sv->rax[0], cm, make_s2<F, 0, 0, V>(sv, q, nullptr, sv->dpart, nullptr, 0), s);
        return dispatch<2>(sv->rax[2], cm, make_s2<F, 2, 0, V>(sv, q, nullptr, sv->dpart, nullptr, 0), s);
    }
    if (k == 1) return dispatch<1>(sv->rax[1], cm, make_s2<F, 1, 1, V>(sv, q, sv->dpart, sv->dpart, nullptr, 0), s);
    if (ax == 2) {
        if (q0) return dispatch<2>(sv->rax[2], cm, make_s2<F, 2, 3, V>(sv, q, sv->dpart, out, q0, div), s);
        return dispatch<2>(sv->rax[2], cm, make_s2<F, 2, 2, V>(sv, q, sv->dpart, out, nullptr, 0), s);
    }
    if (q0) return dispatch<0>(sv->rax[0], cm, make_s2<F, 0, 3, V>(sv, q, sv->dpart, out, q0, div), s);
    return dispatch<0>(sv->rax[0], cm, make_s2<F, 0, 2, V>(sv, q, sv->dpart, out, nullptr, 0), s);
}

// the gradient functor that also reduces the time step, per storage type
template <template <int, bool> class F>
struct DtOf {
    static constexpr bool kOk = false;
    template <int A, bool C>
    using type = F<A, C>;
};
template <>
struct DtOf<S1> {
    static constexpr bool kOk = true;
    template <int A, bool C>
    using type = S1D<A, C>;
};
template <>
struct DtOf<S1F> {
    static constexpr bool kOk = true;
    template <int A, bool C>
    using type = S1FD<A, C>;
};
template <>
struct DtOf<S1C> {
    static constexpr bool kOk = true;
    template <int A, bool C>
    using type = S1CD<A, C>;
};
template <>
struct DtOf<S1CF> {
    static constexpr bool kOk = true;
    template <int A, bool C>
    using type = S1CFD<A, C>;
};

template <template <int, bool> class F = S1>
static int s1_sweep(cs_solver* sv, int k, const SweepCommon& cm, const double* q, hipStream_t s) {
    if (sv->order == 0) {
        if (k == 0) return dispatch<0>(sv->rax[0], cm, make_s1<F, 0, false>(sv, q), s);
        if (k == 1) return dispatch<1>(sv->rax[1], cm, make_s1<F, 1, false>(sv, q), s);
        return dispatch<2>(sv->rax[2], cm, make_s1<F, 2, true>(sv, q), s);
    }
    // the sweep of an RK4 step's first stage that reduces the time step
    // (dt_fold = its position + 1): the first with constant metrics, the
    // closing x sweep with per-point metrics
    if (k == 0) {
        if constexpr (DtOf<F>::kOk && !F<1, false>::kCurv) {
            if (sv->dt_fold == 1) {
                auto f = make_s1<DtOf<F>::template type, 1, false>(sv, q);
                f.dtpart = sv->dtpart;
                return dispatch<1>(sv->rax[1], cm, f, s);
            }
        }
        return dispatch<1>(sv->rax[1], cm, make_s1<F, 1, false>(sv, q), s);
    }
    if (k == 1) return dispatch<2>(sv->rax[2], cm, make_s1<F, 2, false>(sv, q), s);
    if constexpr (DtOf<F>::kOk && F<0, true>::kCurv) {
        if (sv->dt_fold == 3) {
            auto f = make_s1<DtOf<F>::template type, 0, true>(sv, q);
            f.dtpart = sv->dtpart;
            return dispatch<0>(sv->rax[0], cm, f, s);
        }
    }
    return dispatch<0>(sv->rax[0], cm, make_s1<F, 0, true>(sv, q), s);
}

}  // namespace fused
}  // namespace cs
