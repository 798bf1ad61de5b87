// tape.hip -- the adjoint of the fused Euler RK4 step as transposed line
// sweeps, and the forward step that records what the adjoint reads (no
// reference counterpart).
//
// For the inviscid right-hand side R(q) = sum_a [ D_a F^a(q) + C^a(q) ]
// (operators described in vjp.hip),
//   R'(q)^T t = sum_a [ (dF^a/dq)^T D_a^T (t + S_a^T mu_a) + G_a ],
// with D_a^T = B_a^T A_a^-T.  One sweep per axis evaluates, line by line:
// the seed t plus the face seeds mu_a at the two face rows; the solve with
// the LU factors of A^T (the chunked recurrences of the forward sweeps on the
// transposed table); the transposed explicit stencil B^T (inside the line the
// derivative stencil with the opposite sign, plus the closure or seam rows at
// the line ends); and the transposed flux Jacobian at every point,
// accumulated over the three sweeps.  The last sweep of a stage also carries
// the RK4 recurrence.  The filter's transpose is a sweep of the same form
// (A^T solve, then B^T) on closure lines; on cyclic lines the filter is a
// symmetric circulant and is its own transpose.  mu_a and G_a, the adjoint
// of the characteristic BC at the face points, come from a small face kernel
// that runs before the sweeps of a stage.
//
// Tape of one step: the stage states q0..q3 and, for every stage, the
// boundary derivatives of the flux divergence along each axis (what the
// characteristic BC linearises about), [4 stages][3 axes][2 face][5][face].
#include "sweep_parts.hpp"
#include "vjp_parts.hpp"

namespace cs {
namespace fused {

namespace {

// g = (dF^AX/dq)^T nu for the inviscid flux row of flux_row(): with
// ss = m (u_AX + uf) and p = (gamma-1)(E - |m|^2/(2 rho)),
//   F_0 = rho ss, F_1+i = (rho u_i) ss + [i = AX] m p, F_4 = (E + p) ss - p uf m,
// the products collect into cs (the coefficient of d ss) and cp (of d p).
template <int AX>
__device__ __forceinline__ void flux_jt(const double* q, const double* nu, double m, double uf, double* g) {
    const Prim p = primitive(q);
    const double vel = AX == 0 ? p.u : (AX == 1 ? p.v : p.w);
    const double ss = m * (vel + uf);
    const double cs = nu[1] * q[1] + nu[2] * q[2] + nu[3] * q[3] + nu[4] * (q[4] + p.p);
    const double cp = m * nu[1 + AX] + nu[4] * (m * vel);
    const double u[3] = {p.u, p.v, p.w};
    g[0] = nu[0] * (m * uf) - cs * (m * vel * p.rhoi) + cp * (kGamm1 * 0.5 * (p.u * p.u + p.v * p.v + p.w * p.w));
#pragma unroll
    for (int i = 0; i < 3; i++) g[1 + i] = nu[1 + i] * ss - cp * (kGamm1 * u[i]);
    g[1 + AX] += nu[0] * m + cs * (m * p.rhoi);
    g[4] = nu[4] * ss + cp * kGamm1;
}

// the same for a general metric row m (curvilinear grids): ss = m.(u + uf),
// F_1+i = (rho u_i) ss + m_i p, F_4 = (E + p) ss - p (uf.m); with mu = m.u,
//   g_0 = nu_0 (m.uf) - cs mu / rho + cp (gamma-1)|u|^2/2,
//   g_1+i = nu_1+i ss + nu_0 m_i + cs m_i / rho - cp (gamma-1) u_i,
//   g_4 = nu_4 ss + cp (gamma-1),
// cs = nu_1..3 . (rho u) + nu_4 (E + p), cp = nu_1..3 . m + nu_4 mu
__device__ __forceinline__ void flux_jt_m(const double* q, const double* nu, const double* m, cs_point umf,
                                          double* g) {
    const Prim p = primitive(q);
    const double u[3] = {p.u, p.v, p.w};
    const double mu = m[0] * p.u + m[1] * p.v + m[2] * p.w;
    const double mf = m[0] * umf.x + m[1] * umf.y + m[2] * umf.z;
    const double ss = m[0] * (p.u + umf.x) + m[1] * (p.v + umf.y) + m[2] * (p.w + umf.z);
    const double cs = nu[1] * q[1] + nu[2] * q[2] + nu[3] * q[3] + nu[4] * (q[4] + p.p);
    const double cp = m[0] * nu[1] + m[1] * nu[2] + m[2] * nu[3] + nu[4] * mu;
    g[0] = nu[0] * mf - cs * (mu * p.rhoi) + cp * (kGamm1 * 0.5 * (p.u * p.u + p.v * p.v + p.w * p.w));
#pragma unroll
    for (int i = 0; i < 3; i++) g[1 + i] = nu[1 + i] * ss + nu[0] * m[i] + cs * (m[i] * p.rhoi) - cp * (kGamm1 * u[i]);
    g[4] = nu[4] * ss + cp * kGamm1;
}

// AT: the transposed flux-divergence sweep along AX of one stage.
// MODE 0: gout = J^T nu (+ G); 1: gout = gin + J^T nu (+ G); 2 (last sweep of
// stages 4..2): tn = -ck (gin + J^T nu + G), tout = tn, lam += tn; 3 (last
// sweep of stage 1): lam += -ck (gin + J^T nu + G)
// CURV: the point's metric row (curvilinear solvers; their per-point
// Jacobian scales the seed before the stage, see backward_fused)
template <int AX, int MODE, bool CURV = false>
struct AT {
    static constexpr int NF = 5;
    static constexpr int kWpe = 3;        // waves per SIMD of the y/z kernels (168 VGPRs)
    const double* q[5];       // the stage state the RHS was evaluated at
    const double* t[5];       // seed
    const double* gin[5];
    double* gout[5];
    double* lam[5];
    double* tout[5];
    const double* G;          // [2 face][5][nface] face-point terms, or null
    long nface;
    double ck;
    double m, uf;
    cs_point umf;
    Metric mt;
    __device__ __forceinline__ void seed(uoff bo, double* v) const {
#pragma unroll
        for (int f = 0; f < 5; f++) v[f] = ldg(t[f], bo);
    }
    // J^T nu at one point folded into d with the stage's combination (loads
    // only); store() then writes.  The kernels fetch all their points before
    // their first store: the outputs may alias q, gin or lam for all the
    // compiler knows, so a load placed after a store waits for it.
    __device__ __forceinline__ void fetch(uoff bo, double* d, int face, long line) const {
        double qq[5], g[5];
        const double c = ck;
#pragma unroll
        for (int v = 0; v < 5; v++) qq[v] = ldg(q[v], bo);
        if constexpr (CURV) {
            const double mr[3] = {ldg(mt.m[AX][0], bo), ldg(mt.m[AX][1], bo), ldg(mt.m[AX][2], bo)};
            flux_jt_m(qq, d, mr, umf, g);
        } else {
            flux_jt<AX>(qq, d, m, uf, g);
        }
        if (G && face >= 0) {
#pragma unroll
            for (int v = 0; v < 5; v++) g[v] += G[(face * 5 + v) * nface + line];
        }
#pragma unroll
        for (int v = 0; v < 5; v++) {
            if (MODE == 0) d[v] = g[v];
            else if (MODE == 1) d[v] = ldg(gin[v], bo) + g[v];
            else if (MODE == 2) d[v] = -c * (ldg(gin[v], bo) + g[v]);
            else d[v] = ldg(lam[v], bo) - c * (ldg(gin[v], bo) + g[v]);
        }
    }
    __device__ __forceinline__ void store(uoff bo, const double* d) const {
        if (MODE == 2) {
            double l[5];
#pragma unroll
            for (int v = 0; v < 5; v++) l[v] = ldg(lam[v], bo);
#pragma unroll
            for (int v = 0; v < 5; v++) {
                stg(tout[v], bo, d[v]);
                stg(lam[v], bo, l[v] + d[v]);
            }
        } else {
#pragma unroll
            for (int v = 0; v < 5; v++) stg(MODE == 3 ? lam[v] : gout[v], bo, d[v]);
        }
    }
};

// FT: the transposed filter in place on five fields lmx apart; NF = 5 (all
// per workgroup) or 1 (blockIdx.y picks the field)
template <int NFV>
struct FT {
    static constexpr int NF = NFV;
    static constexpr int kWpe = NFV == 1 ? 4 : 2;
    double* q;
    long lmx;
    __device__ __forceinline__ double* field(int f) const {
        return q + (long)(NFV == 1 ? (int)blockIdx.y : f) * lmx;
    }
    __device__ __forceinline__ void seed(uoff bo, double* v) const {
#pragma unroll
        for (int f = 0; f < NFV; f++) v[f] = ldg(field(f), bo);
    }
    __device__ __forceinline__ void fetch(uoff, double*, int, long) const {}
    __device__ __forceinline__ void store(uoff bo, const double* d) const {
#pragma unroll
        for (int f = 0; f < NFV; f++) stg(field(f), bo, d[f]);
    }
};

// B^T of the derivative on the K rows of chunk c from a window w holding rows
// s-2+u (wrapped on seam lines).  Rows i in [2, N-3] are interior rows, so
// column j receives B[i][j] = +bb, +aa, -aa, -bb from i = j-2, j-1, j+1, j+2;
// the closure rows (forward_system in adjoint.hip; build_system,
// src/cuda/kernels/numerics_device_pc.hpp:176-293) or the seam rows
// (pbci on the own side, pbco on the far side of the wrap) add at the ends.
template <int K>
__device__ __forceinline__ void deriv_bt(const double (&w)[K + 4], double (&o)[K], int s, int N, int c, int P,
                                         bool seam, const double* pb) {
    if (c > 0 && c < P - 1) {
#pragma unroll
        for (int t = 0; t < K; t++) o[t] = kAa * (w[t + 1] - w[t + 3]) + kBb * (w[t] - w[t + 4]);
        return;
    }
#pragma unroll
    for (int t = 0; t < K; t++) {
        const int j = s + t;
        double acc = 0.0;
        if (j - 2 >= 2 && j - 2 <= N - 3) acc += kBb * w[t];
        if (j - 1 >= 2 && j - 1 <= N - 3) acc += kAa * w[t + 1];
        if (j + 1 >= 2 && j + 1 <= N - 3) acc -= kAa * w[t + 3];
        if (j + 2 >= 2 && j + 2 <= N - 3) acc -= kBb * w[t + 4];
        o[t] = acc;
    }
    if (!seam) {
        if (c == 0) {
            const double y0 = w[2], y1 = w[3];
            o[0] += kAb00 * y0 - kAb10 * y1;
            o[1] += kAb01 * y0;
            o[2] += kAb02 * y0 + kAb10 * y1;
        } else {
            const double yN1 = w[K + 1], yN2 = w[K];
            o[K - 1] += kAb10 * yN2 - kAb00 * yN1;
            o[K - 2] += -kAb01 * yN1;
            o[K - 3] += -kAb02 * yN1 - kAb10 * yN2;
        }
    } else {
        static_assert(K == kLmd, "seam columns span one 8-point chunk");
        const double* pbco = pb;              // pbco[16] | pbci[16]
        const double* pbci = pb + 2 * kLmd;
        if (c == 0) {
            const double y0 = w[2], y1 = w[3], yN1 = w[1], yN2 = w[0];
#pragma unroll
            for (int t = 0; t < K; t++)
                o[t] += pbci[t] * y0 + pbci[kLmd + t] * y1 - pbco[t] * yN1 - pbco[kLmd + t] * yN2;
        } else {
            const double y0 = w[K + 2], y1 = w[K + 3], yN1 = w[K + 1], yN2 = w[K];
#pragma unroll
            for (int t = 0; t < K; t++) {
                const int i = K - 1 - t;
                o[t] += pbco[i] * y0 + pbco[kLmd + i] * y1 - pbci[i] * yN1 - pbci[kLmd + i] * yN2;
            }
        }
    }
}

// B^T of the closure-line filter from a window w holding rows s-5+u: column j
// receives B[i][j] from rows i = j-5..j+5, where row i has order 2*min(d,5),
// d = min(i, N-1-i) its distance from the nearer end, and is the identity at
// d = 0 (an + 6M = coefficient set M, as in filter_rhs of fused.hip)
template <int K>
__device__ __forceinline__ void filter_bt(const double (&w)[K + 10], double (&o)[K], int s, int N, int c, int P,
                                          const double* an) {
    constexpr int H = 5;
    if (c >= 2 && c <= P - 3) {
#pragma unroll
        for (int t = 0; t < K; t++) {
            double v = an[30] * w[t + H];
#pragma unroll
            for (int m = 1; m <= 5; m++) v += 0.5 * an[30 + m] * (w[t + H + m] + w[t + H - m]);
            o[t] = v;
        }
        return;
    }
#pragma unroll
    for (int t = 0; t < K; t++) {
        const int j = s + t;
        double acc = 0.0;
#pragma unroll
        for (int m = -H; m <= H; m++) {
            const int i = j + m;
            if (i < 0 || i >= N) continue;
            const int d = min(i, N - 1 - i);
            const int am = m < 0 ? -m : m;
            double cf = 0.0;
            if (d == 0) {
                cf = m == 0 ? 1.0 : 0.0;
            } else {
                const int M = min(d, H);
                if (am <= M) cf = am == 0 ? an[6 * M] : 0.5 * an[6 * M + am];
            }
            acc += cf * w[t + H + m];
        }
        o[t] = acc;
    }
}

template <int K, int ST, int W>
__device__ __forceinline__ void stencil_t(const double (&w)[W], double (&o)[K], int s, int N, int c, int P,
                                          bool seam, const double* pb, const double* an) {
    if constexpr (ST == 0)
        deriv_bt<K>(w, o, s, N, c, P, seam, pb);
    else
        filter_bt<K>(w, o, s, N, c, P, an);
}

// ---------------------------------------------------------------------------
// transposed sweep, x lines: one wave holds the chunk groups of 64/GS lines
// (a contiguous run); coalesced loads and stores, LDS transposes field by
// field, carries by wave scans.  ST: 0 = derivative, 1 = closure filter.
// ---------------------------------------------------------------------------
template <int K, class SW, int ST>
__global__ void __launch_bounds__(64) xtT_kernel(AxisDesc ax, SweepCommon cm, SW sw, const double* tabT,
                                                 const double* mu, const double* an) {
    constexpr int NF = SW::NF;
    constexpr int J = K;
    constexpr int H = ST == 0 ? 2 : 5;
    constexpr int W = K + 2 * H;
    static_assert(H <= K, "halo rows come from the neighbouring chunks only");
    extern __shared__ double lds[];       // [64/GS lines][P chunks][K+1]

    const int N = ax.N, P = ax.P, GS = ax.GS;
    const int lane = threadIdx.x;
    const int lpw = 64 / GS;
    const long line0 = (long)blockIdx.x * lpw;
    const int nl = (int)min((long)lpw, ax.nlines - line0);
    const int npts = nl * N;
    const uoff run = line_off(0, (uoff)line0, (uoff)cm.lxi, (uoff)cm.plane);

    double xq[NF][J];
#pragma unroll
    for (int j = 0; j < J; j++) {
        const int p = lane + 64 * j;
        double x[NF];
        if (p < npts) {
            sw.seed(run + (uoff)p * 8u, x);
        } else {
#pragma unroll
            for (int f = 0; f < NF; f++) x[f] = 0.0;
        }
#pragma unroll
        for (int f = 0; f < NF; f++) xq[f][j] = x[f];
    }
    int slot[J], lrw[J];                 // LDS slot; line << 10 | row
    {
        int l = 0, row = lane;
        while (row >= N) { row -= N; l++; }
#pragma unroll
        for (int j = 0; j < J; j++) {
            slot[j] = lane + 64 * j < npts ? xt_slot<K>(l, P, row / K, row % K) : 0;
            lrw[j] = (l << 10) | row;
            row += 64;
            while (row >= N) { row -= N; l++; }
        }
    }

    const int li = lane / GS, c = lane % GS;
    const bool work = li < nl && c < P;
    const int lr = li < nl ? li : 0;
    const int cc = c < P ? c : P - 1;
    const int s = cc * K;
    const long line = line0 + lr;

    double r[NF][K];
#pragma unroll
    for (int f = 0; f < NF; f++) {
        __syncthreads();
#pragma unroll
        for (int j = 0; j < J; j++)
            if (lane + 64 * j < npts) lds[slot[j]] = xq[f][j];
        __syncthreads();
#pragma unroll
        for (int t = 0; t < K; t++) r[f][t] = lds[xt_slot<K>(lr, P, cc, t)];
    }
    if (mu && work) {
        const long nf = ax.nlines;
#pragma unroll
        for (int f = 0; f < NF; f++) {
            if (c == 0) r[f][0] += mu[f * nf + line];
            if (c == P - 1) r[f][K - 1] += mu[(5 + f) * nf + line];
        }
    }

    const bool cv = c < P;
    double carry[NF], G[NF], X[NF];
    forward_local<K, NF, true>(tabT, N, s, K, r, carry);
    scan_forward<NF>(cv ? tabT[3 * N + s + K - 1] : 1.0, carry, cv, c, GS, G);
    apply_carry<K, NF, true>(tabT, 3 * N, s, K, r, G);
    backward_local<K, NF, true>(tabT, N, s, K, r, carry);
    scan_backward<NF>(cv ? tabT[4 * N + s] : 0.0, carry, cv, c, P, GS, X);
    apply_carry<K, NF, true>(tabT, 4 * N, s, K, r, X);

    const bool wrap = ST == 0 && ax.seam;
    const int hl = cc > 0 ? cc - 1 : (wrap ? P - 1 : 0);
    const int hr = cc + 1 < P ? cc + 1 : (wrap ? 0 : P - 1);
#pragma unroll
    for (int f = 0; f < NF; f++) {
        __syncthreads();
        if (work) {
#pragma unroll
            for (int t = 0; t < K; t++) lds[xt_slot<K>(li, P, c, t)] = r[f][t];
        }
        __syncthreads();
        double w[W];
#pragma unroll
        for (int u = 0; u < H; u++) {
            w[u] = lds[xt_slot<K>(lr, P, hl, K - H + u)];
            w[K + H + u] = lds[xt_slot<K>(lr, P, hr, u)];
        }
#pragma unroll
        for (int t = 0; t < K; t++) w[t + H] = r[f][t];
        double o[K];
        stencil_t<K, ST, W>(w, o, s, N, cc, P, ax.seam, cm.pb, an);
#pragma unroll
        for (int t = 0; t < K; t++) r[f][t] = o[t];
    }

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
    // every point's loads before the first store
#pragma unroll
    for (int j = 0; j < J; j++) {
        const int p = lane + 64 * j;
        if (p < npts) {
            double d[NF];
#pragma unroll
            for (int f = 0; f < NF; f++) d[f] = xo[f][j];
            const int row = lrw[j] & 1023;
            const int face = row == 0 ? 0 : (row == N - 1 ? 1 : -1);
            sw.fetch(run + (uoff)p * 8u, d, face, line0 + (lrw[j] >> 10));
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
            sw.store(run + (uoff)p * 8u, d);
        }
    }
}

// ---------------------------------------------------------------------------
// transposed sweep, y/z lines: TL lines x P chunks per workgroup, each thread
// loads and stores its own chunk rows (coalesced across the TL lines), carries
// by LDS folds, halo rows of the solution through LDS
// ---------------------------------------------------------------------------
template <int K, int AX, int TL, class SW, int ST>
__global__ void __launch_bounds__(TL * kMaxP) __attribute__((amdgpu_waves_per_eu(SW::kWpe))) directT_kernel(AxisDesc ax, SweepCommon cm, SW sw, const double* tabT,
                                                             const double* mu, const double* an) {
    constexpr int NF = SW::NF;
    constexpr int H = ST == 0 ? 2 : 5;
    constexpr int W = K + 2 * H;
    static_assert(H <= K, "halo rows come from the neighbouring chunks only");
    extern __shared__ double lds[];

    const int N = ax.N, P = ax.P;
    const int lloc = threadIdx.x % TL;
    const int c = threadIdx.x / TL;
    const long line = group_of(ax.xcd) * TL + lloc;
    const bool act = line < ax.nlines;
    const bool work = act;                // blockDim = TL*P: every thread owns a chunk
    const int s = c * K;
    const uoff st = (uoff)ax.st * 8u;
    const uoff base = line_off(AX, act ? (uoff)line : 0u, (uoff)cm.lxi, (uoff)cm.plane);

    double r[NF][K];
#pragma unroll
    for (int t = 0; t < K; t++) {
        double x[NF];
        if (work) {
            sw.seed(base + (uoff)(s + t) * st, x);
        } else {
#pragma unroll
            for (int f = 0; f < NF; f++) x[f] = 0.0;
        }
#pragma unroll
        for (int f = 0; f < NF; f++) r[f][t] = x[f];
    }
    if (mu && work) {
        const long nf = ax.nlines;
#pragma unroll
        for (int f = 0; f < NF; f++) {
            if (c == 0) r[f][0] += mu[f * nf + line];
            if (c == P - 1) r[f][K - 1] += mu[(5 + f) * nf + line];
        }
    }

    double carry[NF], G[NF], X[NF];
    forward_local<K, NF, true>(tabT, N, s, K, r, carry);
#pragma unroll
    for (int f = 0; f < NF; f++) lds[(f * P + c) * TL + lloc] = carry[f];
    __syncthreads();
    fold_forward<NF, true, K>(lds, tabT, N, P, TL, c, lloc, G);
    apply_carry<K, NF, true>(tabT, 3 * N, s, K, r, G);
    backward_local<K, NF, true>(tabT, N, s, K, r, carry);
    double* xb = lds + NF * P * TL;
#pragma unroll
    for (int f = 0; f < NF; f++) xb[(f * P + c) * TL + lloc] = carry[f];
    __syncthreads();
    fold_backward<NF, true, K>(xb, tabT, N, P, TL, c, lloc, X);
    apply_carry<K, NF, true>(tabT, 4 * N, s, K, r, X);

    // halo rows hb[(f*2H + k)*P + c][TL]: k < H the first H rows of chunk c,
    // k >= H its last H rows (neighbours wrap on seam lines of the
    // derivative); the derivative's five fields in one exchange, the filter
    // one field at a time
    const bool wrap = ST == 0 && ax.seam;
    const int cl = c > 0 ? c - 1 : (wrap ? P - 1 : 0), cr = c + 1 < P ? c + 1 : (wrap ? 0 : P - 1);
    constexpr int FX = ST == 0 ? NF : 1;      // fields per exchange
#pragma unroll
    for (int f0 = 0; f0 < NF; f0 += FX) {
        __syncthreads();                  // earlier reads of the LDS are done
#pragma unroll
        for (int e = 0; e < FX; e++) {
#pragma unroll
            for (int k = 0; k < H; k++) {
                lds[((e * 2 * H + k) * P + c) * TL + lloc] = r[f0 + e][k];
                lds[((e * 2 * H + H + k) * P + c) * TL + lloc] = r[f0 + e][K - H + k];
            }
        }
        __syncthreads();
#pragma unroll
        for (int e = 0; e < FX; e++) {
            const int f = f0 + e;
            double w[W];
#pragma unroll
            for (int k = 0; k < H; k++) {
                w[k] = lds[((e * 2 * H + H + k) * P + cl) * TL + lloc];
                w[K + H + k] = lds[((e * 2 * H + k) * P + cr) * TL + lloc];
            }
#pragma unroll
            for (int t = 0; t < K; t++) w[t + H] = r[f][t];
            double o[K];
            stencil_t<K, ST, W>(w, o, s, N, c, P, ax.seam, cm.pb, an);
#pragma unroll
            for (int t = 0; t < K; t++) r[f][t] = o[t];
        }
    }

    if (!work) return;
    // every row's loads before the first store
#pragma unroll
    for (int t = 0; t < K; t++) {
        const int row = s + t;
        double d[NF];
#pragma unroll
        for (int f = 0; f < NF; f++) d[f] = r[f][t];
        const int face = row == 0 ? 0 : (row == N - 1 ? 1 : -1);
        sw.fetch(base + (uoff)row * st, d, face, line);
#pragma unroll
        for (int f = 0; f < NF; f++) r[f][t] = d[f];
    }
#pragma unroll
    for (int t = 0; t < K; t++) {
        double d[NF];
#pragma unroll
        for (int f = 0; f < NF; f++) d[f] = r[f][t];
        sw.store(base + (uoff)(s + t) * st, d);
    }
}

// adjoint of the characteristic BC at the face points of axis ax (the same
// products as gcbc_vjp_kernel of vjp.hip, into face buffers): lam_f = sum of
// sbcc*lam over the kMbci rows next to the face; mu = (dc/dg)^T lam_f (added
// to the seed at the face row), G = (dc/dq)^T lam_f (added at the face
// point); drv = [2 face][5][nface] boundary derivatives of the forward sweep
__global__ void __launch_bounds__(256) face_adjoint_kernel(const double* q, const double* lam, const double* sbcc,
                                                           const double* drv, double* mu, double* G, int ax, int N,
                                                           long nface, long lxi, long plane, long st, EulerPhys ph,
                                                           long lmx, const double* cmf, const double* yaco) {
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
        gd[v] = drv[((long)face * 5 + v) * nface + f];
    }
    // face normal and Jacobian: the caller's arrays (curvilinear solvers) or the constants
    double cm[3], J = ph.J;
    if (yaco) {
        for (int k = 0; k < 3; k++) cm[k] = cmf[((long)face * 3 + k) * nface + f];
        J = yaco[pt];
    } else {
        const double sg = ph.m[ax] * ph.J < 0.0 ? -1.0 : 1.0;
        for (int k = 0; k < 3; k++) cm[k] = k == ax ? sg : 0.0;
    }
    typedef Dual<5> D;
    {
        D gdd[5], cr[5];
#pragma unroll
        for (int v = 0; v < 5; v++) gdd[v] = D::seed(gd[v], v);
        face_corr_at<double, D>(face, qq, gdd, cm, J, ph.umf, cr);
        D phi(0.0);
#pragma unroll
        for (int v = 0; v < 5; v++) phi = phi + cr[v] * lf[v];
#pragma unroll
        for (int v = 0; v < 5; v++) mu[((long)face * 5 + v) * nface + f] = phi.d[v];
    }
    {
        D qd[5], gdd[5], cr[5];
#pragma unroll
        for (int v = 0; v < 5; v++) {
            qd[v] = D::seed(qq[v], v);
            gdd[v] = D(gd[v]);
        }
        face_corr_at<D, D>(face, qd, gdd, cm, J, ph.umf, cr);
        D phi(0.0);
#pragma unroll
        for (int v = 0; v < 5; v++) phi = phi + cr[v] * lf[v];
#pragma unroll
        for (int v = 0; v < 5; v++) G[((long)face * 5 + v) * nface + f] = phi.d[v];
    }
}

// transposed sweep along AX: x lines one wave per workgroup, y/z lines
// ax.tl lines per workgroup; FT<1> spreads the five fields over grid.y
template <int AX, int ST, class SW>
int launch_t(const AxisDesc& ax, const SweepCommon& cm, const SW& sw, const double* tabT, const double* mu,
             const double* an, hipStream_t s) {
    constexpr int K = 8, NF = SW::NF, H = ST == 0 ? 2 : 5;
    const unsigned gy = NF == 1 ? 5u : 1u;
    if constexpr (AX == 0) {
        const int lpw = 64 / ax.GS;
        const size_t lds = sizeof(double) * lpw * ax.P * (K + 1);
        dim3 block(64), grid((unsigned)((ax.nlines + lpw - 1) / lpw), gy);
        hipLaunchKernelGGL((xtT_kernel<K, SW, ST>), grid, block, lds, s, ax, cm, sw, tabT, mu, an);
    } else {
        const int TL = ax.tl == 8 ? 8 : 16;
        const size_t halo = (size_t)(ST == 0 ? NF : 1) * 2 * H * ax.P * TL;
        const size_t carries = (size_t)2 * NF * ax.P * TL;
        const size_t lds = sizeof(double) * (halo > carries ? halo : carries);
        dim3 block(TL * ax.P), grid((unsigned)((ax.nlines + TL - 1) / TL), gy);
        const void* k = TL == 8 ? (const void*)directT_kernel<K, AX, 8, SW, ST>
                                : (const void*)directT_kernel<K, AX, 16, SW, ST>;
        if (lds > 65536) CS_HIP(hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        if (TL == 8)
            hipLaunchKernelGGL((directT_kernel<K, AX, 8, SW, ST>), grid, block, lds, s, ax, cm, sw, tabT, mu, an);
        else
            hipLaunchKernelGGL((directT_kernel<K, AX, 16, SW, ST>), grid, block, lds, s, ax, cm, sw, tabT, mu, an);
    }
    CS_HIP(hipGetLastError());
    return CS_OK;
}

EulerPhys euler_phys_of(const cs_solver* sv) {
    EulerPhys ph;
    ph.umf = sv->cfg.umf;
    ph.J = sv->cfg.yaco;
    for (int a = 0; a < 3; a++) ph.m[a] = sv->cfg.metric[a];
    return ph;
}

double uf_of(const cs_solver* sv, int a) { return a == 0 ? sv->cfg.umf.x : (a == 1 ? sv->cfg.umf.y : sv->cfg.umf.z); }

// adjoint scratch: seed t and accumulator g (5 fields each), the face
// buffers mu|G of the three axes, and the tape of cs_solver_rk4_vjp
long adj_face_buf(const cs_solver* sv) { return 5 * sv->g.lmx * 2; }
long adj_tape(const cs_solver* sv) { return adj_face_buf(sv) + 2 * faces_size(sv->g); }
long adj_doubles(const cs_solver* sv) { return adj_tape(sv) + 20 * sv->g.lmx + 4 * faces_size(sv->g); }

int upload(const std::vector<double>& h, double** d) {
    CS_HIP(hipMalloc(d, h.size() * sizeof(double)));
    CS_HIP(hipMemcpy(*d, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice));
    return CS_OK;
}

int adjoint_setup(cs_solver* sv) {
    if (sv->adj) return CS_OK;
    int rc;
    for (int a = 0; a < 3; a++) {
        const AxisDesc& ax = sv->ax[a];
        std::vector<double> h;
        if ((rc = build_adjoint_table(OP_DERIV_T, ax.N, ax.seam, ax.seam, sv->cfg.alpha_f, sv->num->pb_host, 8, h)))
            return rc;
        if ((rc = upload(h, &sv->ttabs[a]))) return rc;
        if (sv->cfg.filter && !ax.seam) {
            if ((rc = build_adjoint_table(OP_FILTER_T, ax.N, 0, 0, sv->cfg.alpha_f, sv->num->pb_host, 8, h)))
                return rc;
            if ((rc = upload(h, &sv->fttabs[a]))) return rc;
        }
    }
    CS_HIP(hipMalloc(&sv->adj, sizeof(double) * adj_doubles(sv)));
    return CS_OK;
}

// transposed filter along axis a in place on the five fields of lam
template <int AX>
int filter_t_axis(cs_solver* sv, double* lam, const SweepCommon& cm, hipStream_t s) {
    const AxisDesc& ax = sv->ax[AX];
    if (ax.seam) return filter_sweeps(sv, lam, s, AX);        // cyclic: its own transpose
    const double* an = ax.ftab + 6 * ax.N + 8;
    if (AX == 0) {
        FT<5> f{lam, sv->g.lmx};
        return launch_t<AX, 1>(ax, cm, f, sv->fttabs[AX], nullptr, an, s);
    }
    FT<1> f{lam, sv->g.lmx};
    return launch_t<AX, 1>(ax, cm, f, sv->fttabs[AX], nullptr, an, s);
}

template <int AX, int MODE, bool CURV>
int stage_sweep_t(cs_solver* sv, const SweepCommon& cm, const double* q, const double* t, const double* gin,
                  double* gout, double* lam, double* tout, double ck, const double* fb, hipStream_t s) {
    const AxisDesc& ax = sv->ax[AX];
    const long lmx = sv->g.lmx;
    AT<AX, MODE, CURV> a{};
    fields(q, lmx, a.q);
    fields(t, lmx, a.t);
    fields(gin, lmx, a.gin);
    fields(gout, lmx, a.gout);
    fields(lam, lmx, a.lam);
    fields(tout, lmx, a.tout);
    const double* mu = nullptr;
    if (ax.gcbc) {
        mu = fb + 2 * face_offset(sv->g, AX);
        a.G = mu + 10 * ax.nlines;
    }
    a.nface = ax.nlines;
    a.ck = ck;
    a.m = sv->cfg.metric[AX];
    a.uf = uf_of(sv, AX);
    a.umf = sv->cfg.umf;
    if constexpr (CURV) a.mt = sv->met;
    return launch_t<AX, 0>(ax, cm, a, sv->ttabs[AX], mu, nullptr, s);
}

// ck = dt/div times the constant Jacobian; curvilinear solvers pass a seed
// already scaled by the point's Jacobian (backward_fused) and ck = dt/div
template <int AX, int MODE>
int stage_sweep(cs_solver* sv, const SweepCommon& cm, const double* q, const double* t, const double* gin,
                double* gout, double* lam, double* tout, double ck, const double* fb, hipStream_t s) {
    if (sv->curv) return stage_sweep_t<AX, MODE, true>(sv, cm, q, t, gin, gout, lam, tout, ck, fb, s);
    // (constant Jacobian: scaling the seed or the result is the same)
    return stage_sweep_t<AX, MODE, false>(sv, cm, q, t, gin, gout, lam, tout, ck * sv->cfg.yaco, fb, s);
}

// face seeds of every non-reflective axis for the RHS evaluated at q with
// boundary derivatives `faces` ([3 axes][2][5][face]) and seed t
int face_adjoints(cs_solver* sv, const double* q, const double* t, const double* faces, const double* sbcc,
                  double* fb, hipStream_t s) {
    const EulerPhys ph = euler_phys_of(sv);
    for (int a = 0; a < 3; a++) {
        const AxisDesc& ax = sv->ax[a];
        if (!ax.gcbc) continue;
        const long nf = ax.nlines;
        double* mu = fb + 2 * face_offset(sv->g, a);
        hipLaunchKernelGGL(face_adjoint_kernel, dim3((unsigned)((2 * nf + 255) / 256)), dim3(256), 0, s, q, t, sbcc,
                           faces + face_offset(sv->g, a), mu, mu + 10 * nf, a, ax.N, nf, (long)sv->g.lxi,
                           (long)sv->g.lxi * sv->g.let, ax.st, ph, sv->g.lmx, sv->curv ? sv->met.cm[a] : nullptr,
                           sv->curv ? sv->met.J : nullptr);
        CS_HIP(hipGetLastError());
    }
    return CS_OK;
}

// out = J t, five fields: the seed of a stage's transposed RHS on a
// curvilinear grid (stage k adds -dt/div * J(x) R(q) at each point x, so its
// adjoint applies R^T to J t, the Jacobian on the output side of R)
__global__ void __launch_bounds__(256) scale_seed_kernel(const double* t, const double* J, double* out, long lmx) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= lmx) return;
    const double j = J[i];
#pragma unroll
    for (int v = 0; v < 5; v++) out[i + v * lmx] = j * t[i + v * lmx];
}

// lam <- step'(q0)^T lam from a tape (fused sweeps)
int backward_fused(cs_solver* sv, const double* tape, double* lam, double dt, const double* sbcc, hipStream_t s) {
    const long lmx = sv->g.lmx, n5 = 5 * lmx, F = faces_size(sv->g);
    const double* faces = tape + 20 * lmx;
    double* tb = sv->adj;
    double* gb = tb + n5;
    double* fb = sv->adj + adj_face_buf(sv);
    const SweepCommon cm = common_of(sv, sbcc);
    int rc;
    if (sv->cfg.filter) {
        if ((rc = filter_t_axis<2>(sv, lam, cm, s))) return rc;
        if ((rc = filter_t_axis<1>(sv, lam, cm, s))) return rc;
        if ((rc = filter_t_axis<0>(sv, lam, cm, s))) return rc;
    }
    for (int k = 4; k >= 1; k--) {
        const double* q = tape + (long)(k - 1) * n5;
        const double* t = k == 4 ? lam : tb;
        const double ck = dt / (double)(5 - k);   // times the Jacobian in stage_sweep
        if (sv->curv) {
            // the seed scaled by the point's Jacobian, in the stage buffer qc
            // (unused outside cs_solver_rk4_step)
            hipLaunchKernelGGL(scale_seed_kernel, dim3((unsigned)((lmx + 255) / 256)), dim3(256), 0, s, t, sv->met.J,
                               sv->qc, lmx);
            CS_HIP(hipGetLastError());
            t = sv->qc;
        }
        if ((rc = face_adjoints(sv, q, t, faces + (long)(k - 1) * F, sbcc, fb, s))) return rc;
        if ((rc = stage_sweep<2, 0>(sv, cm, q, t, nullptr, gb, nullptr, nullptr, 0.0, fb, s))) return rc;
        if ((rc = stage_sweep<1, 1>(sv, cm, q, t, gb, gb, nullptr, nullptr, 0.0, fb, s))) return rc;
        if (k > 1)
            rc = stage_sweep<0, 2>(sv, cm, q, t, gb, nullptr, lam, tb, ck, fb, s);
        else
            rc = stage_sweep<0, 3>(sv, cm, q, t, gb, nullptr, lam, nullptr, ck, fb, s);
        if (rc) return rc;
    }
    return CS_OK;
}

// one RK4 step in place on qa recording the tape; dt from the CFL limit when
// dt <= 0 (as cs_solver_rk4_step)
int step_recording(cs_solver* sv, double* qa, const double* sbcc, double dt, double* tape, hipStream_t s) {
    const long lmx = sv->g.lmx, n5 = 5 * lmx, F = faces_size(sv->g);
    double* faces = tape + 20 * lmx;
    int rc = time_step(sv, qa, dt, s);
    if (rc) return rc;
    CS_HIP(hipMemcpyAsync(tape, qa, sizeof(double) * n5, hipMemcpyDeviceToDevice, s));
    for (int k = 1; k <= 4 && !rc; k++) {
        sv->rec = faces + (long)(k - 1) * F;
        rc = rhs_sweeps(sv, tape + (long)(k - 1) * n5, sbcc, k < 4 ? tape + (long)k * n5 : qa, tape, 5 - k, s);
    }
    sv->rec = nullptr;
    if (rc) return rc;
    return step_filter(sv, qa, s);
}

}  // namespace

bool adjoint_fused(const cs_solver* sv) {
    if (sv->adj_mode != 1 || sv->cfg.viscous) return false;
    for (int a = 0; a < 3; a++)
        if (sv->ax[a].K != 8 || sv->ax[a].N != sv->ax[a].P * 8 || sv->ax[a].P > kMaxP) return false;
    return !sv->cfg.filter || filter_fusable(sv);
}

long tape_doubles(const cs_solver* sv) { return 20 * sv->g.lmx + 4 * faces_size(sv->g); }

int rk4_step_tape(cs_solver* sv, double* qa, const double* sbcc, double dt, double* tape, hipStream_t s) {
    if (sv->cfg.viscous) return fail(CS_ERR_UNSUPPORTED, "the adjoint covers the inviscid right-hand side");
    return step_recording(sv, qa, sbcc, dt, tape, s);
}

int rk4_vjp_tape(cs_solver* sv, const double* tape, double* lam, double dt, const double* sbcc, hipStream_t s) {
    if (sv->cfg.viscous) return fail(CS_ERR_UNSUPPORTED, "the adjoint covers the inviscid right-hand side");
    if (!(dt > 0.0)) return fail(CS_ERR_ARG, "the adjoint needs the step's dt > 0");
    if (!adjoint_fused(sv)) {
        const long n5 = 5 * sv->g.lmx;
        return rk4_vjp_lines(sv, tape, tape + n5, lam, dt, sbcc, s);
    }
    int rc = adjoint_setup(sv);
    if (rc) return rc;
    return backward_fused(sv, tape, lam, dt, sbcc, s);
}

// the fused forms of cs_solver_rk4_vjp / cs_solver_rhs_vjp (adjoint_fused)
int rk4_vjp_fused(cs_solver* sv, const double* q0, double* lam, double dt, const double* sbcc, hipStream_t s) {
    int rc = adjoint_setup(sv);
    if (rc) return rc;
    const long lmx = sv->g.lmx, n5 = 5 * lmx, F = faces_size(sv->g);
    double* tape = sv->adj + adj_tape(sv);
    double* faces = tape + 20 * lmx;
    // stage states and faces of q0..q2 by the forward stages, faces of q3 by
    // one more right-hand side; dt into the device slot the stages read
    if ((rc = time_step(sv, q0, dt, s))) return rc;
    CS_HIP(hipMemcpyAsync(tape, q0, sizeof(double) * n5, hipMemcpyDeviceToDevice, s));
    for (int k = 1; k <= 4 && !rc; k++) {
        sv->rec = faces + (long)(k - 1) * F;
        if (k < 4)
            rc = rhs_sweeps(sv, tape + (long)(k - 1) * n5, sbcc, tape + (long)k * n5, tape, 5 - k, s);
        else
            rc = rhs_sweeps(sv, tape + 3 * n5, sbcc, sv->dpart, nullptr, 0, s);
    }
    sv->rec = nullptr;
    if (rc) return rc;
    return backward_fused(sv, tape, lam, dt, sbcc, s);
}

int rhs_vjp_fused(cs_solver* sv, const double* q, const double* lam, double* g, const double* sbcc, hipStream_t s) {
    int rc = adjoint_setup(sv);
    if (rc) return rc;
    double* faces = sv->adj + adj_tape(sv) + 20 * sv->g.lmx;
    double* fb = sv->adj + adj_face_buf(sv);
    sv->rec = faces;
    rc = rhs_sweeps(sv, q, sbcc, sv->dpart, nullptr, 0, s);
    sv->rec = nullptr;
    if (rc) return rc;
    const SweepCommon cm = common_of(sv, sbcc);
    if ((rc = face_adjoints(sv, q, lam, faces, sbcc, fb, s))) return rc;
    if ((rc = stage_sweep<2, 0>(sv, cm, q, lam, nullptr, g, nullptr, nullptr, 0.0, fb, s))) return rc;
    if ((rc = stage_sweep<1, 1>(sv, cm, q, lam, g, g, nullptr, nullptr, 0.0, fb, s))) return rc;
    return stage_sweep<0, 1>(sv, cm, q, lam, g, g, nullptr, nullptr, 0.0, fb, s);
}

}  // namespace fused
}  // namespace cs

extern "C" {

int cs_solver_tape_size(const cs_solver* sv, size_t* doubles) {
    if (!sv || !doubles) return cs::fail(CS_ERR_ARG, "null argument");
    *doubles = (size_t)cs::fused::tape_doubles(sv);
    return CS_OK;
}

int cs_solver_rk4_step_tape(cs_solver* sv, double* qa, const double* sbcc, double dt, double* dt_out, double* tape,
                            cs_stream stream) {
    if (!sv || !qa || !tape) return cs::fail(CS_ERR_ARG, "null argument");
    if (!sbcc && (sv->ax[0].gcbc || sv->ax[1].gcbc || sv->ax[2].gcbc))
        return cs::fail(CS_ERR_ARG, "non-reflective faces need sbcc");
    hipStream_t s = (hipStream_t)stream;
    int rc = cs::fused::rk4_step_tape(sv, qa, sbcc, dt, tape, s);
    if (rc || !dt_out) return rc;
    CS_HIP(hipMemcpyAsync(sv->dt_host, sv->dt, 3 * sizeof(double), hipMemcpyDeviceToHost, s));
    CS_HIP(hipStreamSynchronize(s));
    *dt_out = sv->dt_host[0];
    return CS_OK;
}

int cs_solver_rk4_vjp_tape(cs_solver* sv, const double* tape, double* lam, double dt, const double* sbcc,
                           cs_stream stream) {
    if (!sv || !tape || !lam) return cs::fail(CS_ERR_ARG, "null argument");
    if (!sbcc && (sv->ax[0].gcbc || sv->ax[1].gcbc || sv->ax[2].gcbc))
        return cs::fail(CS_ERR_ARG, "non-reflective faces need sbcc");
    return cs::fused::rk4_vjp_tape(sv, tape, lam, dt, sbcc, (hipStream_t)stream);
}

int cs_solver_adjoint_mode(cs_solver* sv, int mode) {
    if (!sv || (mode != 0 && mode != 1)) return cs::fail(CS_ERR_ARG, "adjoint mode is 0 or 1");
    sv->adj_mode = mode;
    return CS_OK;
}

}  // extern "C"
