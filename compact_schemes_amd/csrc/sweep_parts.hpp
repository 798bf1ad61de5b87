// sweep_parts.hpp -- pieces shared by the line-sweep kernels of the fused
// solver (forward sweeps in fused.hip, transposed sweeps in tape.hip):
// 32-bit field offsets, line addressing, the chunked LU recurrences of the
// line solve and their carries across chunks (wave scans for x lines, LDS
// folds for y/z lines), and the host-side argument helpers.
#pragma once

#include "fused.hpp"

namespace cs {
namespace fused {

typedef unsigned int uoff;    // byte offset inside one field

__device__ __forceinline__ double ldg(const double* p, uoff bo) {
    return *reinterpret_cast<const double*>(reinterpret_cast<const char*>(p) + bo);
}
__device__ __forceinline__ void stg(double* p, uoff bo, double v) {
    *reinterpret_cast<double*>(reinterpret_cast<char*>(p) + bo) = v;
}

struct Phys {
    cs_point umf;
    double mu, J;
    double m[3];   // metric scale per axis: xi_x, eta_y, zeta_z
};

__device__ __forceinline__ Prim prim_at(const double* const* q, uoff bo, double* qq) {
#pragma unroll
    for (int v = 0; v < 5; v++) qq[v] = ldg(q[v], bo);
    return primitive(qq);
}

// the two quantities of the time step at one point with constant diagonal
// metrics (calc_time_step_kernel, src/cuda/kernels/physics.hpp:335-427, D9
// corrected): the convective (c |grad xi| + |u . grad xi|) |J| and the
// viscous mu/rho |grad xi|^2 J^2.  One expression for every kernel that
// reduces them, so that all give the same dt bit for bit.  The viscous one is
// a product of non-negative factors rounded left to right, so its maximum
// over points is its value at the largest 1/rho (rounding is monotone).
__device__ __forceinline__ double dt_conv(const Phys& ph, double u, double v, double w, double c2) {
    const double rr1 = ph.m[0] * ph.m[0] + ph.m[1] * ph.m[1] + ph.m[2] * ph.m[2];
    const double rr2 = fabs(ph.m[0] * (u + ph.umf.x)) + fabs(ph.m[1] * (v + ph.umf.y)) +
                       fabs(ph.m[2] * (w + ph.umf.z));
    return (sqrt(c2 * rr1) + rr2) * fabs(ph.J);
}
__device__ __forceinline__ double dt_visc(const Phys& ph, double rhoi) {
    const double rr1 = ph.m[0] * ph.m[0] + ph.m[1] * ph.m[1] + ph.m[2] * ph.m[2];
    const double ssi = fabs(ph.J);
    return rhoi * ph.mu * rr1 * ssi * ssi;
}
// the same two maxima updated at one point with the point's metric rows M[a]
// = grad xi_a and Jacobian J (calc_time_step_kernel's full form); one
// expression for the separate pass and the sweep that folds it
__device__ __forceinline__ void dt_curv_point(const double (&M)[3][3], double J, const Phys& ph, const Prim& p,
                                              double& m0, double& m1) {
    const double u = p.u + ph.umf.x, v = p.v + ph.umf.y, w = p.w + ph.umf.z;
    double rr1 = 0.0, rr2 = 0.0;
#pragma unroll
    for (int a = 0; a < 3; a++) {
        const double mx = M[a][0], my = M[a][1], mz = M[a][2];
        rr1 += mx * mx + my * my + mz * mz;
        rr2 += fabs(mx * u + my * v + mz * w);
    }
    const double ssi = fabs(J);
    m0 = fmax(m0, (sqrt(p.c2 * rr1) + rr2) * ssi);
    m1 = fmax(m1, p.rhoi * ph.mu * rr1 * ssi * ssi);
}

// byte offset of row 0 of line f of an axis (f runs over the face's points
// with the first remaining index fastest)
__device__ __forceinline__ uoff line_off(int axis, uoff f, uoff lxi, uoff plane) {
    if (axis == 0) return f * lxi * 8u;
    if (axis == 1) return ((f / lxi) * plane + (f % lxi)) * 8u;
    return f * 8u;
}

// recurrences of the solve (EX: every chunk has K rows)
template <int K, int NF, bool EX>
__device__ __forceinline__ void forward_local(const double* tab, int N, int s, int L, double (&r)[NF][K],
                                              double (&carry)[NF]) {
#pragma clang fp contract(fast)
#pragma unroll
    for (int f = 0; f < NF; f++) carry[f] = 0.0;
#pragma unroll
    for (int t = 0; t < K; t++) {
        if (EX || t < L) {
            const double nu = tab[s + t], mu = tab[N + s + t];
#pragma unroll
            for (int f = 0; f < NF; f++) { carry[f] = nu * r[f][t] + mu * carry[f]; r[f][t] = carry[f]; }
        }
    }
}

template <int K, int NF, bool EX>
__device__ __forceinline__ void backward_local(const double* tab, int N, int s, int L, double (&r)[NF][K],
                                               double (&carry)[NF]) {
#pragma clang fp contract(fast)
#pragma unroll
    for (int f = 0; f < NF; f++) carry[f] = 0.0;
#pragma unroll
    for (int t = K - 1; t >= 0; t--) {
        if (EX || t < L) {
            const double rh = tab[2 * N + s + t];
#pragma unroll
            for (int f = 0; f < NF; f++) { carry[f] = r[f][t] + rh * carry[f]; r[f][t] = carry[f]; }
        }
    }
}

template <int K, int NF, bool EX>
__device__ __forceinline__ void apply_carry(const double* tab, int off, int s, int L, double (&r)[NF][K],
                                            const double (&G)[NF]) {
#pragma clang fp contract(fast)
#pragma unroll
    for (int t = 0; t < K; t++) {
        if (EX || t < L) {
            const double M = tab[off + s + t];
#pragma unroll
            for (int f = 0; f < NF; f++) r[f][t] += M * G[f];
        }
    }
}

// carries across the chunks of a line through LDS: buf = [NF][P][TL]
template <int NF, bool EX, int K>
__device__ __forceinline__ void fold_forward(const double* buf, const double* tab, int N, int P, int TL, int c,
                                             int lloc, double (&G)[NF]) {
#pragma clang fp contract(fast)
#pragma unroll
    for (int f = 0; f < NF; f++) G[f] = 0.0;
    for (int cc = 0; cc < c; cc++) {
        const int e = EX ? cc * K + K - 1 : (int)(((long)(cc + 1) * N) / P) - 1;
        const double M = tab[3 * N + e];
#pragma unroll
        for (int f = 0; f < NF; f++) G[f] = buf[(f * P + cc) * TL + lloc] + M * G[f];
    }
}

template <int NF, bool EX, int K>
__device__ __forceinline__ void fold_backward(const double* buf, const double* tab, int N, int P, int TL, int c,
                                              int lloc, double (&X)[NF]) {
#pragma clang fp contract(fast)
#pragma unroll
    for (int f = 0; f < NF; f++) X[f] = 0.0;
    for (int cc = P - 1; cc > c; cc--) {
        const int sc = EX ? cc * K : (int)(((long)cc * N) / P);
        const double R = tab[4 * N + sc];
#pragma unroll
        for (int f = 0; f < NF; f++) X[f] = buf[(f * P + cc) * TL + lloc] + R * X[f];
    }
}

// carries across the chunks of a line held by GS consecutive lanes of a wave:
// segmented Kogge-Stone scans of the affine maps x -> A*x + B
template <int NF>
__device__ __forceinline__ void scan_forward(double A, const double (&carry)[NF], bool cv, int c, int GS,
                                             double (&G)[NF]) {
#pragma clang fp contract(fast)
    double B[NF];
#pragma unroll
    for (int f = 0; f < NF; f++) B[f] = cv ? carry[f] : 0.0;
    if (!cv) A = 1.0;
    for (int d = 1; d < GS; d <<= 1) {
        const double Au = __shfl_up(A, d, GS);
        double Bu[NF];
#pragma unroll
        for (int f = 0; f < NF; f++) Bu[f] = __shfl_up(B[f], d, GS);
        if (c >= d) {
#pragma unroll
            for (int f = 0; f < NF; f++) B[f] += A * Bu[f];
            A *= Au;
        }
    }
#pragma unroll
    for (int f = 0; f < NF; f++) {
        const double g = __shfl_up(B[f], 1, GS);
        G[f] = (c == 0) ? 0.0 : g;
    }
}

template <int NF>
__device__ __forceinline__ void scan_backward(double A, const double (&carry)[NF], bool cv, int c, int P, int GS,
                                              double (&X)[NF]) {
#pragma clang fp contract(fast)
    double B[NF];
#pragma unroll
    for (int f = 0; f < NF; f++) B[f] = cv ? carry[f] : 0.0;
    if (!cv) A = 0.0;
    for (int d = 1; d < GS; d <<= 1) {
        const double Ad = __shfl_down(A, d, GS);
        double Bd[NF];
#pragma unroll
        for (int f = 0; f < NF; f++) Bd[f] = __shfl_down(B[f], d, GS);
        if (c + d < GS) {
#pragma unroll
            for (int f = 0; f < NF; f++) B[f] += A * Bd[f];
            A *= Ad;
        }
    }
#pragma unroll
    for (int f = 0; f < NF; f++) {
        const double x = __shfl_down(B[f], 1, GS);
        X[f] = (c + 1 < P) ? x : 0.0;
    }
}

// workgroup -> line group, neighbouring groups on the same XCD (workgroups
// are dispatched round-robin over the 8 XCDs, each with its own L2)
__device__ __forceinline__ long group_of(int xcd) {
    const long w = blockIdx.x;
    if (!xcd) return w;
    const long G8 = ((long)gridDim.x / 8) * 8;
    if (w >= G8) return w;
    return (w % 8) * (G8 / 8) + w / 8;
}

// LDS slot of row t of chunk c of line li in the transposing x kernels
template <int K>
__device__ __forceinline__ int xt_slot(int li, int P, int c, int t) { return (li * P + c) * (K + 1) + t; }

// ---------------------------------------------------------------------------
// host helpers
// ---------------------------------------------------------------------------
inline Phys phys_of(const cs_solver* sv) {
    Phys ph;
    ph.umf = sv->cfg.umf;
    ph.mu = sv->cfg.mu;
    ph.J = sv->cfg.yaco;
    for (int a = 0; a < 3; a++) ph.m[a] = sv->cfg.metric[a];
    return ph;
}

inline SweepCommon common_of(const cs_solver* sv, const double* sbcc) {
    SweepCommon cm;
    cm.lxi = sv->g.lxi;
    cm.plane = (long)sv->g.lxi * sv->g.let;
    cm.pb = sv->num->pb;
    cm.sbcc = sbcc;
    return cm;
}

// offset of axis a's [2 face][5 var][face] block in a buffer holding the
// three axes' boundary-derivative faces, and that buffer's size
inline long face_offset(const Geo& g, int a) {
    long o = 0;
    for (int b = 0; b < a; b++) o += 10 * g.face_size(b);
    return o;
}
inline long faces_size(const Geo& g) { return face_offset(g, 3); }

template <class T, int N>
inline void fields(T* base, long lmx, T* (&out)[N]) {
    for (int k = 0; k < N; k++) out[k] = base ? base + (long)k * lmx : nullptr;
}

}  // namespace fused
}  // namespace cs
