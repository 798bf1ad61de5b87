// adjoint.hip -- transposed line operators (no reference counterpart).
//
// A compact derivative or filter along one axis is x = A^-1 B f per line,
// with A the tridiagonal left-hand side and B the explicit right-hand-side
// stencil, boundary closure or seam rows included (line_solver.hip; the
// reference's build_system, src/cuda/kernels/numerics_device_pc.hpp:176-293).
// Its transpose is f' = B^T A^-T x': a solve with the LU factors of A^T
// (line_kernel with OP_SOLVE, the chunked recurrences of the forward solve)
// followed by the transposed explicit stencil, bt_kernel.  In the interior
// B^T is the stencil itself with the sign of the odd (derivative) or even
// (filter) part; columns near the line ends receive the boundary rows, so
// there the coefficients come from a dense block of B^T built on the host
// from the forward rows (or, for lines of at most 64 points, the whole B^T).
// The cyclic filter is a symmetric circulant (A^T = A, B^T = B, and
// circulants commute), so its transpose is the forward filter.
#include "cs_internal.hpp"

#include <algorithm>
#include <cmath>

namespace cs {

namespace {

constexpr int kDenseMax = 64;        // lines up to this length keep the whole B^T

// end-block sizes: columns [0, E) and [N-E, N) use explicit coefficients over
// rows [0, W) and [N-W, N); the interior formula holds for the other columns
int end_cols(int op) { return op == OP_DERIV_T ? 8 : 11; }
int end_rows(int op) { return op == OP_DERIV_T ? 10 : 16; }

// forward rows of the operator as a dense n x n matrix B (row-major) and the
// tridiagonal a, b, c of A
void forward_system(int op, int n, int start, int end, double af, const double* pb, std::vector<double>& B,
                    std::vector<double>& a, std::vector<double>& b, std::vector<double>& c) {
    B.assign((size_t)n * n, 0.0);
    a.assign(n, 0.0);
    b.assign(n, 1.0);
    c.assign(n, 0.0);
    auto at = [&](int i, int j) -> double& { return B[(size_t)i * n + j]; };
    if (op == OP_DERIV_T) {
        for (int i = 0; i < n; i++) { a[i] = kAlpha; c[i] = kAlpha; }
        for (int i = 2; i < n - 2; i++) {
            at(i, i - 2) = -kBb;
            at(i, i - 1) = -kAa;
            at(i, i + 1) = kAa;
            at(i, i + 2) = kBb;
        }
        const double* pbco = pb;
        const double* pbci = pb + 16;
        if (start == 0) {
            a[0] = 0; c[0] = kAlpha01; a[1] = kAlpha10; c[1] = kAlpha10;
            at(0, 0) = kAb00; at(0, 1) = kAb01; at(0, 2) = kAb02;
            at(1, 0) = -kAb10; at(1, 2) = kAb10;
        } else {
            a[0] = 0; c[0] = kAlpha;
            for (int i = 0; i < kLmd; i++) {
                at(0, i) += pbci[i];
                at(0, n - 1 - i) += pbco[i];
                at(1, i) += pbci[kLmd + i];
                at(1, n - 1 - i) += pbco[kLmd + i];
            }
        }
        if (end == 0) {
            a[n - 1] = kAlpha01; c[n - 1] = 0; a[n - 2] = kAlpha10; c[n - 2] = kAlpha10;
            at(n - 1, n - 1) = -kAb00; at(n - 1, n - 2) = -kAb01; at(n - 1, n - 3) = -kAb02;
            at(n - 2, n - 3) = -kAb10; at(n - 2, n - 1) = kAb10;
        } else {
            a[n - 1] = kAlpha; c[n - 1] = 0;
            for (int i = 0; i < kLmd; i++) {
                at(n - 1, n - 1 - i) -= pbci[i];
                at(n - 1, i) -= pbco[i];
                at(n - 2, n - 1 - i) -= pbci[kLmd + i];
                at(n - 2, i) -= pbco[kLmd + i];
            }
        }
    } else {
        // closure filter (the cyclic one never gets here): identity end rows,
        // reduced-order centred rows with graded alpha near the ends, 10th
        // order inside
        double an[6][6] = {};
        for (int M = 1; M <= 5; M++) filter_coefficients(M, filter_row_alpha(M, af), an[M]);
        for (int i = 0; i < n; i++) a[i] = c[i] = filter_row_alpha(std::min(i, n - 1 - i), af);
        a[0] = 0; c[0] = 0; a[n - 1] = 0; c[n - 1] = 0;
        for (int j = 0; j < n; j++) {
            const int dist = std::min(j, n - 1 - j);
            if (dist == 0) { at(j, j) = 1.0; continue; }
            const int M = std::min(dist, 5);
            at(j, j) += an[M][0];
            for (int m = 1; m <= M; m++) {
                at(j, j + m) += 0.5 * an[M][m];
                at(j, j - m) += 0.5 * an[M][m];
            }
        }
    }
}

void thomas(int n, const std::vector<double>& a, const std::vector<double>& b, const std::vector<double>& c,
            double* nu, double* mu, double* rho) {
    double cp = 0.0;
    for (int i = 0; i < n; i++) {
        const double w = 1.0 / (b[i] - (i > 0 ? a[i] * cp : 0.0));
        nu[i] = w;
        mu[i] = i > 0 ? -a[i] * w : 0.0;
        cp = c[i] * w;
        rho[i] = -cp;
    }
    rho[n - 1] = 0.0;
}

}  // namespace

// Table of a transposed operator: [0, 9n+44) the layout of build_line_table
// for A^T (nu, mu, rho, Mf, Rb, z, ...); then at 9n+44: dense flag, E, W and
// either B^T (n x n, row j = column j of B) or the end blocks
// [2E columns][2W rows] (rows [0, W) then [n-W, n)).
int build_adjoint_table(int op, int n, int start, int end, double af, const double* pb, int K,
                        std::vector<double>& h) {
    if (op == OP_FILTER_T && start == 1 && end == 1)
        return fail(CS_ERR_ARG, "the cyclic filter is its own transpose");
    if ((start == 1) != (end == 1) && op == OP_FILTER_T)
        return fail(CS_ERR_UNSUPPORTED, "filter: a seam on one face only has no single-GPU wrap");
    if (n < 16) return fail(CS_ERR_SHAPE, "transposed operators need lines of at least 16 points");
    std::vector<double> B, a, b, c;
    forward_system(op, n, start, end, af, pb, B, a, b, c);
    // A^T: sub-diagonal c[i-1], super-diagonal a[i+1]
    std::vector<double> at(n, 0.0), ct(n, 0.0);
    for (int i = 1; i < n; i++) at[i] = c[i - 1];
    for (int i = 0; i + 1 < n; i++) ct[i] = a[i + 1];
    const bool dense = n <= kDenseMax;
    const int E = end_cols(op), W = end_rows(op);
    const size_t tail = 3 + (dense ? (size_t)n * n : (size_t)2 * E * 2 * W);
    h.assign(9 * (size_t)n + 44 + tail, 0.0);
    double* nu = &h[0];
    double* mu = &h[n];
    double* rho = &h[2 * (size_t)n];
    double* Mf = &h[3 * (size_t)n];
    double* Rb = &h[4 * (size_t)n];
    thomas(n, at, b, ct, nu, mu, rho);
    for (int s = 0; s < n; s += K) {
        const int e = std::min(s + K, n) - 1;
        double p = 1.0;
        for (int j = s; j <= e; j++) { p *= mu[j]; Mf[j] = p; }
        p = 1.0;
        for (int j = e; j >= s; j--) { p *= rho[j]; Rb[j] = p; }
    }
    double* T = &h[9 * (size_t)n + 44];
    T[0] = dense ? 1.0 : 0.0;
    T[1] = E;
    T[2] = W;
    double* Bt = T + 3;
    if (dense) {
        for (int j = 0; j < n; j++)
            for (int i = 0; i < n; i++) Bt[(size_t)j * n + i] = B[(size_t)i * n + j];
    } else {
        for (int q = 0; q < 2 * E; q++) {
            const int j = q < E ? q : n - 2 * E + q;
            for (int r = 0; r < 2 * W; r++) {
                const int i = r < W ? r : n - 2 * W + r;
                Bt[(size_t)q * 2 * W + r] = B[(size_t)i * n + j];
            }
        }
        // the interior formula must reproduce B^T away from the ends
        double an5[6] = {};
        if (op == OP_FILTER_T) filter_coefficients(5, af, an5);
        for (int j = E; j < n - E; j++)
            for (int i = std::max(0, j - 6); i <= std::min(n - 1, j + 6); i++) {
                double want = 0.0;
                const int m = i - j;                       // B^T[j][i] = B[i][j]
                if (op == OP_DERIV_T) {
                    if (m == 1) want = -kAa;
                    if (m == -1) want = kAa;
                    if (m == 2) want = -kBb;
                    if (m == -2) want = kBb;
                } else {
                    const int am = std::abs(m);
                    if (am == 0) want = an5[0];
                    else if (am <= 5) want = 0.5 * an5[am];
                }
                if (std::fabs(B[(size_t)i * n + j] - want) > 1e-14)
                    return fail(CS_ERR_STATE, "transposed stencil: interior pattern mismatch");
            }
    }
    return CS_OK;
}

template <int AX>
__global__ void __launch_bounds__(256) bt_kernel(const double* y, double* out, const double* T, int op, int N,
                                                 long lmx, int lxi, int let, double a0, double a1, double a2,
                                                 double a3, double a4, double a5) {
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= lmx) return;
    const long i = idx % lxi, jk = idx / lxi, jj = jk % let, k = jk / let;
    const int row = (int)(AX == 0 ? i : (AX == 1 ? jj : k));
    const long st = AX == 0 ? 1 : (AX == 1 ? (long)lxi : (long)lxi * let);
    const long base = idx - (long)row * st;
    const bool dense = T[0] > 0.5;
    const int E = (int)T[1], W = (int)T[2];
    const double* Bt = T + 3;
    double v = 0.0;
    if (dense) {
        const double* col = Bt + (long)row * N;
        for (int r = 0; r < N; r++) v += col[r] * y[base + r * st];
    } else if (row < E || row >= N - E) {
        const int q = row < E ? row : row - (N - 2 * E);
        const double* col = Bt + (long)q * 2 * W;
        for (int r = 0; r < 2 * W; r++) {
            const int ri = r < W ? r : N - 2 * W + r;
            v += col[r] * y[base + ri * st];
        }
    } else if (op == OP_DERIV_T) {
        auto Y = [&](int d) { return y[base + (row + d) * st]; };
        v = kAa * (Y(-1) - Y(1)) + kBb * (Y(-2) - Y(2));
    } else {
        auto Y = [&](int d) { return y[base + (row + d) * st]; };
        v = a0 * Y(0) + 0.5 * a1 * (Y(1) + Y(-1)) + 0.5 * a2 * (Y(2) + Y(-2)) + 0.5 * a3 * (Y(3) + Y(-3)) +
            0.5 * a4 * (Y(4) + Y(-4)) + 0.5 * a5 * (Y(5) + Y(-5));
    }
    out[idx] = v;
}

int adjoint_workspace(cs_numerics* num, double** w) {
    if (!num->work) CS_HIP(hipMalloc(&num->work, sizeof(double) * num->g.lmx));
    *w = num->work;
    return CS_OK;
}

// out = (operator along axis)^T in, for op OP_DERIV_T / OP_FILTER_T
static int adjoint_apply(cs_numerics* num, int op, int axis, const double* in, double* out, hipStream_t s) {
    if (axis < 0 || axis > 2) return fail(CS_ERR_ARG, "axis must be 0, 1 or 2");
    if (!in || !out) return fail(CS_ERR_ARG, "null field pointer");
    if (in == out) return fail(CS_ERR_ARG, "the transposed operators need out != in");
    const int start = num->ndf[0][axis], end = num->ndf[1][axis];
    if (op == OP_FILTER_T && start == 1 && end == 1)
        return line_apply(num, OP_FILTER, axis, in, out, 1, 1, nullptr, 0, false, s);
    if (op == OP_DERIV_T && (start == 1 || end == 1) && !num->table_set)
        return fail(CS_ERR_STATE, "seam rows need deriv_setup (or an interface table) first");
    LineOpTable* t = nullptr;
    int rc = get_line_op(num, op, axis, start, end, &t);
    if (rc) return rc;
    double* y = nullptr;
    if ((rc = adjoint_workspace(num, &y))) return rc;
    // y = A^-T in
    LineArgs a{};
    a.in = in;
    a.out = y;
    a.tab = t->dev;
    a.pb = num->pb;
    a.N = t->n;
    a.nlines = (int)num->g.face_size(axis);
    a.lxi = num->g.lxi;
    a.plane = (long)num->g.lxi * num->g.let;
    a.st = num->g.stride(axis);
    a.P = t->P;
    a.tab32 = t->dev32;
    if ((rc = launch_line(OP_SOLVE, axis, t->K, a, s))) return rc;
    // out = B^T y
    double an5[6] = {};
    if (op == OP_FILTER_T) filter_coefficients(5, num->alpha_f, an5);
    const double* T = t->dev + 9 * (size_t)t->n + 44;
    const long lmx = num->g.lmx;
    const dim3 grid((unsigned)((lmx + 255) / 256)), block(256);
    const int lxi = num->g.lxi, let = num->g.let;
    if (axis == 0)
        hipLaunchKernelGGL(bt_kernel<0>, grid, block, 0, s, y, out, T, op, t->n, lmx, lxi, let, an5[0], an5[1], an5[2],
                           an5[3], an5[4], an5[5]);
    else if (axis == 1)
        hipLaunchKernelGGL(bt_kernel<1>, grid, block, 0, s, y, out, T, op, t->n, lmx, lxi, let, an5[0], an5[1], an5[2],
                           an5[3], an5[4], an5[5]);
    else
        hipLaunchKernelGGL(bt_kernel<2>, grid, block, 0, s, y, out, T, op, t->n, lmx, lxi, let, an5[0], an5[1], an5[2],
                           an5[3], an5[4], an5[5]);
    CS_HIP(hipGetLastError());
    return CS_OK;
}

int deriv_adjoint(cs_numerics* num, int axis, const double* in, double* out, hipStream_t s) {
    return adjoint_apply(num, OP_DERIV_T, axis, in, out, s);
}

int filter_adjoint(cs_numerics* num, int axis, const double* in, double* out, hipStream_t s) {
    return adjoint_apply(num, OP_FILTER_T, axis, in, out, s);
}

}  // namespace cs

extern "C" {

int cs_deriv_adjoint(cs_numerics* num, int axis, const double* in, double* out, cs_stream stream) {
    if (!num) return cs::fail(CS_ERR_ARG, "null context");
    return cs::deriv_adjoint(num, axis, in, out, (hipStream_t)stream);
}

int cs_filter_adjoint(cs_numerics* num, int axis, const double* in, double* out, cs_stream stream) {
    if (!num) return cs::fail(CS_ERR_ARG, "null context");
    return cs::filter_adjoint(num, axis, in, out, (hipStream_t)stream);
}

}  // extern "C"
