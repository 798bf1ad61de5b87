// line_solver.hip -- compact derivative and compact filter along one axis.
//
// Math: the tridiagonal compact first derivative of the reference
// (build_system, src/cuda/kernels/numerics_device_pc.hpp:176-293: closure
// rows for nstart/nend = 0, quasi-disjoint seam rows for 1) and a 10th-order
// compact filter (absent upstream).  Solve: not the reference's PCR
// (numerics_device_pc.hpp:121-174, power-of-two lines only).  The system
// matrix is constant per line length and end condition, so its LU
// factorisation is computed once on the host and the per-line solve is two
// first-order linear recurrences (forward with nu/mu, backward with rho).
// A line is cut into P chunks of K points, one thread per chunk: each thread
// runs both recurrences on its chunk with zero carry-in, and the carries are
// propagated across chunks through LDS using the precomputed chunk products
// (Mf, Rb).  Any line length 8..1024 works (x lines: as long as the LDS
// tile fits).
//
// Layouts (one workgroup = TL lines x P chunks, threads line-fastest):
//   y/z lines (AX = 1, 2): TL = 64 up to 16 chunks, 16 for 17..64 chunks,
//     32 for 32-point chunks; lanes of a wave are consecutive lines, so every
//     load and store of a wave touches consecutive doubles of one row; with
//     TL = 64 the chunk index is wave-uniform and the tables are read with
//     scalar loads;
//   x lines   (AX = 0): TL = 16, the 16 lines (contiguous in memory) are
//     staged through LDS with coalesced loads and stores, transposed so
//     that a thread reads its chunk conflict-free.
#include "cs_internal.hpp"

#include <algorithm>
#include <array>
#include <cmath>
#include <map>
#include <mutex>

namespace cs {

constexpr int kXLines = 16;   // lines per workgroup for x lines
constexpr int kMaxP   = 64;   // chunks per line (8- and 16-point chunks)
constexpr int kWideP  = 16;   // chunks per line of 32-point chunks, and of 64-line y/z workgroups

// lines per workgroup: 16 for x (LDS tile); for y/z lines 64 (one chunk per
// wave, 512-byte rows) up to kWideP chunks and 16 beyond (up to kMaxP: 1024
// threads), 32 with 32-point chunks so that the register file holds them
// (512 threads)
__host__ __device__ constexpr int lines_per_block(int K, int AX, int P) {
    return AX == 0 ? kXLines : (K == 32 ? 32 : (P <= kWideP ? 64 : 16));
}
__host__ __device__ constexpr int max_threads(int K) { return K == 32 ? 512 : 1024; }
int max_chunks(int K) { return K == 32 ? kWideP : kMaxP; }

// ---------------------------------------------------------------------------
// host: tables
// ---------------------------------------------------------------------------
int choose_chunk(int n, int* K, int* P) {
    // the shortest chunk that keeps 64-line y/z workgroups (n <= 256), else
    // 16-point chunks in 16-line workgroups (the 32-point kernels hold one
    // 5-wave workgroup per CU)
    for (int k : {8, 16}) {
        const int p = (n + k - 1) / k;
        if (p <= kWideP) { *K = k; *P = p; return CS_OK; }
    }
    const int p = (n + 15) / 16;
    if (p <= kMaxP) { *K = 16; *P = p; return CS_OK; }
    return fail(CS_ERR_SHAPE, "line length " + std::to_string(n) + " exceeds 1024");
}

// Order-2M centred filter coefficients: sum a_n = 1+2af, sum_{n>=1} a_n n^{2m}
// = 2af (m = 1..M-1), sum (-1)^n a_n = 0; solved by Gaussian elimination.
void filter_coefficients(int M, double af, double* an) {
    const int n = M + 1;
    double A[6][6] = {}, r[6] = {};
    for (int j = 0; j <= M; j++) A[0][j] = 1.0;
    r[0] = 1.0 + 2.0 * af;
    for (int m = 1; m < M; m++) {
        for (int j = 1; j <= M; j++) A[m][j] = std::pow((double)j, 2.0 * m);
        r[m] = 2.0 * af;
    }
    for (int j = 0; j <= M; j++) A[M][j] = (j % 2 == 0) ? 1.0 : -1.0;
    for (int k = 0; k < n; k++) {
        int p = k;
        for (int i = k + 1; i < n; i++) if (std::fabs(A[i][k]) > std::fabs(A[p][k])) p = i;
        if (p != k) { for (int j = 0; j < n; j++) std::swap(A[k][j], A[p][j]); std::swap(r[k], r[p]); }
        for (int i = k + 1; i < n; i++) {
            double f = A[i][k] / A[k][k];
            for (int j = k; j < n; j++) A[i][j] -= f * A[k][j];
            r[i] -= f * r[k];
        }
    }
    for (int i = n - 1; i >= 0; i--) {
        double s = r[i];
        for (int j = i + 1; j < n; j++) s -= A[i][j] * an[j];
        an[i] = s / A[i][i];
    }
}

// Closure lines of the filter: the row at distance d = 1..4 from the nearer
// end is the order-2d centred filter with alpha 0.2, 0.3, 0.4, 0.45 (never
// above alpha_f), so that the rows next to the one-sided derivative closures
// damp grid-scale modes more than the interior's; the end point itself is not
// filtered.  Inside (d >= 5) the 10th-order filter with alpha_f.
double filter_row_alpha(int dist, double af) {
    static constexpr double kEdge[5] = {0.0, 0.2, 0.3, 0.4, 0.45};
    return dist >= 5 ? af : std::min(af, kEdge[dist < 1 ? 1 : dist]);
}

static void thomas_lu(int n, const std::vector<double>& a, const std::vector<double>& b,
                      const std::vector<double>& c, double* nu, double* mu, double* rho) {
    double cp_prev = 0.0;
    for (int i = 0; i < n; i++) {
        double w = 1.0 / (b[i] - (i > 0 ? a[i] * cp_prev : 0.0));
        nu[i] = w;
        mu[i] = (i > 0) ? -a[i] * w : 0.0;
        double cp = c[i] * w;
        rho[i] = -cp;
        cp_prev = cp;
    }
    rho[n - 1] = 0.0;
}

int build_line_table(int op, int n, int start, int end, double af, int K, std::vector<double>& h) {
    h.assign(6 * (size_t)n + 44 + 3 * (size_t)n, 0.0);
    std::vector<double> a(n), b(n, 1.0), c(n);
    bool cyclic = false;
    if (op == OP_DERIV) {
        for (int i = 0; i < n; i++) { a[i] = kAlpha; c[i] = kAlpha; }
        if (start == 0) { a[0] = 0; c[0] = kAlpha01; a[1] = kAlpha10; c[1] = kAlpha10; }
        else            { a[0] = 0; c[0] = kAlpha; }
        if (end == 0)   { a[n - 1] = kAlpha01; c[n - 1] = 0; a[n - 2] = kAlpha10; c[n - 2] = kAlpha10; }
        else            { a[n - 1] = kAlpha; c[n - 1] = 0; }
    } else {
        cyclic = (start == 1 && end == 1);
        if (!cyclic && (start == 1 || end == 1))
            return fail(CS_ERR_UNSUPPORTED, "filter: a seam on one face only has no single-GPU wrap");
        for (int i = 0; i < n; i++) a[i] = c[i] = cyclic ? af : filter_row_alpha(std::min(i, n - 1 - i), af);
        if (cyclic) {
            // Sherman-Morrison: A = A' + u v^T, gamma = -b_0.
            double gamma = -1.0;
            b[0] = 1.0 - gamma;
            b[n - 1] = 1.0 - af * af / gamma;
            a[0] = 0.0; c[n - 1] = 0.0;
        } else {
            a[0] = 0; c[0] = 0; a[n - 1] = 0; c[n - 1] = 0;   // identity end rows
        }
        // coefficient set M serves the rows at distance M (M = 5: every row inside)
        for (int M = 1; M <= 5; M++) filter_coefficients(M, filter_row_alpha(M, af), &h[6 * (size_t)n + 8 + 6 * M]);
    }
    double* nu = &h[0];
    double* mu = &h[n];
    double* rho = &h[2 * (size_t)n];
    double* Mf = &h[3 * (size_t)n];
    double* Rb = &h[4 * (size_t)n];
    double* z = &h[5 * (size_t)n];
    // the system matrix itself (rows a x[j-1] + b x[j] + c x[j+1]; for the
    // cyclic filter a[0] couples to x[n-1] and c[n-1] to x[0]), for the
    // residual of the mixed-precision solve
    double* A = &h[6 * (size_t)n + 44];
    for (int i = 0; i < n; i++) {
        A[i] = a[i];
        A[n + i] = b[i];
        A[2 * n + i] = c[i];
    }
    if (cyclic) {
        A[0] = af;
        A[n] = 1.0;
        A[2 * n - 1] = 1.0;
        A[3 * n - 1] = af;
    }
    thomas_lu(n, a, b, c, nu, mu, rho);
    for (int s = 0; s < n; s += K) {
        int e = std::min(s + K, n) - 1;
        double p = 1.0;
        for (int j = s; j <= e; j++) { p *= mu[j]; Mf[j] = p; }
        p = 1.0;
        for (int j = e; j >= s; j--) { p *= rho[j]; Rb[j] = p; }
    }
    if (cyclic) {
        double gamma = -1.0;
        std::vector<double> u(n, 0.0), g(n);
        u[0] = gamma; u[n - 1] = af;
        for (int i = 0; i < n; i++) g[i] = nu[i] * u[i] + (i > 0 ? mu[i] * g[i - 1] : 0.0);
        for (int i = n - 1; i >= 0; i--) z[i] = g[i] + (i < n - 1 ? rho[i] * z[i + 1] : 0.0);
        double kappa = af / gamma;
        h[6 * (size_t)n + 0] = 1.0 / (1.0 + z[0] + kappa * z[n - 1]);
        h[6 * (size_t)n + 1] = kappa;
    }
    return CS_OK;
}

// ---------------------------------------------------------------------------
// device
// ---------------------------------------------------------------------------
// Chunked solve of one line system in arithmetic T with tables in T (layout
// of build_line_table): forward recurrence with zero carry-in, carries across
// chunks through LDS (gend/xbeg, [P][TL] entries of T), backward
// recurrence, carries, and the Sherman-Morrison correction of a cyclic
// system.  Two workgroup barriers.
template <int K, typename T>
__device__ __forceinline__ void chunk_solve(T (&r)[K], const T* tab, int N, int P, int s, int c, bool cvalid,
                                            bool cyclic, int TL, int lloc, T* gend, T* xbeg) {
    const T* nu = tab;
    const T* mu = tab + N;
    const T* rho = tab + 2 * N;
    const T* Mf = tab + 3 * N;
    const T* Rb = tab + 4 * N;
    T prev = T(0);
#pragma unroll
    for (int t = 0; t < K; t++) {
        const int j = s + t;
        if (j < N) { prev = nu[j] * r[t] + mu[j] * prev; r[t] = prev; }
    }
    if (cvalid) gend[c * TL + lloc] = prev;
    __syncthreads();
    T G = T(0);
    for (int cc = 0; cc < c; cc++) G = gend[cc * TL + lloc] + Mf[cc * K + K - 1] * G;
    T yN1 = T(0);
    if (cyclic) {
        T Ga = T(0);
        for (int cc = 0; cc < P; cc++) Ga = gend[cc * TL + lloc] + Mf[min(cc * K + K, N) - 1] * Ga;
        yN1 = Ga;
    }
#pragma unroll
    for (int t = 0; t < K; t++) {
        const int j = s + t;
        if (j < N) r[t] += Mf[j] * G;
    }
    T nxt = T(0);
#pragma unroll
    for (int t = K - 1; t >= 0; t--) {
        const int j = s + t;
        if (j < N) { nxt = r[t] + rho[j] * nxt; r[t] = nxt; }
    }
    if (cvalid) xbeg[c * TL + lloc] = nxt;
    __syncthreads();
    T X = T(0);
    for (int cc = P - 1; cc > c; cc--) X = xbeg[cc * TL + lloc] + Rb[cc * K] * X;
#pragma unroll
    for (int t = 0; t < K; t++) {
        const int j = s + t;
        if (j < N) r[t] += Rb[j] * X;
    }
    if (cyclic) {
        T X0 = T(0);
        for (int cc = P - 1; cc >= 0; cc--) X0 = xbeg[cc * TL + lloc] + Rb[cc * K] * X0;
        const T* z = tab + 5 * N;
        const T beta = (X0 + tab[6 * N + 1] * yN1) * tab[6 * N];
#pragma unroll
        for (int t = 0; t < K; t++) {
            const int j = s + t;
            if (j < N) r[t] -= beta * z[j];
        }
    }
}

// One workgroup's share of a line operator: lines blk*TL .. blk*TL+TL-1.
// Threads beyond TL*P of the workgroup are inert (their chunk index is past
// the line) but take part in the barriers.  The per-line arithmetic does not
// depend on TL, so any TL gives the same results bit for bit.
template <int K, int OP, int AX, int TL, bool MIX = false>
__device__ __forceinline__ void line_body(const LineArgs& a, long blk, double* smem) {
    constexpr int H = (OP == OP_DERIV) ? 2 : (OP == OP_FILTER ? 5 : 0);   // stencil half-width
    constexpr int W = K + 2 * H;
    constexpr int TP = TL + 1;                      // padded LDS row (x tile)

    double* gend = smem;                            // [P][TL]
    double* xbeg = smem + a.P * TL;                 // [P][TL]
    double* tile = smem + 2 * a.P * TL;             // x only: [N][TP]

    const int tid = threadIdx.x;
    const int lloc = tid % TL;
    int c = tid / TL;
    if (AX != 0 && TL == 64) c = __builtin_amdgcn_readfirstlane(c);   // one chunk per wave
    const int N = a.N;
    const int P = a.P;
    const long line0 = blk * TL;
    const long line = line0 + lloc;
    const bool active = line < a.nlines;
    const int s = c * K;
    const bool cvalid = s < N;
    const double* tab = a.tab;

    long base = 0;
    if (AX != 0) base = line_base(AX, active ? line : 0, a.lxi, a.plane);

    // ---- stage x lines in LDS (coalesced, transposed) ----
    if (AX == 0) {
        const int nl = (int)min((long)TL, (long)a.nlines - line0);
        const double* src = a.in + line0 * N;
        for (int i = tid; i < N; i += blockDim.x) {
            double v[TL];
#pragma unroll
            for (int l = 0; l < TL; l++) v[l] = (l < nl) ? src[(long)l * N + i] : 0.0;
#pragma unroll
            for (int l = 0; l < TL; l++) tile[i * TP + l] = v[l];
        }
        __syncthreads();
    }

    auto fetch = [&](int row) -> double {
        if (AX == 0) return tile[row * TP + lloc];
        return active ? a.in[base + (long)row * a.st] : 0.0;
    };

    // ---- right-hand side of the chunk rows ----
    double r[K];
    double f0[MIX && OP == OP_FILTER ? K : 1];      // the filtered values (mixed filter: x = f + d)
    if (cvalid) {
        double w[W];
#pragma unroll
        for (int t = 0; t < W; t++) {
            int row = s - H + t;
            if (OP == OP_FILTER && a.cyclic) {
                // wrap without a division: the rows an output of this chunk
                // reads lie in [-H, N + H); rows past those (for t beyond the
                // chunk's last row) are clamped and not used
                row = row < 0 ? row + N : (row >= N ? row - N : row);
                row = min(row, N - 1);
            }
            w[t] = (row >= 0 && row < N) ? fetch(row) : 0.0;
        }
        if constexpr (OP == OP_SOLVE) {
#pragma unroll
            for (int t = 0; t < K; t++) r[t] = w[t];
        } else if constexpr (OP == OP_DERIV) {
#pragma unroll
            for (int t = 0; t < K; t++)
                r[t] = kAa * (w[t + H + 1] - w[t + H - 1]) + kBb * (w[t + H + 2] - w[t + H - 2]);
            const double* pbco = a.pb;
            const double* pbci = a.pb + 16;
            if (c == 0) {
                if (a.seam0) {
                    double own[kLmd], oth[kLmd];
#pragma unroll
                    for (int i = 0; i < kLmd; i++) { own[i] = fetch(i); oth[i] = fetch(N - 1 - i); }
                    double rv0, rv1;
                    if (a.recv) {
                        rv0 = active ? a.recv[line] : 0.0;
                        rv1 = active ? a.recv[a.nlines + line] : 0.0;
                    } else {
                        rv0 = 0.0; rv1 = 0.0;
#pragma unroll
                        for (int i = 0; i < kLmd; i++) { rv0 += pbco[i] * oth[i]; rv1 += pbco[kLmd + i] * oth[i]; }
                    }
                    double s0 = 0.0, s1 = 0.0;
#pragma unroll
                    for (int i = 0; i < kLmd; i++) { s0 += pbci[i] * own[i]; s1 += pbci[kLmd + i] * own[i]; }
                    r[0] = s0 + rv0;
                    r[1] = s1 + rv1;
                } else {
                    r[0] = kAb00 * w[H] + kAb01 * w[H + 1] + kAb02 * w[H + 2];
                    r[1] = kAb10 * (w[H + 2] - w[H]);
                }
            }
            if (s + K > N - 2) {            // the chunk holds row N-2 and/or N-1
                const int tl = N - 1 - s;   // chunk-local index of row N-1
                double r1, r2;
                if (a.seam1) {
                    double own[kLmd], oth[kLmd];
#pragma unroll
                    for (int i = 0; i < kLmd; i++) { own[i] = fetch(N - 1 - i); oth[i] = fetch(i); }
                    double rv2, rv3;
                    if (a.recv) {
                        rv2 = active ? a.recv[2L * a.nlines + line] : 0.0;
                        rv3 = active ? a.recv[3L * a.nlines + line] : 0.0;
                    } else {
                        rv2 = 0.0; rv3 = 0.0;
#pragma unroll
                        for (int i = 0; i < kLmd; i++) { rv2 += pbco[i] * oth[i]; rv3 += pbco[kLmd + i] * oth[i]; }
                    }
                    double s0 = 0.0, s1 = 0.0;
#pragma unroll
                    for (int i = 0; i < kLmd; i++) { s0 += pbci[i] * own[i]; s1 += pbci[kLmd + i] * own[i]; }
                    r1 = -(s0 + rv2);
                    r2 = -(s1 + rv3);
                } else {
                    double fN1 = fetch(N - 1), fN2 = fetch(N - 2), fN3 = fetch(N - 3);
                    r1 = -(kAb00 * fN1 + kAb01 * fN2 + kAb02 * fN3);
                    r2 = -kAb10 * (fN3 - fN1);
                }
#pragma unroll
                for (int t = 0; t < K; t++) {
                    if (t == tl) r[t] = r1;
                    if (t == tl - 1) r[t] = r2;
                }
            }
        } else {
            const double* an = tab + 6 * N + 8;
            if (a.cyclic) {
#pragma unroll
                for (int t = 0; t < K; t++) {
                    double v = an[30] * w[t + H];
#pragma unroll
                    for (int m = 1; m <= 5; m++) v += 0.5 * an[30 + m] * (w[t + H + m] + w[t + H - m]);
                    r[t] = v;
                }
            } else {
#pragma unroll
                for (int t = 0; t < K; t++) {
                    const int j = s + t;
                    const int dist = min(j, N - 1 - j);
                    if (dist >= 5) {
                        double v = an[30] * w[t + H];
#pragma unroll
                        for (int m = 1; m <= 5; m++) v += 0.5 * an[30 + m] * (w[t + H + m] + w[t + H - m]);
                        r[t] = v;
                    } else if (dist <= 0) {
                        r[t] = w[t + H];
                    } else {
                        // unrolled with a bound test so that w stays in registers
                        const double* cf = an + 6 * dist;
                        double v = cf[0] * w[t + H];
#pragma unroll
                        for (int m = 1; m < H; m++)
                            if (m <= dist) v += 0.5 * cf[m] * (w[t + H + m] + w[t + H - m]);
                        r[t] = v;
                    }
                }
            }
        }
        if constexpr (MIX && OP == OP_FILTER) {
            // the mixed filter solves for the correction d = x - f, A d = B f - A f:
            // its FP32 rounding is relative to the (small) correction, not to f
            const double* Am = tab + 6 * N + 44;
            const double* Bm = Am + N;
            const double* Cm = Am + 2 * N;
#pragma unroll
            for (int t = 0; t < K; t++) {
                const int j = s + t;
                f0[t] = w[t + H];
                if (j < N) r[t] -= Am[j] * w[t + H - 1] + Bm[j] * w[t + H] + Cm[j] * w[t + H + 1];
            }
        }
    } else {
#pragma unroll
        for (int t = 0; t < K; t++) r[t] = 0.0;
        if constexpr (MIX && OP == OP_FILTER) {
#pragma unroll
            for (int t = 0; t < K; t++) f0[t] = 0.0;
        }
    }

    const bool cyc = OP == OP_FILTER && a.cyclic;
    if constexpr (!MIX) {
        chunk_solve<K, double>(r, tab, N, P, s, c, cvalid, cyc, TL, lloc, gend, xbeg);
    } else {
        // FP32 solve, then `mixed` refinement sweeps: FP64 residual of the
        // FP64 system, FP32 solve of it, FP64 update
        float x32[K];
#pragma unroll
        for (int t = 0; t < K; t++) x32[t] = (float)r[t];
        chunk_solve<K, float>(x32, a.tab32, N, P, s, c, cvalid, cyc, TL, lloc, (float*)gend, (float*)xbeg);
        double x[K];
#pragma unroll
        for (int t = 0; t < K; t++) x[t] = (double)x32[t];
        double* first = smem + 2 * P * TL + (AX == 0 ? (size_t)N * TP : 0);   // [P][TL]
        double* last = first + P * TL;
        const int L = min(K, N - s);
        const double* Am = tab + 6 * N + 44;
        const double* Bm = Am + N;
        const double* Cm = Am + 2 * N;
        const int cl = c > 0 ? c - 1 : P - 1, cr = c + 1 < P ? c + 1 : 0;    // wraps: cyclic corners
        for (int sweep = 0; sweep < a.mixed; sweep++) {
            __syncthreads();          // the previous sweep's reads of first/last are done
            if (cvalid) {
                double xl = 0.0;
#pragma unroll
                for (int t = 0; t < K; t++)
                    if (t == L - 1) xl = x[t];
                first[c * TL + lloc] = x[0];
                last[c * TL + lloc] = xl;
            }
            __syncthreads();
            const double xprev = cvalid ? last[cl * TL + lloc] : 0.0;
            const double xnext = cvalid ? first[cr * TL + lloc] : 0.0;
            float dx[K];
#pragma unroll
            for (int t = 0; t < K; t++) {
                const int j = s + t;
                double res = 0.0;
                if (j < N) {
                    const double xm = t > 0 ? x[t > 0 ? t - 1 : 0] : xprev;
                    double xp = xnext;
#pragma unroll
                    for (int u = 0; u < K; u++)
                        if (u == t + 1 && u < L) xp = x[u];
                    res = r[t] - (Am[j] * xm + Bm[j] * x[t] + Cm[j] * xp);
                }
                dx[t] = (float)res;
            }
            chunk_solve<K, float>(dx, a.tab32, N, P, s, c, cvalid, cyc, TL, lloc, (float*)gend, (float*)xbeg);
#pragma unroll
            for (int t = 0; t < K; t++) x[t] += (double)dx[t];
        }
        if constexpr (OP == OP_FILTER) {
#pragma unroll
            for (int t = 0; t < K; t++) r[t] = f0[t] + x[t];
        } else {
#pragma unroll
            for (int t = 0; t < K; t++) r[t] = x[t];
        }
    }

    // ---- boundary values for the characteristic BCs ----
    if (a.drva && active && cvalid) {
        const long fs = a.nlines;
        if (c == 0) a.drva[(long)a.var * fs + line] = r[0];
        if (c == P - 1) {
#pragma unroll
            for (int t = 0; t < K; t++)
                if (s + t == N - 1) a.drva[(long)kNV * fs + (long)a.var * fs + line] = r[t];
        }
    }

    // ---- store ----
    if (AX == 0) {
        // every thread finished reading the tile before the second barrier
        if (cvalid) {
#pragma unroll
            for (int t = 0; t < K; t++)
                if (s + t < N) tile[(s + t) * TP + lloc] = r[t];
        }
        __syncthreads();
        const int nl = (int)min((long)TL, (long)a.nlines - line0);
        double* dst = a.out + line0 * N;
        for (int i = tid; i < N; i += blockDim.x) {
#pragma unroll
            for (int l = 0; l < TL; l++)
                if (l < nl) dst[(long)l * N + i] = tile[i * TP + l];
        }
    } else if (active && cvalid) {
#pragma unroll
        for (int t = 0; t < K; t++)
            if (s + t < N) a.out[base + (long)(s + t) * a.st] = r[t];
    }
}

// grid y runs a.nfields fields a.fstride apart with the same operator
template <int K, int OP, int AX, bool MIX, int TL>
__global__ void __launch_bounds__(max_threads(K)) line_kernel(LineArgs a) {
    extern __shared__ double smem[];
    if (blockIdx.y > 0) {
        a.in += (long)blockIdx.y * a.fstride;
        a.out += (long)blockIdx.y * a.fstride;
    }
    line_body<K, OP, AX, TL, MIX>(a, blockIdx.x, smem);
}

template <int K, int OP, int AX, bool MIX, int TL>
static int launch_tl(const LineArgs& a, hipStream_t s) {
    dim3 block(TL * a.P);
    dim3 grid((unsigned)((a.nlines + TL - 1) / TL), (unsigned)(a.nfields > 1 ? a.nfields : 1));
    size_t lds = sizeof(double) * 2 * a.P * TL;
    if (AX == 0) lds += sizeof(double) * (size_t)a.N * (TL + 1);
    if (MIX) lds += sizeof(double) * 2 * a.P * TL;
    if (lds > 64 * 1024)
        CS_HIP(hipFuncSetAttribute((const void*)line_kernel<K, OP, AX, MIX, TL>,
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL((line_kernel<K, OP, AX, MIX, TL>), grid, block, lds, s, a);
    CS_HIP(hipGetLastError());
    return CS_OK;
}

template <int K, int OP, int AX, bool MIX>
static int launch_mix(const LineArgs& a, hipStream_t s) {
    if constexpr (AX != 0 && K != 32) {
        if (lines_per_block(K, AX, a.P) == 16) return launch_tl<K, OP, AX, MIX, 16>(a, s);
    }
    return launch_tl<K, OP, AX, MIX, lines_per_block(K, AX, 1)>(a, s);
}

template <int K, int OP, int AX>
static int launch_one(const LineArgs& a, hipStream_t s) {
    if constexpr (OP != OP_SOLVE) {
        if (a.mixed) return launch_mix<K, OP, AX, true>(a, s);
    }
    return launch_mix<K, OP, AX, false>(a, s);
}

template <int K, int OP>
static int launch_axis(int axis, const LineArgs& a, hipStream_t s) {
    if (axis == 0) return launch_one<K, OP, 0>(a, s);
    if (axis == 1) return launch_one<K, OP, 1>(a, s);
    return launch_one<K, OP, 2>(a, s);
}

int launch_line(int op, int axis, int K, const LineArgs& a, hipStream_t s) {
    if (a.P < 1 || a.P > max_chunks(K)) return fail(CS_ERR_SHAPE, "chunk count out of range");
    if (a.nfields > 1 && (a.drva || a.recv)) return fail(CS_ERR_ARG, "multi-field launches write no boundary data");
    if ((long)a.P * K < a.N) return fail(CS_ERR_SHAPE, "chunks do not cover the line");
    if (axis == 0 && (size_t)a.N * (kXLines + 1) * 8 + 4 * (size_t)a.P * kXLines * 8 > 160 * 1024)
        return fail(CS_ERR_SHAPE, "x line too long for the LDS tile");
    if (a.mixed && !a.tab32) return fail(CS_ERR_STATE, "mixed precision needs the FP32 tables");
    if (op == OP_DERIV) {
        if (K == 8) return launch_axis<8, OP_DERIV>(axis, a, s);
        if (K == 16) return launch_axis<16, OP_DERIV>(axis, a, s);
        if (K == 32) return launch_axis<32, OP_DERIV>(axis, a, s);
    } else if (op == OP_FILTER) {
        if (K == 8) return launch_axis<8, OP_FILTER>(axis, a, s);
        if (K == 16) return launch_axis<16, OP_FILTER>(axis, a, s);
        if (K == 32) return launch_axis<32, OP_FILTER>(axis, a, s);
    } else if (op == OP_SOLVE) {
        if (a.mixed) return fail(CS_ERR_UNSUPPORTED, "transposed solves run in FP64");
        if (K == 8) return launch_axis<8, OP_SOLVE>(axis, a, s);
        if (K == 16) return launch_axis<16, OP_SOLVE>(axis, a, s);
        if (K == 32) return launch_axis<32, OP_SOLVE>(axis, a, s);
    }
    return fail(CS_ERR_SHAPE, "unsupported chunk length");
}

// ---------------------------------------------------------------------------
// many line operators of one chunk length and axis in one persistent launch
// (batched subdomains): item i runs workgroup share items[i].y of operator
// args[items[i].x]; workgroups take items round-robin.  Same body and lines
// per workgroup as line_kernel, so the same results bit for bit.
// ---------------------------------------------------------------------------
template <int K, int OP, int AX, int TL>
__global__ void __launch_bounds__(max_threads(K)) batch_line_kernel(const LineArgs* args, const int2* items,
                                                                    int nitems) {
    extern __shared__ double smem[];
    static_assert(sizeof(LineArgs) % 4 == 0, "LineArgs moves as 32-bit words");
    constexpr int kWords = sizeof(LineArgs) / 4;
    for (int it = blockIdx.x; it < nitems; it += gridDim.x) {
        // the item and its operator are the same for the whole workgroup:
        // move them into scalar registers word by word, as the by-value
        // kernel argument of line_kernel is
        const int op = __builtin_amdgcn_readfirstlane(items[it].x);
        const int blk = __builtin_amdgcn_readfirstlane(items[it].y);
        const unsigned* src = reinterpret_cast<const unsigned*>(args + op);
        unsigned words[kWords];
#pragma unroll
        for (int i = 0; i < kWords; i++) words[i] = __builtin_amdgcn_readfirstlane(src[i]);
        LineArgs a;
        __builtin_memcpy(&a, words, sizeof(LineArgs));
        line_body<K, OP, AX, TL>(a, blk, smem);
        __syncthreads();   // the LDS of this item is free for the next
    }
}

int batch_lines_per_item(int K, int axis, int P) { return lines_per_block(K, axis, P); }

size_t batch_lds_bytes(int K, int axis, int N) {
    const int P = (N + K - 1) / K;
    const int TL = lines_per_block(K, axis, P);
    size_t lds = sizeof(double) * 2 * (size_t)P * TL;
    if (axis == 0) lds += sizeof(double) * (size_t)N * (TL + 1);
    return lds;
}

template <int K, int OP, int TL>
static const void* batch_symbol_tl(int axis) {
    if (axis == 0) return (const void*)batch_line_kernel<K, OP, 0, lines_per_block(K, 0, 1)>;
    if (axis == 1) return (const void*)batch_line_kernel<K, OP, 1, TL>;
    return (const void*)batch_line_kernel<K, OP, 2, TL>;
}

template <int K, int OP>
static const void* batch_symbol(int axis, int P) {
    if (K != 32 && lines_per_block(K, axis, P) == 16) return batch_symbol_tl<K, OP, 16>(axis);
    return batch_symbol_tl<K, OP, lines_per_block(K, 1, 1)>(axis);
}

static const void* batch_symbol(int op, int K, int axis, int P) {
    if (op == OP_DERIV) {
        if (K == 8) return batch_symbol<8, OP_DERIV>(axis, P);
        if (K == 16) return batch_symbol<16, OP_DERIV>(axis, P);
        return batch_symbol<32, OP_DERIV>(axis, P);
    }
    if (K == 8) return batch_symbol<8, OP_FILTER>(axis, P);
    if (K == 16) return batch_symbol<16, OP_FILTER>(axis, P);
    return batch_symbol<32, OP_FILTER>(axis, P);
}

int batch_blocks_per_cu(int op, int K, int axis, int P, int threads, size_t lds, int* nb) {
    if (lds > 160 * 1024) return fail(CS_ERR_SHAPE, "batched x lines too long for the LDS tile");
    const void* k = batch_symbol(op, K, axis, P);
    if (lds > 64 * 1024) CS_HIP(hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    int n = 0;
    CS_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k, threads, lds));
    *nb = n < 1 ? 1 : n;
    return CS_OK;
}

template <int K, int OP, int TL>
static void batch_launch_tl(int axis, dim3 g, dim3 b, size_t lds, hipStream_t s, const LineArgs* args,
                            const int2* items, int n) {
    if (axis == 0)
        hipLaunchKernelGGL((batch_line_kernel<K, OP, 0, lines_per_block(K, 0, 1)>), g, b, lds, s, args, items, n);
    else if (axis == 1) hipLaunchKernelGGL((batch_line_kernel<K, OP, 1, TL>), g, b, lds, s, args, items, n);
    else hipLaunchKernelGGL((batch_line_kernel<K, OP, 2, TL>), g, b, lds, s, args, items, n);
}

template <int K, int OP>
static void batch_launch(int axis, int P, dim3 g, dim3 b, size_t lds, hipStream_t s, const LineArgs* args,
                         const int2* items, int n) {
    if (K != 32 && lines_per_block(K, axis, P) == 16) batch_launch_tl<K, OP, 16>(axis, g, b, lds, s, args, items, n);
    else batch_launch_tl<K, OP, lines_per_block(K, 1, 1)>(axis, g, b, lds, s, args, items, n);
}

int launch_batch_lines(int op, int K, int axis, int P, const LineArgs* args, const int2* items, int nitems,
                       int grid, int threads, size_t lds, hipStream_t s) {
    if (nitems <= 0) return CS_OK;
    if (lds > 160 * 1024) return fail(CS_ERR_SHAPE, "batched x lines too long for the LDS tile");
    if (threads > max_threads(K) || P > max_chunks(K)) return fail(CS_ERR_SHAPE, "too many chunks for the batched launch");
    const dim3 g(grid), b(threads);
    if (op == OP_DERIV) {
        if (K == 8) batch_launch<8, OP_DERIV>(axis, P, g, b, lds, s, args, items, nitems);
        else if (K == 16) batch_launch<16, OP_DERIV>(axis, P, g, b, lds, s, args, items, nitems);
        else batch_launch<32, OP_DERIV>(axis, P, g, b, lds, s, args, items, nitems);
    } else {
        if (K == 8) batch_launch<8, OP_FILTER>(axis, P, g, b, lds, s, args, items, nitems);
        else if (K == 16) batch_launch<16, OP_FILTER>(axis, P, g, b, lds, s, args, items, nitems);
        else batch_launch<32, OP_FILTER>(axis, P, g, b, lds, s, args, items, nitems);
    }
    CS_HIP(hipGetLastError());
    return CS_OK;
}

// ---------------------------------------------------------------------------
// context helpers
// ---------------------------------------------------------------------------
// Device table of one line operator with chunk length K (0 = the tuned or
// default choice).
int make_line_op(const cs_numerics* num, int op, int n, int start, int end, int K, LineOpTable* t) {
    const int minlen = (op == OP_FILTER && start == 1) ? 11 : 8;
    if (n < minlen) return fail(CS_ERR_SHAPE, "line length " + std::to_string(n) + " below the stencil size");
    t->n = n;
    if (K == 0 && (op == OP_DERIV || op == OP_FILTER)) K = tuned_chunk(op, n, start, end, num->precision);
    if (K != 0) {
        if (K != 8 && K != 16 && K != 32) return fail(CS_ERR_ARG, "chunk length must be 8, 16 or 32");
        t->K = K;
        t->P = (n + K - 1) / K;
        if (t->P > max_chunks(K)) return fail(CS_ERR_SHAPE, "chunk length too short for this line");
    } else {
        int rc = choose_chunk(n, &t->K, &t->P);
        if (rc) return rc;
    }
    std::vector<double> h;
    int rc = (op == OP_DERIV_T || op == OP_FILTER_T)
                 ? build_adjoint_table(op, n, start, end, num->alpha_f, num->pb_host, t->K, h)
                 : build_line_table(op, n, start, end, num->alpha_f, t->K, h);
    if (rc) return rc;
    // FP32 copy of the factorisation (nu, mu, rho, Mf, Rb, z, 1/den, kappa)
    // for the mixed-precision solve
    std::vector<float> h32(6 * (size_t)n + 2);
    for (size_t i = 0; i < h32.size(); i++) h32[i] = (float)h[i];
    hipError_t e = hipMalloc(&t->dev, h.size() * sizeof(double));
    if (e == hipSuccess) e = hipMemcpy(t->dev, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMalloc(&t->dev32, h32.size() * sizeof(float));
    if (e == hipSuccess) e = hipMemcpy(t->dev32, h32.data(), h32.size() * sizeof(float), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        free_line_op(*t);      // the caller keeps nothing of a failed table
        return fail(CS_ERR_HIP, std::string("operator table upload failed: ") + hipGetErrorString(e));
    }
    t->count = h.size();
    return CS_OK;
}

int get_line_op(cs_numerics* num, int op, int axis, int start, int end, LineOpTable** out) {
    const int n = num->g.n(axis);
    LineOpKey key{op, n, start, end};
    auto it = num->ops.find(key);
    if (it != num->ops.end()) { *out = &it->second; return CS_OK; }
    LineOpTable t;
    int rc = make_line_op(num, op, n, start, end, 0, &t);
    if (rc) return rc;
    auto res = num->ops.emplace(key, t);
    *out = &res.first->second;
    return CS_OK;
}

int line_apply(cs_numerics* num, int op, int axis, const double* in, double* out, int start, int end,
               const double* recv, int var, bool write_drva, hipStream_t s, int nfields) {
    if (axis < 0 || axis > 2) return fail(CS_ERR_ARG, "axis must be 0, 1 or 2");
    if (!in || !out) return fail(CS_ERR_ARG, "null field pointer");
    if (op == OP_DERIV && (start == 1 || end == 1) && !num->table_set)
        return fail(CS_ERR_STATE, "seam rows need deriv_setup (or an interface table) first");
    LineOpTable* t = nullptr;
    int rc = get_line_op(num, op, axis, start, end, &t);
    if (rc) return rc;
    LineArgs a{};
    a.in = in;
    a.out = out;
    a.tab = t->dev;
    a.pb = num->pb;
    a.recv = recv;
    a.drva = write_drva ? num->drva[axis] : nullptr;
    a.var = var;
    a.N = t->n;
    a.nlines = (int)num->g.face_size(axis);
    a.lxi = num->g.lxi;
    a.plane = (long)num->g.lxi * num->g.let;
    a.st = num->g.stride(axis);
    a.P = t->P;
    a.seam0 = (op == OP_DERIV) ? start : 0;
    a.seam1 = (op == OP_DERIV) ? end : 0;
    a.cyclic = (op == OP_FILTER) ? (start == 1 && end == 1) : 0;
    a.tab32 = t->dev32;
    a.mixed = num->precision;
    a.nfields = nfields;
    a.fstride = num->g.lmx;
    return launch_line(op, axis, t->K, a, s);
}

}  // namespace cs
