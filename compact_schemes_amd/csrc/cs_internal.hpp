// cs_internal.hpp -- shared declarations of the gfx950 compact-scheme library.
//
// Scheme and gas constants restate src/common/parameters.hpp:33-74 of the
// reference; the field layout is the reference's (SoA, variable-major,
// x fastest).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>
#include <map>
#include <array>

#include "../../include/compact_schemes.h"

namespace cs {

// ---- compact first derivative (Lele 6th order, tridiagonal) ----------------
constexpr double kAlpha   = 1.0 / 3.0;
constexpr double kAlpha01 = 2.0;
constexpr double kAlpha10 = 1.0 / 4.0;
constexpr double kAa      = 7.0 / 9.0;
constexpr double kBb      = 1.0 / 36.0;
constexpr double kAb00    = -5.0 / 2.0;
constexpr double kAb01    = 2.0;
constexpr double kAb02    = 1.0 / 2.0;
constexpr double kAb10    = 3.0 / 4.0;
constexpr int    kLmd     = 8;     // points per side of a seam stencil
// depth of the characteristic-BC update: the face row only.  The reference's
// mbci = 4 (src/common/parameters.hpp:46) weights rows 0..3 by the point array
// sbcc; a point weight cannot tell faces apart, so along an edge line the
// deeper rows of one face land on the boundary points of the adjacent face
// and the scheme blows up (1e-8 noise -> NaN in 25 RK4 steps at 32^3).  With
// sbcc = 0 inward those edge writes are all the deeper rows ever did (D24).
constexpr int    kMbci    = 1;
constexpr int    kNV      = 5;     // conserved variables

// ---- gas ---------------------------------------------------------------------
constexpr double kGam          = 1.4;
constexpr double kGamm1        = kGam - 1.0;
constexpr double kHam          = 1.0 / kGam;
constexpr double kHamm1        = 1.0 / kGamm1;
constexpr double kHamHamm1     = kHam * kHamm1;
constexpr double kPrndtl       = 0.71;
constexpr double kGamm1Prndtli = 1.0 / (kGamm1 * kPrndtl);

// ---- error handling ------------------------------------------------------------
void set_error(const std::string& msg);
int fail(int code, const std::string& msg);
#define CS_HIP(expr)                                                                 \
    do {                                                                             \
        hipError_t e_ = (expr);                                                      \
        if (e_ != hipSuccess)                                                        \
            return ::cs::fail(CS_ERR_HIP, std::string(#expr " failed: ") +           \
                                              hipGetErrorString(e_));                \
    } while (0)

// ---- geometry -----------------------------------------------------------------
struct Geo {
    int lxi = 0, let = 0, lze = 0;
    long lmx = 0;
    int n(int a) const { return a == 0 ? lxi : (a == 1 ? let : lze); }
    long face_size(int a) const {
        return a == 0 ? (long)let * lze : (a == 1 ? (long)lxi * lze : (long)lxi * let);
    }
    long stride(int a) const { return a == 0 ? 1 : (a == 1 ? (long)lxi : (long)lxi * let); }
};

__host__ __device__ inline long line_base(int axis, long f, int lxi, long plane) {
    if (axis == 0) return f * lxi;
    if (axis == 1) return (f / lxi) * plane + (f % lxi);
    return f;
}

// ---- line operators ------------------------------------------------------------
// OP_SOLVE: the solve alone (right-hand side = the input) with the table of
// a transposed operator, keyed as OP_DERIV_T / OP_FILTER_T
enum LineOp { OP_DERIV = 0, OP_FILTER = 1, OP_SOLVE = 2, OP_DERIV_T = 3, OP_FILTER_T = 4 };

// A line operator is the constant tridiagonal system of one line length and
// one pair of end conditions, stored as the coefficients of two first-order
// linear recurrences (LU of the system) plus the chunk-carry products the
// chunked solve needs.  Layout of the device table (doubles):
//   [0N..1N) nu   forward  g_j = nu_j r_j + mu_j g_{j-1}
//   [1N..2N) mu
//   [2N..3N) rho  backward x_j = g_j + rho_j x_{j+1}
//   [3N..4N) Mf   prod mu_t, t = chunk start .. j
//   [4N..5N) Rb   prod rho_t, t = j .. chunk end
//   [5N..6N) z    cyclic correction vector (zeros otherwise)
//   [6N+0]   inv_den, [6N+1] kappa   (cyclic: x = y - (y_0 + kappa y_{N-1}) inv_den z)
//   [6N+8 + 6*M + m]  filter coefficient a_m of the order-2M stencil (M = 1..5)
struct LineOpKey {
    int op, n, start, end;   // start/end: deriv 0 closure / 1 seam; filter 0 closure / 1 cyclic
    bool operator<(const LineOpKey& o) const {
        if (op != o.op) return op < o.op;
        if (n != o.n) return n < o.n;
        if (start != o.start) return start < o.start;
        return end < o.end;
    }
};

struct LineOpTable {
    int n = 0, K = 0, P = 0;
    double* dev = nullptr;     // FP64 table (build_line_table)
    float* dev32 = nullptr;    // FP32 factorisation for the mixed-precision solve
    size_t count = 0;
};

inline void free_line_op(LineOpTable& t) {
    if (t.dev) (void)hipFree(t.dev);
    if (t.dev32) (void)hipFree(t.dev32);
    t.dev = nullptr;
    t.dev32 = nullptr;
}

int choose_chunk(int n, int* K, int* P);
// tables of the transposed operators (adjoint.hip): LU of A^T in the layout
// of build_line_table, then the transposed explicit stencil B^T
int build_adjoint_table(int op, int n, int start, int end, double alpha_f, const double* pb, int K,
                        std::vector<double>& h);
int build_line_table(int op, int n, int start, int end, double alpha_f, int K,
                     std::vector<double>& host);
void filter_coefficients(int M, double af, double* an);
// alpha of the closure filter's row at distance dist >= 1 from the nearer end
double filter_row_alpha(int dist, double af);

struct LineArgs {
    const double* in;
    double* out;
    const double* tab;
    const double* pb;     // pbco[16] followed by pbci[16]
    const double* recv;   // seam sums of this variable, [2 face][2 row][face]; null = wrap inline
    double* drva;         // [2 face][5 var][face] or null
    int var;
    int N;
    int nlines;
    int lxi;
    long plane;
    long st;
    int P;
    int seam0, seam1;     // deriv: seam rows at the start / end of the line
    int cyclic;           // filter: cyclic system
    const float* tab32;   // FP32 factorisation (mixed precision)
    int mixed;            // k > 0: FP32 solve + k FP64-residual refinement sweeps
    int nfields;          // fields processed by one launch (grid y), 0 or 1 = one
    long fstride;         // distance between those fields
};

int launch_line(int op, int axis, int K, const LineArgs& a, hipStream_t s);

// most chunks per line of a K-point chunked line operator
int max_chunks(int K);

// batched line operators (batch.hip drives them, line_solver.hip launches);
// one launch holds operators of one chunk length K, axis and chunk count P
int batch_lines_per_item(int K, int axis, int P);
size_t batch_lds_bytes(int K, int axis, int N);
int batch_blocks_per_cu(int op, int K, int axis, int P, int threads, size_t lds, int* nb);
int launch_batch_lines(int op, int K, int axis, int P, const LineArgs* args, const int2* items, int nitems,
                       int grid, int threads, size_t lds, hipStream_t s);

// launch-shape choices (tune.hip): chunk length per (operator, line length,
// start/end rows, precision), 0 = none (the default rule of choose_chunk);
// fused right-hand-side sweep shapes per (axis, line length, seam axis,
// precision of the intermediates, constant 0 / per-point 1 metrics), fused
// filter shapes per (axis, line length, seam axis)
int tuned_chunk(int op, int n, int start, int end, int prec);
void set_tuned_chunk(int op, int n, int start, int end, int prec, int K);
bool tuned_sweep(int axis, int n, int seam, int prec, int metrics, int* variant_lines_xcd);
void set_tuned_sweep(int axis, int n, int seam, int prec, int metrics, const int* variant_lines_xcd);
bool tuned_filter(int axis, int n, int seam, int* fields_lines_chunk);
void set_tuned_filter(int axis, int n, int seam, const int* fields_lines_chunk);

// ---- contexts -----------------------------------------------------------------
}  // namespace cs

struct cs_numerics {
    cs::Geo g;
    int nbc[2][3];
    int ndf[2][3];
    int device = 0;
    double* drva[3] = {nullptr, nullptr, nullptr};
    double* send[3] = {nullptr, nullptr, nullptr};
    double* recv[3] = {nullptr, nullptr, nullptr};
    double* pb = nullptr;        // pbco[16] | pbci[16] on the device
    double pb_host[32];
    bool table_set = false;
    double alpha_f = 0.49;
    int precision = 0;           // 0: FP64 line solves; k > 0: FP32 solves + k FP64-residual refinement sweeps
    double* work = nullptr;      // [lmx] scratch of the adjoint line operators (allocated on first use)
    std::map<cs::LineOpKey, cs::LineOpTable> ops;
};

struct cs_physics {
    cs::Geo g;
    int viscous = 1;
    cs_point umf{0, 0, 0};
    cs_point dudtmf{0, 0, 0};
    double* buffer = nullptr;        // [5][3][lmx]
    double* buffer_deriv = nullptr;  // [5][3][lmx]
    double* ss = nullptr;            // [3][lmx]
    double* stress = nullptr;        // [6][lmx]
    double* heat = nullptr;          // [3][lmx]
    double* red = nullptr;           // time-step partial maxima
    double* red_host = nullptr;      // pinned
    int red_blocks = 0;
};

namespace cs {
int get_line_op(cs_numerics* num, int op, int axis, int start, int end, LineOpTable** t);
int make_line_op(const cs_numerics* num, int op, int n, int start, int end, int K, LineOpTable* t);
int adjoint_workspace(cs_numerics* num, double** w);
int deriv_adjoint(cs_numerics* num, int axis, const double* in, double* out, hipStream_t s);
int filter_adjoint(cs_numerics* num, int axis, const double* in, double* out, hipStream_t s);
int line_apply(cs_numerics* num, int op, int axis, const double* in, double* out, int start, int end,
               const double* recv, int var, bool write_drva, hipStream_t s, int nfields = 1);
int check_dims(const cs_dims* d, Geo* g);
}  // namespace cs
