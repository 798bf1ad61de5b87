// solver.hip -- cs_solver: the fused right-hand side and the RK4 step.
//
// The RK4 step is the CANARD low-storage scheme the reference driver loop
// sketches (driver/canard.cu:132-205, weights dtk = dt/(nkrk-nk+1) for
// 1-based nk): q_k = q0 - dtk*yaco*de(q_{k-1}).  Stage buffers ping-pong
// between the caller's q and two scratch fields so that q0 is never copied:
// A -> B, (B, A) -> C, (C, A) -> B, (B, A) -> A.
#include "fused.hpp"

#include <algorithm>
#include <cstring>

namespace cs {
namespace fused {

// Derivative operator of one axis with balanced chunks (the table layout of
// the line solver; closure rows for 0, seam rows for 1).
int build_axis_table(int n, int seam, int K, int P, std::vector<double>& h) {
    (void)K;
    h.assign(5 * (size_t)n, 0.0);
    std::vector<double> a(n, kAlpha), b(n, 1.0), c(n, kAlpha);
    if (!seam) { a[0] = 0; c[0] = kAlpha01; a[1] = kAlpha10; c[1] = kAlpha10; }
    else       { a[0] = 0; c[0] = kAlpha; }
    if (!seam) { a[n - 1] = kAlpha01; c[n - 1] = 0; a[n - 2] = kAlpha10; c[n - 2] = kAlpha10; }
    else       { a[n - 1] = kAlpha; c[n - 1] = 0; }
    double* nu = &h[0];
    double* mu = &h[n];
    double* rho = &h[2 * (size_t)n];
    double* Mf = &h[3 * (size_t)n];
    double* Rb = &h[4 * (size_t)n];
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
    for (int ch = 0; ch < P; ch++) {
        int s = (int)(((long)ch * n) / P), e = (int)(((long)(ch + 1) * n) / P) - 1;
        double p = 1.0;
        for (int j = s; j <= e; j++) { p *= mu[j]; Mf[j] = p; }
        p = 1.0;
        for (int j = e; j >= s; j--) { p *= rho[j]; Rb[j] = p; }
    }
    return CS_OK;
}

// the filter that closes an RK4 step (cfg.filter): one five-variable sweep
// per axis when every axis splits into 8-point chunks, else the line solver
// with the five variables in one launch per axis
int step_filter(cs_solver* sv, double* qa, hipStream_t s) {
    if (!sv->cfg.filter) return CS_OK;
    for (int a = 0; a < 3; a++) {
        int rc;
        if (filter_axis_fused(sv, a)) {
            rc = filter_sweeps(sv, qa, s, a);
        } else {
            const int sp = sv->ax[a].seam;
            rc = line_apply(sv->num, OP_FILTER, a, qa, qa, sp, sp, nullptr, 0, false, s, 5);
        }
        if (rc) return rc;
    }
    return CS_OK;
}

}  // namespace fused
}  // namespace cs

using namespace cs;
using namespace cs::fused;

static int axis_mode(const int* nbc, int axis, int* seam, int* gcbc) {
    int f0 = nbc[axis], f1 = nbc[3 + axis];
    auto is_seam = [](int v) {
        return v == CS_BC_INTER_STRAIGHT || v == CS_BC_INTER_SUBDOMAINS || v == CS_BC_PERIODIC;
    };
    if (is_seam(f0) && is_seam(f1)) { *seam = 1; *gcbc = 0; return CS_OK; }
    if (f0 == CS_BC_NON_REFLECTIVE && f1 == CS_BC_NON_REFLECTIVE) { *seam = 0; *gcbc = 1; return CS_OK; }
    return fail(CS_ERR_UNSUPPORTED, "fused path: each axis needs two non-reflective faces or two seam faces "
                                    "(walls and curved interfaces use the per-operator API)");
}

namespace {
struct EventPair {
    hipEvent_t a = nullptr, b = nullptr;
    ~EventPair() {
        if (a) (void)hipEventDestroy(a);
        if (b) (void)hipEventDestroy(b);
    }
};

// mean time of one call of `fn` over `reps` calls after one untimed call
template <class F>
int time_calls(F&& fn, int reps, hipStream_t s, EventPair& ev, float* ms) {
    int rc = fn();
    if (rc) return rc;
    CS_HIP(hipEventRecord(ev.a, s));
    for (int i = 0; i < reps; i++)
        if ((rc = fn())) return rc;
    CS_HIP(hipEventRecord(ev.b, s));
    CS_HIP(hipEventSynchronize(ev.b));
    CS_HIP(hipEventElapsedTime(ms, ev.a, ev.b));
    *ms /= reps;
    return CS_OK;
}
}  // namespace

extern "C" {

int cs_solver_destroy(cs_solver* sv) {
    if (!sv) return CS_OK;
    double* bufs[] = {sv->ga, sv->gb, sv->tau, sv->hf, sv->dpart, sv->qb, sv->qc, sv->red, sv->dt, sv->vjp,
                      sv->tabs[0], sv->tabs[1], sv->tabs[2], sv->tabs16[0], sv->tabs16[1], sv->tabs16[2],
                      sv->ftabs16[0], sv->ftabs16[1], sv->ftabs16[2],
                      sv->ftabs[0], sv->ftabs[1], sv->ftabs[2],
                      sv->ttabs[0], sv->ttabs[1], sv->ttabs[2], sv->fttabs[0], sv->fttabs[1], sv->fttabs[2],
                      sv->adj, sv->dtpart};
    for (double* p : bufs)
        if (p) (void)hipFree(p);
    if (sv->dt_host) (void)hipHostFree(sv->dt_host);
    if (sv->ev_fork) (void)hipEventDestroy(sv->ev_fork);
    if (sv->ev_join) (void)hipEventDestroy(sv->ev_join);
    if (sv->aux) (void)hipStreamDestroy(sv->aux);
    if (sv->num) cs_numerics_destroy(sv->num);
    delete sv;
    return CS_OK;
}

// the right-hand-side sweeps of `axis` on 16-point chunks (lines of 32..256
// points, a multiple of 16); the filter and adjoint sweeps keep 8-point chunks
static int use_chunks16(cs_solver* sv, int axis, int lines, int xcd) {
    const AxisDesc& ax = sv->ax[axis];
    if (axis == 0 || ax.K != 8 || ax.N % 16 != 0 || ax.N < 32)
        return fail(CS_ERR_UNSUPPORTED, "16-point chunks: y or z lines of 32..256 points, a multiple of 16");
    if (!sv->tabs16[axis]) {
        std::vector<double> h;
        int rc = build_axis_table(ax.N, ax.seam, 16, ax.N / 16, h);
        if (rc) return rc;
        CS_HIP(hipMalloc(&sv->tabs16[axis], h.size() * sizeof(double)));
        CS_HIP(hipMemcpy(sv->tabs16[axis], h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice));
    }
    AxisDesc r = ax;
    r.K = 16;
    r.P = ax.N / 16;
    r.GS = 1;
    while (r.GS < r.P) r.GS <<= 1;
    r.tab = sv->tabs16[axis];
    r.variant = 2;
    r.tl = lines;
    r.xcd = xcd ? 1 : 0;
    sv->rax[axis] = r;
    return CS_OK;
}

// the filter sweep of `axis` (y or z) on 16-point chunks (lines of 32..256
// points, a multiple of 16) or back on the 8-point chunks of ax
static int filter_chunks(cs_solver* sv, int axis, int chunk) {
    const AxisDesc& ax = sv->ax[axis];
    AxisDesc f = ax;
    f.ffields = sv->fax[axis].ffields;
    f.ftl = sv->fax[axis].ftl;
    if (chunk != ax.K && (ax.K != 8 || axis == 0))
        return fail(CS_ERR_UNSUPPORTED, "this axis's filter sweep runs " + std::to_string(ax.K) + "-point chunks");
    if (chunk == 16 && ax.K == 8) {
        if (!ax.ftab || ax.N % 16 != 0 || ax.N < 32)
            return fail(CS_ERR_UNSUPPORTED, "16-point filter chunks: y or z lines of 32..256 points, a multiple of 16");
        if (!sv->ftabs16[axis]) {
            std::vector<double> hf;
            int rc = build_line_table(OP_FILTER, ax.N, ax.seam, ax.seam, sv->cfg.alpha_f, 16, hf);
            if (rc) return rc;
            CS_HIP(hipMalloc(&sv->ftabs16[axis], hf.size() * sizeof(double)));
            CS_HIP(hipMemcpy(sv->ftabs16[axis], hf.data(), hf.size() * sizeof(double), hipMemcpyHostToDevice));
        }
        f.K = 16;
        f.P = ax.N / 16;
        f.GS = 1;
        while (f.GS < f.P) f.GS <<= 1;
        f.ftab = sv->ftabs16[axis];
    }
    sv->fax[axis] = f;
    return CS_OK;
}

static void apply_tuned(cs_solver* sv, bool filters);

int cs_solver_create(const cs_solver_config* cfg, cs_solver** out) {
    if (!cfg || !out) return fail(CS_ERR_ARG, "null argument");
    *out = nullptr;
    Geo g;
    int rc = check_dims(&cfg->dims, &g);
    if (rc) return rc;
    int seam[3], gc[3];
    for (int a = 0; a < 3; a++) {
        if ((rc = axis_mode(cfg->nbc, a, &seam[a], &gc[a]))) return rc;
        if (g.n(a) < 16) return fail(CS_ERR_SHAPE, "fused path needs at least 16 points per axis");
        if (g.n(a) > 512) return fail(CS_ERR_SHAPE, "fused path supports at most 512 points per axis");
    }
    if (g.lmx > (1L << 29)) return fail(CS_ERR_SHAPE, "fused path: a field must stay below 4 GiB (2^29 points)");
    cs_solver* sv = new cs_solver();
    sv->cfg = *cfg;
    sv->g = g;
    if ((rc = cs_numerics_create(&cfg->dims, cfg->nbc, &sv->num)) || (rc = cs_deriv_setup(sv->num)) ||
        (rc = cs_set_filter_alpha(sv->num, cfg->alpha_f))) {
        std::string msg = cs_last_error();
        cs_solver_destroy(sv);
        return fail(rc, msg);
    }
    const long lmx = g.lmx;
    for (int a = 0; a < 3; a++) {
        AxisDesc& ax = sv->ax[a];
        ax.N = g.n(a);
        // 8-point chunks up to 256 points, and along x up to 512 when they are
        // equal (one line per wave of the transposing x kernel, 64 chunks);
        // 16-point chunks otherwise
        ax.K = (ax.N <= 256 || (a == 0 && ax.N % 8 == 0)) ? 8 : 16;
        ax.P = (ax.N + ax.K - 1) / ax.K;
        ax.GS = 1;
        while (ax.GS < ax.P) ax.GS <<= 1;
        ax.nlines = g.face_size(a);
        ax.st = g.stride(a);
        ax.seam = seam[a];
        ax.gcbc = gc[a];
        // measured defaults (256^3, MI355X): x lines run the one-wave transposing
        // kernel; y and z lines 16 per workgroup, y with XCD-grouped workgroups
        ax.variant = 0;
        ax.tl = 16;
        ax.xcd = a == 1 ? 1 : 0;
        // filter sweeps: x all five variables per wave, y/z one per workgroup
        ax.ffields = a == 0 ? 5 : 1;
        ax.ftl = 16;
        std::vector<double> h;
        build_axis_table(ax.N, ax.seam, ax.K, ax.P, h);
        if (hipMalloc(&sv->tabs[a], h.size() * sizeof(double)) != hipSuccess ||
            hipMemcpy(sv->tabs[a], h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) {
            cs_solver_destroy(sv);
            return fail(CS_ERR_HIP, "operator table upload failed");
        }
        ax.tab = sv->tabs[a];
        ax.ftab = nullptr;
        if (cfg->filter && ax.N == ax.P * ax.K && (a == 0 ? ax.K == 8 : ax.P <= kMaxP)) {
            // the filter of the five variables in one sweep per axis (equal chunks)
            std::vector<double> hf;
            const int sp = ax.seam;
            if ((rc = build_line_table(OP_FILTER, ax.N, sp, sp, cfg->alpha_f, ax.K, hf))) {
                cs_solver_destroy(sv);
                return rc;
            }
            if (hipMalloc(&sv->ftabs[a], hf.size() * sizeof(double)) != hipSuccess ||
                hipMemcpy(sv->ftabs[a], hf.data(), hf.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) {
                cs_solver_destroy(sv);
                return fail(CS_ERR_HIP, "filter table upload failed");
            }
            ax.ftab = sv->ftabs[a];
        }
    }
    for (int a = 0; a < 3; a++) sv->rax[a] = sv->fax[a] = sv->ax[a];
    // measured default (256^3, MI355X): the y and z filter sweeps run 16-point chunks
    for (int a = 1; a < 3 && cfg->filter; a++)
        if (sv->ax[a].ftab && sv->ax[a].K == 8 && sv->ax[a].N % 16 == 0 && sv->ax[a].N >= 32 &&
            (rc = filter_chunks(sv, a, 16))) {
            std::string msg = cs_last_error();
            cs_solver_destroy(sv);
            return fail(rc, msg);
        }
    // measured default (256^3, MI355X): z lines of up to 256 points run the
    // right-hand-side sweeps on 16-point chunks (S1z 0.345 -> 0.305 ms, S2z
    // 0.525 -> 0.518 ms; y lines measured the same or slower)
    if (sv->ax[2].K == 8 && sv->ax[2].N % 16 == 0 && sv->ax[2].N >= 32 && (rc = use_chunks16(sv, 2, 16, 0))) {
        std::string msg = cs_last_error();
        cs_solver_destroy(sv);
        return fail(rc, msg);
    }
    sv->red_blocks = (int)std::min<long>(2048, (lmx + 255) / 256);
    bool ok = hipMalloc(&sv->dpart, sizeof(double) * 5 * lmx) == hipSuccess &&
              hipMalloc(&sv->qb, sizeof(double) * 5 * lmx) == hipSuccess &&
              hipMalloc(&sv->qc, sizeof(double) * 5 * lmx) == hipSuccess &&
              hipMalloc(&sv->red, sizeof(double) * 2 * sv->red_blocks) == hipSuccess &&
              hipMalloc(&sv->dt, sizeof(double) * 4) == hipSuccess &&
              hipHostMalloc(&sv->dt_host, sizeof(double) * 4) == hipSuccess;
    // one pair of time-step maxima per workgroup of a y or z gradient sweep
    // (at least 8 lines per workgroup) or of the x sweep (at least one line)
    const long fold_blocks = std::max((std::max(g.face_size(1), g.face_size(2)) + 7) / 8, g.face_size(0));
    if (ok && cfg->viscous)
        ok = hipMalloc(&sv->ga, sizeof(double) * 4 * lmx) == hipSuccess &&
             hipMalloc(&sv->gb, sizeof(double) * 4 * lmx) == hipSuccess &&
             hipMalloc(&sv->tau, sizeof(double) * 6 * lmx) == hipSuccess &&
             hipMalloc(&sv->hf, sizeof(double) * 3 * lmx) == hipSuccess &&
             hipMalloc(&sv->dtpart, sizeof(double) * 2 * fold_blocks) == hipSuccess &&
             hipStreamCreateWithFlags(&sv->aux, hipStreamNonBlocking) == hipSuccess &&
             hipEventCreateWithFlags(&sv->ev_fork, hipEventDisableTiming) == hipSuccess &&
             hipEventCreateWithFlags(&sv->ev_join, hipEventDisableTiming) == hipSuccess;
    if (!ok || hipMemset(sv->dt, 0, sizeof(double) * 4) != hipSuccess) {
        cs_solver_destroy(sv);
        return fail(CS_ERR_HIP, "allocation of solver buffers failed");
    }
    for (int a = 0; a < 3; a++) {
        const AxisDesc &r = sv->rax[a], &x = sv->ax[a];
        const int v[6] = {r.variant, r.tl, r.xcd, x.variant, x.tl, x.xcd};
        for (int i = 0; i < 6; i++) sv->shape0[a][i] = v[i];
    }
    apply_tuned(sv, true);
    *out = sv;
    return CS_OK;
}

int cs_solver_rhs(cs_solver* sv, const double* qa, double* de, const double* sbcc, cs_stream stream) {
    if (!sv || !qa || !de) return fail(CS_ERR_ARG, "null argument");
    if (!sbcc && (sv->ax[0].gcbc || sv->ax[1].gcbc || sv->ax[2].gcbc))
        return fail(CS_ERR_ARG, "non-reflective faces need sbcc");
    return rhs_sweeps(sv, qa, sbcc, de, nullptr, 0, (hipStream_t)stream);
}

int cs_solver_sweep(cs_solver* sv, int which, const double* qa, double* de, const double* sbcc, cs_stream stream) {
    if (!sv || !qa || !de) return fail(CS_ERR_ARG, "null argument");
    if (!sbcc && (sv->ax[0].gcbc || sv->ax[1].gcbc || sv->ax[2].gcbc))
        return fail(CS_ERR_ARG, "non-reflective faces need sbcc");
    return one_sweep(sv, which, qa, sbcc, de, (hipStream_t)stream);
}

// launch shape of the right-hand-side sweeps along `axis`
static int sweep_shape(cs_solver* sv, int axis, int variant, int lines, int xcd) {
    if (variant < 0 || variant > 2 || (lines != 8 && lines != 16))
        return fail(CS_ERR_ARG, "variant 0/1/2, lines 8/16");
    if (variant == 1 && (sv->ax[axis].K != 8 || sv->ax[axis].P > kMaxP))
        return fail(CS_ERR_UNSUPPORTED, "tiled variant needs 8-point chunks, at most 256 points");
    if (variant == 2) return use_chunks16(sv, axis, lines, xcd);
    AxisDesc& ax = sv->ax[axis];
    ax.variant = variant;
    ax.tl = lines;
    ax.xcd = xcd ? 1 : 0;
    sv->rax[axis] = ax;
    return CS_OK;
}

// launch shape of the fused filter sweep along `axis`
static int filter_shape(cs_solver* sv, int axis, int fields, int lines, int chunk) {
    if ((fields != 1 && fields != 5) || (lines != 8 && lines != 16) || (chunk != 8 && chunk != 16))
        return fail(CS_ERR_ARG, "fields 1/5, lines 8/16, chunk 8/16");
    if (!filter_axis_fused(sv, axis)) return fail(CS_ERR_UNSUPPORTED, "the filter of this axis has no fused sweep");
    int rc = filter_chunks(sv, axis, chunk);
    if (rc) return rc;
    sv->ax[axis].ffields = sv->fax[axis].ffields = fields;
    sv->ax[axis].ftl = sv->fax[axis].ftl = lines;
    return CS_OK;
}

// applies the process-wide launch-shape choices for this solver's axes: the
// right-hand-side sweeps take the entry of the current precision and metric
// kind, or their default shape when there is none; the filter sweeps take
// theirs when `filters`.  A choice that this solver cannot run (a cache
// written for another shape) is skipped.
static void apply_tuned(cs_solver* sv, bool filters) {
    for (int a = 0; a < 3; a++) {
        AxisDesc& ax = sv->ax[a];
        const int* d = sv->shape0[a];
        ax.variant = d[3];
        ax.tl = d[4];
        ax.xcd = d[5];
        int v[3];
        if (!tuned_sweep(a, ax.N, ax.seam, sv->cfg.viscous ? sv->prec : 0, sv->curv, v) ||
            sweep_shape(sv, a, v[0], v[1], v[2]))
            (void)sweep_shape(sv, a, d[0], d[1], d[2]);
        if (filters && filter_axis_fused(sv, a) && tuned_filter(a, ax.N, ax.seam, v))
            (void)filter_shape(sv, a, v[0], v[1], v[2]);
    }
}

int cs_solver_tune(cs_solver* sv, int axis, int variant, int lines, int xcd) {
    if (!sv || axis < 0 || axis > 2) return fail(CS_ERR_ARG, "bad solver or axis");
    return sweep_shape(sv, axis, variant, lines, xcd);
}

int cs_solver_tune_filter(cs_solver* sv, int axis, int fields, int lines, int chunk) {
    if (!sv || axis < 0 || axis > 2) return fail(CS_ERR_ARG, "bad solver or axis");
    return filter_shape(sv, axis, fields, lines, chunk);
}

int cs_solver_get_tune(const cs_solver* sv, int axis, int* shape6) {
    if (!sv || axis < 0 || axis > 2 || !shape6) return fail(CS_ERR_ARG, "bad solver, axis or output");
    const AxisDesc& r = sv->rax[axis];
    const AxisDesc& f = sv->fax[axis];
    const bool fused = filter_axis_fused(sv, axis);
    const int v[6] = {r.variant, r.tl, r.xcd, fused ? f.ffields : 0, fused ? f.ftl : 0, fused ? f.K : 0};
    for (int i = 0; i < 6; i++) shape6[i] = v[i];
    return CS_OK;
}


int cs_solver_autotune(cs_solver* sv, const double* qa, const double* sbcc, int reps, cs_stream stream) {
    if (!sv || !qa) return fail(CS_ERR_ARG, "null argument");
    if (!sbcc && (sv->ax[0].gcbc || sv->ax[1].gcbc || sv->ax[2].gcbc))
        return fail(CS_ERR_ARG, "non-reflective faces need sbcc");
    if (reps < 1 || reps > 1000) return fail(CS_ERR_ARG, "reps in 1..1000");
    hipStream_t s = (hipStream_t)stream;
    EventPair ev;
    CS_HIP(hipEventCreate(&ev.a));
    CS_HIP(hipEventCreate(&ev.b));
    double* out = sv->qb;   // stage buffer, free outside cs_solver_rk4_step
    const int prec = sv->cfg.viscous ? sv->prec : 0;
    // one whole RHS so that every intermediate a single sweep reads is written
    int rc = rhs_sweeps(sv, qa, sbcc, out, nullptr, 0, s);
    if (rc) return rc;
    // a candidate replaces the shape in use only when it is 2% faster, so
    // that timing noise does not flip between equal shapes
    constexpr float kGain = 0.98f;
    for (int a = 0; a < 3; a++) {
        // the gradient and flux sweeps of this axis, by launch position
        int pos[2], np = 0;
        for (int k = 0; k < 3; k++) {
            if (sv->cfg.viscous && kS1Axes[sv->order][k] == a) pos[np++] = k;
            if (kS2Axes[sv->order][k] == a) pos[np++] = 3 + k;
        }
        auto sweeps = [&]() {
            for (int i = 0; i < np; i++) {
                const int r = one_sweep(sv, pos[i], qa, sbcc, out, s);
                if (r) return r;
            }
            return (int)CS_OK;
        };
        const AxisDesc cur = sv->rax[a], ax0 = sv->ax[a];
        int best[3] = {cur.variant, cur.tl, cur.xcd};
        float tbest;
        if ((rc = time_calls(sweeps, reps, s, ev, &tbest))) return rc;
        for (int v = 0; v < 3; v++)
            for (int l : {8, 16})
                for (int x = 0; x < 2; x++) {
                    // x lines: the transposing kernel ignores lines and XCD
                    // grouping, the tiled one takes lines only
                    if (a == 0 && (x == 1 || (v == 0 && l == 8))) continue;
                    if (v == best[0] && l == best[1] && x == best[2]) continue;
                    if (sweep_shape(sv, a, v, l, x)) continue;   // not admissible on this axis
                    float t;
                    if ((rc = time_calls(sweeps, reps, s, ev, &t))) return rc;
                    if (t < kGain * tbest) {
                        tbest = t;
                        best[0] = v, best[1] = l, best[2] = x;
                    }
                }
        sv->ax[a] = ax0;
        if ((rc = sweep_shape(sv, a, best[0], best[1], best[2]))) return rc;
        set_tuned_sweep(a, sv->ax[a].N, sv->ax[a].seam, prec, sv->curv, best);
    }
    if (!sv->cfg.filter) return CS_OK;
    CS_HIP(hipMemcpyAsync(out, qa, sizeof(double) * 5 * sv->g.lmx, hipMemcpyDeviceToDevice, s));
    for (int a = 0; a < 3; a++) {
        if (!filter_axis_fused(sv, a)) continue;
        auto filt = [&]() { return filter_sweeps(sv, out, s, a); };
        const AxisDesc cur = sv->fax[a];
        int best[3] = {cur.ffields, cur.ftl, cur.K};
        float tbest;
        if ((rc = time_calls(filt, reps, s, ev, &tbest))) return rc;
        for (int f : {1, 5})
            for (int l : {8, 16})
                for (int c : {8, 16}) {
                    if (a == 0 && l == 8) continue;   // the x filter ignores lines
                    if (f == best[0] && l == best[1] && c == best[2]) continue;
                    if (filter_shape(sv, a, f, l, c)) continue;
                    float t;
                    if ((rc = time_calls(filt, reps, s, ev, &t))) return rc;
                    if (t < kGain * tbest) {
                        tbest = t;
                        best[0] = f, best[1] = l, best[2] = c;
                    }
                }
        if ((rc = filter_shape(sv, a, best[0], best[1], best[2]))) return rc;
        set_tuned_filter(a, sv->ax[a].N, sv->ax[a].seam, best);
    }
    return CS_OK;
}

int cs_solver_set_metrics(cs_solver* sv, const double* xim, const double* etm, const double* zem,
                          const double* yaco, const double* cm0, const double* cm1, const double* cm2) {
    if (!sv) return fail(CS_ERR_ARG, "null solver");
    const double* rows[3] = {xim, etm, zem};
    const double* cms[3] = {cm0, cm1, cm2};
    const int given = (xim != nullptr) + (etm != nullptr) + (zem != nullptr) + (yaco != nullptr) +
                      (cm0 != nullptr) + (cm1 != nullptr) + (cm2 != nullptr);
    if (given == 0) {            // back to the constants of the configuration
        const bool was = sv->curv != 0;
        sv->curv = 0;
        sv->met = Metric{};
        if (was) apply_tuned(sv, false);   // the sweep shapes of constant metrics
        return CS_OK;
    }
    if (given != 7) return fail(CS_ERR_ARG, "metrics: xim, etm, zem, yaco and the three face-normal arrays");
    Metric m{};
    const long lmx = sv->g.lmx;
    for (int a = 0; a < 3; a++) {
        for (int k = 0; k < 3; k++) m.m[a][k] = rows[a] + (long)k * lmx;
        m.cm[a] = cms[a];
    }
    m.J = yaco;
    const bool was = sv->curv != 0;
    sv->met = m;
    sv->curv = 1;
    if (!was) apply_tuned(sv, false);   // the sweep shapes of per-point metrics
    return CS_OK;
}

int cs_solver_set_precision(cs_solver* sv, int mode) {
    if (!sv || (mode != 0 && mode != 1)) return fail(CS_ERR_ARG, "precision mode is 0 or 1");
    if (mode == sv->prec) return CS_OK;
    sv->prec = mode;
    apply_tuned(sv, false);
    return CS_OK;
}

int cs_solver_set_order(cs_solver* sv, int order) {
    if (!sv || (order != 0 && order != 1)) return fail(CS_ERR_ARG, "order is 0 or 1");
    sv->order = order;
    return CS_OK;
}

int cs_solver_time_step(cs_solver* sv, const double* qa, double* dte, cs_stream stream) {
    if (!sv || !qa) return fail(CS_ERR_ARG, "null argument");
    hipStream_t s = (hipStream_t)stream;
    int rc = time_step(sv, qa, 0.0, s);
    if (rc || !dte) return rc;
    CS_HIP(hipMemcpyAsync(sv->dt_host, sv->dt, 3 * sizeof(double), hipMemcpyDeviceToHost, s));
    CS_HIP(hipStreamSynchronize(s));
    *dte = sv->dt_host[0];
    return CS_OK;
}

int cs_solver_rk4_step(cs_solver* sv, double* qa, const double* sbcc, double dt, double* dt_out, cs_stream stream) {
    if (!sv || !qa) return fail(CS_ERR_ARG, "null argument");
    if (!sbcc && (sv->ax[0].gcbc || sv->ax[1].gcbc || sv->ax[2].gcbc))
        return fail(CS_ERR_ARG, "non-reflective faces need sbcc");
    hipStream_t s = (hipStream_t)stream;
    // a computed time step is reduced inside a gradient sweep of the first
    // stage where one can (dt_fold_pos; dt is first read by the stage update
    // at the end of that stage), else in a pass of its own
    const int fold = dt > 0.0 ? -1 : dt_fold_pos(sv);
    int rc = fold >= 0 ? CS_OK : time_step(sv, qa, dt, s);
    if (rc) return rc;
    sv->dt_fold = fold + 1;
    rc = rhs_sweeps(sv, qa, sbcc, sv->qb, qa, 4, s);
    sv->dt_fold = 0;
    if (rc) return rc;
    if ((rc = rhs_sweeps(sv, sv->qb, sbcc, sv->qc, qa, 3, s))) return rc;
    if ((rc = rhs_sweeps(sv, sv->qc, sbcc, sv->qb, qa, 2, s))) return rc;
    if ((rc = rhs_sweeps(sv, sv->qb, sbcc, qa, qa, 1, s))) return rc;
    if ((rc = step_filter(sv, qa, s))) return rc;
    if (dt_out) {
        CS_HIP(hipMemcpyAsync(sv->dt_host, sv->dt, 3 * sizeof(double), hipMemcpyDeviceToHost, s));
        CS_HIP(hipStreamSynchronize(s));
        *dt_out = sv->dt_host[0];
    }
    return CS_OK;
}

}  // extern "C"
