// batch.hip -- many independent subdomains through a few persistent launches.
//
// The plain method of the reference (compact first derivative along x, y, z,
// deriv1d/deriv2d of src/numerics_pc.hpp:74-168 with the seam wrap of a
// single periodic rank) followed by the compact filter along x, y, z, applied
// to a set of subdomains of unrelated shapes.  Every (subdomain, axis) line
// operator is cut into workgroup-sized items, and the items of all
// subdomains that share a kernel variant (chunk length K, axis) and a chunk
// count P (so a workgroup has exactly the threads its lines need) run in one
// launch: the derivative phase has one launch per such class present (the
// three derivatives are independent), the filter one per class and axis, x,
// then y, then z (each filters the previous result in place).  Items
// are ordered longest first and dealt round-robin to a grid sized to fill the
// device, so the uneven subdomains share the CUs instead of each paying the
// tail of a small launch.
//
// An item runs the same device body as the single-subdomain kernels
// (line_body in line_solver.hip) with the same operator table, so every
// subdomain's result is bit-identical to cs_deriv / cs_filter on it alone.
//
// The launches can be recorded into a hipGraph (cs_batch_capture) and
// replayed with one call; nothing in them allocates or clears memory.
#include "cs_internal.hpp"

#include <algorithm>
#include <cstdio>

using namespace cs;

namespace {

constexpr int kLanes = 8;      // streams of a grid-mode-2 step

struct Launch {                // one kernel variant (K, axis) and chunk count P of a phase
    int K = 8, axis = 0, P = 1;
    int first = 0, nitems = 0;  // range in the phase's item list
    int threads = 0, grid = 0;
    size_t lds = 0;
};

struct Phase {
    int op = OP_DERIV;
    LineArgs* args = nullptr;   // device
    int2* items = nullptr;      // device
    int nitems = 0;
    std::vector<Launch> launches;
};

}  // namespace

struct cs_batch {
    int count = 0;
    std::vector<cs_numerics*> num;
    std::vector<double*> f;     // bound fields
    std::vector<double*> d;     // bound derivative outputs [3*count]
    Phase ph[4];                // derivative (all axes), filter x, y, z
    bool bound = false;
    int persistent = 0;         // 1: grid = resident workgroups, items round-robin;
                                // 0: one workgroup per item (the dispatcher balances);
                                // 2: one launch per subdomain and operator (line_kernel)
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    // grid mode 2, whole steps: the subdomains dealt over independent streams
    // (default priority) so that the small launches of different subdomains
    // overlap; fork/join through events (nothing allocates in a capture)
    std::vector<hipStream_t> lanes;
    std::vector<hipEvent_t> lane_ev;   // [0]: fork, [1 + l]: join of lane l
};

static void free_phase(Phase& p) {
    if (p.args) (void)hipFree(p.args);
    if (p.items) (void)hipFree(p.items);
    p = Phase();
}

// streams and events of the lanes, created once outside any capture
static int make_lanes(cs_batch* b) {
    if (!b->lanes.empty()) return CS_OK;
    const int n = std::min(kLanes, b->count);
    b->lanes.assign(n, nullptr);
    b->lane_ev.assign(n + 1, nullptr);
    for (int l = 0; l < n; l++) CS_HIP(hipStreamCreateWithFlags(&b->lanes[l], hipStreamNonBlocking));
    for (int e = 0; e <= n; e++) CS_HIP(hipEventCreateWithFlags(&b->lane_ev[e], hipEventDisableTiming));
    return CS_OK;
}

static void drop_lanes(cs_batch* b) {
    for (hipStream_t l : b->lanes)
        if (l) (void)hipStreamDestroy(l);
    for (hipEvent_t e : b->lane_ev)
        if (e) (void)hipEventDestroy(e);
    b->lanes.clear();
    b->lane_ev.clear();
}

static void drop_graph(cs_batch* b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
}

// line operator `op` along `axis` of subdomain `s`, reading `in`, writing `out`
static int line_args(cs_numerics* num, int op, int axis, const double* in, double* out, LineArgs* a, int* K) {
    const int start = num->ndf[0][axis], end = num->ndf[1][axis];
    if (op == OP_DERIV && (start == 1 || end == 1) && !num->table_set)
        return fail(CS_ERR_STATE, "seam rows need an interface table");
    LineOpTable* t = nullptr;
    int rc = get_line_op(num, op, axis, start, end, &t);
    if (rc) return rc;
    *a = LineArgs{};
    a->in = in;
    a->out = out;
    a->tab = t->dev;
    a->pb = num->pb;
    a->recv = nullptr;     // seam sums formed inside the kernel (single-subdomain wrap)
    a->drva = nullptr;
    a->var = 0;
    a->N = t->n;
    a->nlines = (int)num->g.face_size(axis);
    a->lxi = num->g.lxi;
    a->plane = (long)num->g.lxi * num->g.let;
    a->st = num->g.stride(axis);
    a->P = t->P;
    a->seam0 = op == OP_DERIV ? start : 0;
    a->seam1 = op == OP_DERIV ? end : 0;
    a->cyclic = op == OP_FILTER ? (start == 1 && end == 1) : 0;
    a->tab32 = t->dev32;
    a->mixed = 0;              // batched operators solve in FP64
    *K = t->K;
    return CS_OK;
}

static int build_phase(cs_batch* b, Phase& p, int op, int axis_lo, int axis_hi) {
    struct Item { int variant; double cost; int2 it; };
    std::vector<LineArgs> args;
    std::vector<Item> work;
    std::vector<Launch> var;                       // one per (K, axis) present
    p.op = op;
    for (int s = 0; s < b->count; s++) {
        for (int axis = axis_lo; axis <= axis_hi; axis++) {
            LineArgs a;
            int K = 0;
            double* out = op == OP_DERIV ? b->d[3 * s + axis] : b->f[s];
            int rc = line_args(b->num[s], op, axis, b->f[s], out, &a, &K);
            if (rc) return rc;
            int v = -1;
            for (size_t k = 0; k < var.size(); k++)
                if (var[k].K == K && var[k].axis == axis && var[k].P == a.P) v = (int)k;
            if (v < 0) {
                Launch l;
                l.K = K;
                l.axis = axis;
                l.P = a.P;
                var.push_back(l);
                v = (int)var.size() - 1;
            }
            const int TL = batch_lines_per_item(K, axis, a.P);
            var[v].threads = std::max(var[v].threads, TL * a.P);
            var[v].lds = std::max(var[v].lds, batch_lds_bytes(K, axis, a.N));
            const int idx = (int)args.size();
            args.push_back(a);
            const int nblk = (a.nlines + TL - 1) / TL;
            const double cost = (double)a.N * std::min(TL, a.nlines);
            for (int k = 0; k < nblk; k++) work.push_back({v, cost, make_int2(idx, k)});
        }
    }
    std::stable_sort(work.begin(), work.end(), [](const Item& x, const Item& y) {
        return x.variant != y.variant ? x.variant < y.variant : x.cost > y.cost;
    });
    std::vector<int2> items(work.size());
    for (size_t i = 0; i < work.size(); i++) items[i] = work[i].it;
    p.nitems = (int)items.size();

    int dev = 0, ncu = 0;
    CS_HIP(hipGetDevice(&dev));
    CS_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    size_t pos = 0;
    for (size_t v = 0; v < var.size(); v++) {
        Launch& l = var[v];
        l.first = (int)pos;
        while (pos < work.size() && work[pos].variant == (int)v) pos++;
        l.nitems = (int)pos - l.first;
        int nb = 1;
        int rc = batch_blocks_per_cu(op, l.K, l.axis, l.P, l.threads, l.lds, &nb);
        if (rc) return rc;
        l.grid = std::max(1, std::min(l.nitems, ncu * nb));
    }
    p.launches = var;

    CS_HIP(hipMalloc(&p.args, sizeof(LineArgs) * std::max<size_t>(1, args.size())));
    CS_HIP(hipMalloc(&p.items, sizeof(int2) * std::max<size_t>(1, items.size())));
    CS_HIP(hipMemcpy(p.args, args.data(), sizeof(LineArgs) * args.size(), hipMemcpyHostToDevice));
    CS_HIP(hipMemcpy(p.items, items.data(), sizeof(int2) * items.size(), hipMemcpyHostToDevice));
    return CS_OK;
}

// grid mode 2: the single-subdomain kernels (cs_deriv / cs_filter), one launch
// per subdomain and operator -- shape-specialised, and cheap to launch once
// recorded into the batch's hipGraph
static int run_direct(cs_batch* b, int k, hipStream_t s) {
    for (int sd = 0; sd < b->count; sd++) {
        cs_numerics* num = b->num[sd];
        const int a0 = k == 0 ? 0 : k - 1, a1 = k == 0 ? 2 : k - 1;
        for (int axis = a0; axis <= a1; axis++) {
            const int start = num->ndf[0][axis], end = num->ndf[1][axis];
            const int rc = k == 0 ? line_apply(num, OP_DERIV, axis, b->f[sd], b->d[3 * sd + axis], start, end,
                                               nullptr, 0, false, s)
                                  : line_apply(num, OP_FILTER, axis, b->f[sd], b->f[sd], start, end, nullptr, 0,
                                               false, s);
            if (rc) return rc;
        }
    }
    return CS_OK;
}

// a whole step in grid mode 2: subdomain sd's derivatives and then its
// filter run in order on lane sd % lanes; the lanes fork from and join back
// into s.  Each subdomain sees the same launches as run_direct, in the same
// order, so the results are the same bit for bit.
static int run_direct_lanes(cs_batch* b, hipStream_t s) {
    int rc = make_lanes(b);
    if (rc) return rc;
    const int n = (int)b->lanes.size();
    CS_HIP(hipEventRecord(b->lane_ev[0], s));
    for (int l = 0; l < n; l++) CS_HIP(hipStreamWaitEvent(b->lanes[l], b->lane_ev[0], 0));
    for (int sd = 0; sd < b->count && !rc; sd++) {
        cs_numerics* num = b->num[sd];
        hipStream_t ls = b->lanes[sd % n];
        for (int axis = 0; axis < 3 && !rc; axis++)
            rc = line_apply(num, OP_DERIV, axis, b->f[sd], b->d[3 * sd + axis], num->ndf[0][axis],
                            num->ndf[1][axis], nullptr, 0, false, ls);
        for (int axis = 0; axis < 3 && !rc; axis++)
            rc = line_apply(num, OP_FILTER, axis, b->f[sd], b->f[sd], num->ndf[0][axis], num->ndf[1][axis],
                            nullptr, 0, false, ls);
    }
    // join every lane, also after an error, so that a capture can end cleanly
    for (int l = 0; l < n; l++) {
        CS_HIP(hipEventRecord(b->lane_ev[1 + l], b->lanes[l]));
        CS_HIP(hipStreamWaitEvent(s, b->lane_ev[1 + l], 0));
    }
    return rc;
}

// the operator tables mode 2 launches with, built now so that nothing
// allocates while a graph is being captured
static int warm_tables(cs_batch* b) {
    for (cs_numerics* num : b->num)
        for (int axis = 0; axis < 3; axis++)
            for (int op : {OP_DERIV, OP_FILTER}) {
                LineOpTable* t = nullptr;
                const int rc = get_line_op(num, op, axis, num->ndf[0][axis], num->ndf[1][axis], &t);
                if (rc) return rc;
            }
    return CS_OK;
}

static int run_phase(cs_batch* b, int k, int persistent, hipStream_t s) {
    if (persistent == 2) return run_direct(b, k, s);
    const Phase& p = b->ph[k];
    for (const Launch& l : p.launches) {
        const int grid = persistent ? l.grid : l.nitems;
        int rc = launch_batch_lines(p.op, l.K, l.axis, l.P, p.args, p.items + l.first, l.nitems, grid, l.threads,
                                    l.lds, s);
        if (rc) return rc;
    }
    return CS_OK;
}

extern "C" {

int cs_batch_destroy(cs_batch* b) {
    if (!b) return CS_OK;
    drop_graph(b);
    drop_lanes(b);
    for (Phase& p : b->ph) free_phase(p);
    for (cs_numerics* n : b->num) cs_numerics_destroy(n);
    delete b;
    return CS_OK;
}

int cs_batch_create(int count, const cs_dims* dims, const int* nbc6, cs_batch** out) {
    if (!out || !dims || !nbc6 || count < 1) return fail(CS_ERR_ARG, "null argument or empty batch");
    *out = nullptr;
    cs_batch* b = new cs_batch();
    b->count = count;
    for (int s = 0; s < count; s++) {
        cs_numerics* n = nullptr;
        int rc = cs_numerics_create(&dims[s], nbc6 + 6 * s, &n);
        if (!rc) rc = cs_deriv_setup(n);
        if (rc) {
            std::string msg = "subdomain " + std::to_string(s) + ": " + cs_last_error();
            if (n) cs_numerics_destroy(n);
            cs_batch_destroy(b);
            return fail(rc, msg);
        }
        b->num.push_back(n);
    }
    *out = b;
    return CS_OK;
}

int cs_batch_bind(cs_batch* b, double* const* f, double* const* d) {
    if (!b || !f || !d) return fail(CS_ERR_ARG, "null argument");
    drop_graph(b);
    for (Phase& p : b->ph) free_phase(p);
    b->f.assign(f, f + b->count);
    b->d.assign(d, d + 3 * b->count);
    for (int s = 0; s < b->count; s++)
        if (!b->f[s] || !b->d[3 * s] || !b->d[3 * s + 1] || !b->d[3 * s + 2])
            return fail(CS_ERR_ARG, "null field pointer for subdomain " + std::to_string(s));
    b->bound = false;
    int rc = build_phase(b, b->ph[0], OP_DERIV, 0, 2);
    for (int axis = 0; axis < 3 && !rc; axis++) rc = build_phase(b, b->ph[1 + axis], OP_FILTER, axis, axis);
    if (rc) {
        for (Phase& p : b->ph) free_phase(p);
        return rc;
    }
    b->bound = true;
    return CS_OK;
}

int cs_batch_deriv(cs_batch* b, cs_stream stream) {
    if (!b || !b->bound) return fail(CS_ERR_STATE, "batch has no bound fields");
    return run_phase(b, 0, b->persistent, (hipStream_t)stream);
}

int cs_batch_filter(cs_batch* b, cs_stream stream) {
    if (!b || !b->bound) return fail(CS_ERR_STATE, "batch has no bound fields");
    for (int axis = 0; axis < 3; axis++) {
        int rc = run_phase(b, 1 + axis, b->persistent, (hipStream_t)stream);
        if (rc) return rc;
    }
    return CS_OK;
}

int cs_batch_step(cs_batch* b, cs_stream stream) {
    if (b && b->bound && b->persistent == 2) return run_direct_lanes(b, (hipStream_t)stream);
    int rc = cs_batch_deriv(b, stream);
    if (rc) return rc;
    return cs_batch_filter(b, stream);
}

int cs_batch_set_persistent(cs_batch* b, int persistent) {
    if (!b) return fail(CS_ERR_ARG, "null batch");
    if (persistent < 0 || persistent > 2) return fail(CS_ERR_ARG, "grid mode is 0, 1 or 2");
    if (persistent == 2) {
        int rc = warm_tables(b);
        if (!rc) rc = make_lanes(b);
        if (rc) return rc;
    }
    b->persistent = persistent;
    drop_graph(b);      // a captured graph holds the old grid
    return CS_OK;
}

int cs_batch_get_persistent(const cs_batch* b, int* mode) {
    if (!b || !mode) return fail(CS_ERR_ARG, "null argument");
    *mode = b->persistent;
    return CS_OK;
}

int cs_batch_capture(cs_batch* b, cs_stream stream) {
    if (!b || !b->bound) return fail(CS_ERR_STATE, "batch has no bound fields");
    if (!stream) return fail(CS_ERR_ARG, "capture needs a created stream (not the null stream)");
    drop_graph(b);
    hipStream_t s = (hipStream_t)stream;
    CS_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    int rc = cs_batch_step(b, stream);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(s, &g);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    CS_HIP(e);
    b->graph = g;
    CS_HIP(hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0));
    return CS_OK;
}

int cs_batch_replay(cs_batch* b, cs_stream stream) {
    if (!b || !b->exec) return fail(CS_ERR_STATE, "no captured graph (cs_batch_capture)");
    CS_HIP(hipGraphLaunch(b->exec, (hipStream_t)stream));
    return CS_OK;
}

int cs_batch_info(const cs_batch* b, int* items, int* launches) {
    if (!b || !b->bound) return fail(CS_ERR_STATE, "batch has no bound fields");
    for (int k = 0; k < 4; k++) {
        if (items) items[k] = b->ph[k].nitems;
        if (launches) launches[k] = (int)b->ph[k].launches.size();
    }
    return CS_OK;
}

// ---- autotuner -------------------------------------------------------------

// two timing events, destroyed on every path out of the tuner
struct EventPair {
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int create() {
        CS_HIP(hipEventCreate(&e0));
        CS_HIP(hipEventCreate(&e1));
        return CS_OK;
    }
    ~EventPair() {
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
    }
};

// Times every admissible chunk length for each (operator, line length) the
// batch holds, on one of its subdomains, and records the fastest; then
// rebuilds the batch with the choices.  The choices apply to every operator
// table built afterwards in this process (cs_deriv / cs_filter on a new
// context included), so batched and single-subdomain results stay identical.
// The bound fields f are only read: the filter is timed out of place into the
// subdomain's derivative buffer of that axis (which the batch's derivative
// phase overwrites anyway).
int cs_batch_autotune(cs_batch* b, int reps, cs_stream stream) {
    if (!b || !b->bound) return fail(CS_ERR_STATE, "batch has no bound fields");
    if (reps < 1) reps = 3;
    hipStream_t s = (hipStream_t)stream;
    // one case per (operator, line length, start/end rows, precision)
    struct Case { int op, n, start, end, prec, s, axis; };
    std::vector<Case> cases;
    for (int sd = 0; sd < b->count; sd++)
        for (int axis = 0; axis < 3; axis++)
            for (int op : {OP_DERIV, OP_FILTER}) {
                const cs_numerics* num = b->num[sd];
                const Case c{op, num->g.n(axis), num->ndf[0][axis], num->ndf[1][axis], num->precision, sd, axis};
                bool seen = false;
                for (auto& o : cases)
                    seen |= (o.op == c.op && o.n == c.n && o.start == c.start && o.end == c.end && o.prec == c.prec);
                if (!seen) cases.push_back(c);
            }
    int rc = CS_OK;
    {
    EventPair ev;
    if ((rc = ev.create())) return rc;
    hipEvent_t e0 = ev.e0, e1 = ev.e1;
    for (auto& c : cases) {
        cs_numerics* num = b->num[c.s];
        const int start = c.start, end = c.end;
        float best = 1e30f;
        int bestK = 0;
        for (int K : {8, 16, 32}) {
            if ((c.n + K - 1) / K > max_chunks(K)) continue;
            LineOpTable t;
            if ((rc = make_line_op(num, c.op, c.n, start, end, K, &t))) break;
            LineArgs a;
            int Kd = 0;
            double* out = b->d[3 * c.s + c.axis];
            if ((rc = line_args(num, c.op, c.axis, b->f[c.s], out, &a, &Kd))) { free_line_op(t); break; }
            a.tab = t.dev;
            a.P = t.P;
            float ms = 0.f;
            rc = launch_line(c.op, c.axis, K, a, s);
            if (!rc && hipEventRecord(e0, s) != hipSuccess) rc = fail(CS_ERR_HIP, "event record failed");
            for (int r = 0; r < reps && !rc; r++) rc = launch_line(c.op, c.axis, K, a, s);
            if (!rc && (hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                        hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
                rc = fail(CS_ERR_HIP, "event timing failed");
            (void)hipStreamSynchronize(s);
            free_line_op(t);
            if (rc) break;
            if (ms < best) { best = ms; bestK = K; }
        }
        if (rc) break;
        set_tuned_chunk(c.op, c.n, c.start, c.end, c.prec, bestK);
    }
    }
    if (rc) return rc;
    // drop the cached tables of the batch and rebuild with the choices
    for (cs_numerics* num : b->num) {
        for (auto& kv : num->ops) free_line_op(kv.second);
        num->ops.clear();
    }
    std::vector<double*> f = b->f, d = b->d;
    if ((rc = cs_batch_bind(b, f.data(), d.data()))) return rc;
    // then the grid mode, on the derivative phase (it leaves f unchanged)
    float t[3] = {0.f, 0.f, 0.f};
    {
    EventPair ev;
    if ((rc = ev.create())) return rc;
    hipEvent_t e0 = ev.e0, e1 = ev.e1;
    for (int mode = 0; mode < 3 && !rc; mode++) {
        rc = run_phase(b, 0, mode, s);
        if (!rc && hipEventRecord(e0, s) != hipSuccess) rc = fail(CS_ERR_HIP, "event record failed");
        for (int r = 0; r < reps && !rc; r++) rc = run_phase(b, 0, mode, s);
        if (!rc && (hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                    hipEventElapsedTime(&t[mode], e0, e1) != hipSuccess))
            rc = fail(CS_ERR_HIP, "event timing failed");
    }
    }
    if (rc) return rc;
    int best = 0;
    for (int mode = 1; mode < 3; mode++)
        if (t[mode] < t[best]) best = mode;
    b->persistent = best;
    return best == 2 ? warm_tables(b) : CS_OK;
}


}  // extern "C"
