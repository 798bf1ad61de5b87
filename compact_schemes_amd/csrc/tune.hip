// tune.hip -- the process-wide cache of launch-shape choices and its text
// file (host code only).
//
// The reference specialises its NVRTC kernels per line with -D macros
// (src/numerics_rtc.hpp:105-126: ITERS, NSTART/NEND, BLOCKSIZE_LOCAL,
// LXI_LOCAL), i.e. per line length, boundary pair and block shape.  The
// choices here are data: three kinds of entries, each keyed by what the
// kernels' speed depends on.
//   line:   (operator, line length, start/end rows, precision) -> chunk length
//           of the per-operator and batched line kernels;
//   sweep:  (axis, line length, seam axis, precision of the intermediates,
//           metrics: 0 constant per axis, 1 per point) -> kernel variant, lines
//           per workgroup, XCD grouping of the fused right-hand-side sweeps;
//   filter: (axis, line length, seam axis) -> fields per workgroup, lines per
//           workgroup, chunk length of the fused filter sweeps.
// Entries are written by the autotuners (cs_batch_autotune,
// cs_solver_autotune) or by hand, and read when operator tables and solvers
// are built.  File format, one entry per line ('#' starts a comment):
//   line <op> <length> <start> <end> <precision> <K>
//   sweep <axis> <length> <seam> <precision> <metrics> <variant> <lines> <xcd>
//   filter <axis> <length> <seam> <fields> <lines> <chunk>
// A line of three integers "<op> <length> <K>" (the earlier format) is a line
// entry for FP64 and both closure and seam ends; a sweep entry of seven
// integers (no metrics field) is one for constant metrics.
#include "cs_internal.hpp"

#include <cstdio>
#include <cstring>
#include <mutex>

namespace cs {
namespace {

std::mutex g_lock;
std::map<std::array<int, 5>, int> g_line;                   // op, n, start, end, prec -> K
std::map<std::array<int, 5>, std::array<int, 3>> g_sweep;   // axis, n, seam, prec, metrics -> variant, lines, xcd
std::map<std::array<int, 3>, std::array<int, 3>> g_filter;  // axis, n, seam -> fields, lines, chunk

int check_line(int op, int n, int start, int end, int prec, int K) {
    if ((op != OP_DERIV && op != OP_FILTER) || n < 8 || start < 0 || start > 1 || end < 0 || end > 1 || prec < 0 ||
        prec > 4 || (K != 0 && K != 8 && K != 16 && K != 32))
        return fail(CS_ERR_ARG, "line entry: op 0/1, length >= 8, start/end 0/1, precision 0..4, K in {0, 8, 16, 32}");
    if (K != 0 && (n + K - 1) / K > max_chunks(K))
        return fail(CS_ERR_SHAPE, "chunk length too short for a line of " + std::to_string(n) + " points");
    return CS_OK;
}

int check_sweep(int axis, int n, int seam, int prec, int metrics, int variant, int lines, int xcd) {
    if (axis < 0 || axis > 2 || n < 16 || n > 512 || seam < 0 || seam > 1 || prec < 0 || prec > 1 || metrics < 0 ||
        metrics > 1 || variant < 0 || variant > 2 || (lines != 8 && lines != 16) || xcd < 0 || xcd > 1)
        return fail(CS_ERR_ARG, "sweep entry: axis 0..2, length 16..512, seam 0/1, precision 0/1, metrics 0/1, "
                                "variant 0..2, lines 8/16, xcd 0/1");
    return CS_OK;
}

int check_filter(int axis, int n, int seam, int fields, int lines, int chunk) {
    if (axis < 0 || axis > 2 || n < 16 || n > 512 || seam < 0 || seam > 1 || (fields != 1 && fields != 5) ||
        (lines != 8 && lines != 16) || (chunk != 8 && chunk != 16))
        return fail(CS_ERR_ARG, "filter entry: axis 0..2, length 16..512, seam 0/1, fields 1/5, lines 8/16, chunk 8/16");
    return CS_OK;
}

}  // namespace

int tuned_chunk(int op, int n, int start, int end, int prec) {
    std::lock_guard<std::mutex> g(g_lock);
    auto it = g_line.find({op, n, start, end, prec});
    return it == g_line.end() ? 0 : it->second;
}

void set_tuned_chunk(int op, int n, int start, int end, int prec, int K) {
    std::lock_guard<std::mutex> g(g_lock);
    if (K == 0) g_line.erase({op, n, start, end, prec});
    else g_line[{op, n, start, end, prec}] = K;
}

bool tuned_sweep(int axis, int n, int seam, int prec, int metrics, int* v3) {
    std::lock_guard<std::mutex> g(g_lock);
    auto it = g_sweep.find({axis, n, seam, prec, metrics});
    if (it == g_sweep.end()) return false;
    for (int i = 0; i < 3; i++) v3[i] = it->second[i];
    return true;
}

void set_tuned_sweep(int axis, int n, int seam, int prec, int metrics, const int* v3) {
    std::lock_guard<std::mutex> g(g_lock);
    g_sweep[{axis, n, seam, prec, metrics}] = {v3[0], v3[1], v3[2]};
}

bool tuned_filter(int axis, int n, int seam, int* v3) {
    std::lock_guard<std::mutex> g(g_lock);
    auto it = g_filter.find({axis, n, seam});
    if (it == g_filter.end()) return false;
    for (int i = 0; i < 3; i++) v3[i] = it->second[i];
    return true;
}

void set_tuned_filter(int axis, int n, int seam, const int* v3) {
    std::lock_guard<std::mutex> g(g_lock);
    g_filter[{axis, n, seam}] = {v3[0], v3[1], v3[2]};
}

}  // namespace cs

using namespace cs;

extern "C" {

int cs_tune_set(int op, int n, int start, int end, int precision, int K) {
    const int rc = check_line(op, n, start, end, precision, K);
    if (rc) return rc;
    set_tuned_chunk(op, n, start, end, precision, K);
    return CS_OK;
}

int cs_tune_get(int op, int n, int start, int end, int precision, int* K) {
    if (!K) return fail(CS_ERR_ARG, "null argument");
    *K = tuned_chunk(op, n, start, end, precision);
    return CS_OK;
}

int cs_tune_set_sweep(int axis, int n, int seam, int precision, int metrics, int variant, int lines, int xcd) {
    const int rc = check_sweep(axis, n, seam, precision, metrics, variant, lines, xcd);
    if (rc) return rc;
    const int v[3] = {variant, lines, xcd};
    set_tuned_sweep(axis, n, seam, precision, metrics, v);
    return CS_OK;
}

int cs_tune_get_sweep(int axis, int n, int seam, int precision, int metrics, int* found, int* shape3) {
    if (!found || !shape3) return fail(CS_ERR_ARG, "null argument");
    *found = tuned_sweep(axis, n, seam, precision, metrics, shape3) ? 1 : 0;
    return CS_OK;
}

int cs_tune_set_filter_sweep(int axis, int n, int seam, int fields, int lines, int chunk) {
    const int rc = check_filter(axis, n, seam, fields, lines, chunk);
    if (rc) return rc;
    const int v[3] = {fields, lines, chunk};
    set_tuned_filter(axis, n, seam, v);
    return CS_OK;
}

int cs_tune_get_filter_sweep(int axis, int n, int seam, int* found, int* shape3) {
    if (!found || !shape3) return fail(CS_ERR_ARG, "null argument");
    *found = tuned_filter(axis, n, seam, shape3) ? 1 : 0;
    return CS_OK;
}

int cs_tune_clear(void) {
    std::lock_guard<std::mutex> g(g_lock);
    g_line.clear();
    g_sweep.clear();
    g_filter.clear();
    return CS_OK;
}

int cs_tune_save(const char* path) {
    if (!path) return fail(CS_ERR_ARG, "null path");
    FILE* fp = std::fopen(path, "w");
    if (!fp) return fail(CS_ERR_ARG, std::string("cannot write ") + path);
    std::fprintf(fp, "# compact_schemes_amd launch-shape choices\n");
    std::fprintf(fp, "# line <op> <length> <start> <end> <precision> <K>\n");
    std::fprintf(fp, "# sweep <axis> <length> <seam> <precision> <metrics> <variant> <lines> <xcd>\n");
    std::fprintf(fp, "# filter <axis> <length> <seam> <fields> <lines> <chunk>\n");
    {
        std::lock_guard<std::mutex> g(g_lock);
        for (auto& e : g_line)
            std::fprintf(fp, "line %d %d %d %d %d %d\n", e.first[0], e.first[1], e.first[2], e.first[3], e.first[4],
                         e.second);
        for (auto& e : g_sweep)
            std::fprintf(fp, "sweep %d %d %d %d %d %d %d %d\n", e.first[0], e.first[1], e.first[2], e.first[3],
                         e.first[4], e.second[0], e.second[1], e.second[2]);
        for (auto& e : g_filter)
            std::fprintf(fp, "filter %d %d %d %d %d %d\n", e.first[0], e.first[1], e.first[2], e.second[0],
                         e.second[1], e.second[2]);
    }
    const bool bad = std::ferror(fp) != 0;
    if (std::fclose(fp) != 0 || bad) return fail(CS_ERR_ARG, std::string("write to ") + path + " failed");
    return CS_OK;
}

int cs_tune_load(const char* path) {
    if (!path) return fail(CS_ERR_ARG, "null path");
    FILE* fp = std::fopen(path, "r");
    if (!fp) return fail(CS_ERR_ARG, std::string("cannot read ") + path);
    char buf[256];
    int rc = CS_OK, lineno = 0;
    while (!rc && std::fgets(buf, sizeof buf, fp)) {
        lineno++;
        char* hash = std::strchr(buf, '#');
        if (hash) *hash = '\0';
        char kind[16] = {0};
        int v[9], k = 0;
        if (std::sscanf(buf, "%15s", kind) != 1) continue;      // blank line
        if (!std::strcmp(kind, "line") || !std::strcmp(kind, "sweep") || !std::strcmp(kind, "filter")) {
            const char* rest = std::strstr(buf, kind) + std::strlen(kind);
            const bool sweep = kind[0] == 's';
            k = std::sscanf(rest, "%d %d %d %d %d %d %d %d %d", &v[0], &v[1], &v[2], &v[3], &v[4], &v[5], &v[6], &v[7],
                            &v[8]);
            if (sweep ? (k != 7 && k != 8) : k != 6) {
                rc = fail(CS_ERR_ARG, std::string(path) + ":" + std::to_string(lineno) + ": expected " +
                                          (sweep ? "7 or 8" : "6") + " integers after '" + kind + "'");
                break;
            }
            if (sweep && k == 7) {   // no metrics field: constant metrics
                for (int i = 7; i > 4; i--) v[i] = v[i - 1];
                v[4] = 0;
            }
            if (kind[0] == 'l') rc = cs_tune_set(v[0], v[1], v[2], v[3], v[4], v[5]);
            else if (sweep) rc = cs_tune_set_sweep(v[0], v[1], v[2], v[3], v[4], v[5], v[6], v[7]);
            else rc = cs_tune_set_filter_sweep(v[0], v[1], v[2], v[3], v[4], v[5]);
        } else if (std::sscanf(buf, "%d %d %d %d", &v[0], &v[1], &v[2], &v[3]) == 3) {
            // "<op> <length> <K>": FP64, closure and seam ends
            for (int e = 0; e < 2 && !rc; e++) rc = cs_tune_set(v[0], v[1], e, e, 0, v[2]);
        } else {
            rc = fail(CS_ERR_ARG, std::string(path) + ":" + std::to_string(lineno) + ": unknown entry");
        }
        if (rc) {
            std::string msg = cs_last_error();
            if (msg.find(path) != 0) rc = fail(rc, std::string(path) + ":" + std::to_string(lineno) + ": " + msg);
        }
    }
    std::fclose(fp);
    return rc;
}

}  // extern "C"
