// capi.hip -- the extern "C" boundary: contexts and the operator API of the
// reference's L5 layer (numerics_base / numerics_pc / physics / gcbc_go) on
// one GPU subdomain.  Orchestration follows the reference call order
// (src/physics_pc.hpp:83-256, src/gcbc.hpp:86-228); the per-variable host
// stream synchronisation of mpigo does not exist here because the seam sums
// of one subdomain are formed on the device.
#include "physics_kernels.hpp"

#include <cmath>
#include <cstring>

namespace cs {

static thread_local std::string g_last_error;

void set_error(const std::string& msg) { g_last_error = msg; }

int fail(int code, const std::string& msg) {
    g_last_error = msg;
    return code;
}

int check_dims(const cs_dims* d, Geo* g) {
    if (!d) return fail(CS_ERR_ARG, "null dims");
    if (d->lxi < 1 || d->let < 1 || d->lze < 1) return fail(CS_ERR_SHAPE, "dimensions must be positive");
    long lmx = (long)d->lxi * d->let * d->lze;
    if (lmx * 15 >= (1L << 40)) return fail(CS_ERR_SHAPE, "grid too large");
    g->lxi = d->lxi;
    g->let = d->let;
    g->lze = d->lze;
    g->lmx = lmx;
    return CS_OK;
}

static int ndf_of(int code) {
    switch (code) {
        case CS_BC_NON_REFLECTIVE: case CS_BC_WALL_INVISCID: case CS_BC_WALL_VISCOUS: case CS_BC_INTER_CURV:
            return 0;
        case CS_BC_INTER_STRAIGHT: case CS_BC_INTER_SUBDOMAINS: case CS_BC_PERIODIC:
            return 1;
        default:
            return -1;
    }
}

static bool is_char_bc(int code) {
    return code == CS_BC_NON_REFLECTIVE || code == CS_BC_WALL_INVISCID || code == CS_BC_WALL_VISCOUS ||
           code == CS_BC_INTER_CURV;
}

// interface table in FP64: numerics_base::deriv_setup, src/numerics_base.hpp:182-297
// (Gauss-Jordan inverse of the 16x16 model P, rows 8 and 9 of R P^-1 Q).
static void interface_table(double* pbco, double* pbci) {
    const int n = 16, lmd = kLmd;
    double Q[16][16] = {}, Pm[16][16] = {}, inv[16][16], tmp[16][16], rs[16][16];
    Q[0][0] = kAb00; Q[0][1] = kAb01; Q[0][2] = kAb02;
    Q[1][0] = -kAb10; Q[1][2] = kAb10;
    Q[n - 1][n - 1] = -kAb02; Q[n - 1][n - 2] = -kAb01; Q[n - 1][n - 3] = -kAb00;   // as shipped (D1)
    Q[n - 2][n - 1] = kAb10; Q[n - 2][n - 3] = -kAb10;
    for (int i = 2; i < n - 2; i++) { Q[i][i - 1] = -kAa; Q[i][i + 1] = kAa; Q[i][i - 2] = -kBb; Q[i][i + 2] = kBb; }
    Pm[0][0] = 1.0; Pm[0][1] = kAlpha01;
    Pm[n - 1][n - 1] = 1.0; Pm[n - 1][n - 2] = kAlpha01;
    Pm[1][0] = kAlpha10; Pm[1][1] = 1.0; Pm[1][2] = kAlpha10;
    Pm[n - 2][n - 1] = kAlpha10; Pm[n - 2][n - 2] = 1.0; Pm[n - 2][n - 3] = kAlpha10;
    for (int i = 2; i < n - 2; i++) { Pm[i][i] = 1.0; Pm[i][i - 1] = kAlpha; Pm[i][i + 1] = kAlpha; }
    // Gauss-Jordan with partial pivoting on an augmented copy
    double A[16][32];
    for (int i = 0; i < n; i++)
        for (int j = 0; j < 2 * n; j++) A[i][j] = (j < n) ? Pm[i][j] : (j - n == i ? 1.0 : 0.0);
    for (int k = 0; k < n; k++) {
        int p = k;
        for (int i = k + 1; i < n; i++) if (std::fabs(A[i][k]) > std::fabs(A[p][k])) p = i;
        if (p != k) for (int j = 0; j < 2 * n; j++) std::swap(A[k][j], A[p][j]);
        double d = 1.0 / A[k][k];
        for (int j = 0; j < 2 * n; j++) A[k][j] *= d;
        for (int i = 0; i < n; i++) {
            if (i == k || A[i][k] == 0.0) continue;
            double f = A[i][k];
            for (int j = 0; j < 2 * n; j++) A[i][j] -= f * A[k][j];
        }
    }
    for (int i = 0; i < n; i++) for (int j = 0; j < n; j++) inv[i][j] = A[i][j + n];
    // R: P without the coupling between rows 7 and 8
    Pm[7][8] = 0.0; Pm[8][7] = 0.0;
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) {
            double s = 0.0;
            for (int k = 0; k < n; k++) s += inv[i][k] * Q[k][j];
            tmp[i][j] = s;
        }
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) {
            double s = 0.0;
            for (int k = 0; k < n; k++) s += Pm[i][k] * tmp[k][j];
            rs[i][j] = s;
        }
    for (int t = 0; t < lmd; t++) {
        pbco[lmd - 1 - t] = rs[8][t];
        pbci[t] = rs[8][lmd + t];
        pbco[2 * lmd - 1 - t] = rs[9][t];
        pbci[lmd + t] = rs[9][lmd + t];
    }
}

static int seam_pair(const cs_numerics* num, int axis) {
    // 1 = both faces seam (single-GPU wrap), 0 = both closures, -1 = mixed
    int a = num->ndf[0][axis], b = num->ndf[1][axis];
    if (a == 1 && b == 1) return 1;
    if (a == 0 && b == 0) return 0;
    return -1;
}

}  // namespace cs

using namespace cs;

extern "C" {

const char* cs_last_error(void) { return g_last_error.c_str(); }
const char* cs_version(void) { return "compact_schemes_amd 0.1 (gfx950)"; }

int cs_numerics_create(const cs_dims* dims, const int* nbc, cs_numerics** out) {
    if (!out || !nbc) return fail(CS_ERR_ARG, "null argument");
    *out = nullptr;
    Geo g;
    int rc = check_dims(dims, &g);
    if (rc) return rc;
    cs_numerics* num = new cs_numerics();
    num->g = g;
    for (int ip = 0; ip < 2; ip++)
        for (int nn = 0; nn < 3; nn++) {
            num->nbc[ip][nn] = nbc[ip * 3 + nn];
            num->ndf[ip][nn] = ndf_of(num->nbc[ip][nn]);
            if (num->ndf[ip][nn] < 0) {
                delete num;
                return fail(CS_ERR_ARG, "unknown boundary-condition code " + std::to_string(nbc[ip * 3 + nn]));
            }
        }
    (void)hipGetDevice(&num->device);
    for (int a = 0; a < 3; a++) {
        long fs = g.face_size(a);
        if (hipMalloc(&num->drva[a], sizeof(double) * 2 * kNV * fs) != hipSuccess ||
            hipMalloc(&num->send[a], sizeof(double) * 4 * kNV * fs) != hipSuccess ||
            hipMalloc(&num->recv[a], sizeof(double) * 4 * kNV * fs) != hipSuccess) {
            cs_numerics_destroy(num);
            return fail(CS_ERR_HIP, "hipMalloc of face buffers failed");
        }
        (void)hipMemset(num->drva[a], 0, sizeof(double) * 2 * kNV * fs);
        (void)hipMemset(num->send[a], 0, sizeof(double) * 4 * kNV * fs);
        (void)hipMemset(num->recv[a], 0, sizeof(double) * 4 * kNV * fs);
    }
    if (hipMalloc(&num->pb, sizeof(double) * 32) != hipSuccess) {
        cs_numerics_destroy(num);
        return fail(CS_ERR_HIP, "hipMalloc of the interface table failed");
    }
    (void)hipMemset(num->pb, 0, sizeof(double) * 32);
    std::memset(num->pb_host, 0, sizeof num->pb_host);
    *out = num;
    return CS_OK;
}

int cs_numerics_destroy(cs_numerics* num) {
    if (!num) return CS_OK;
    for (int a = 0; a < 3; a++) {
        if (num->drva[a]) (void)hipFree(num->drva[a]);
        if (num->send[a]) (void)hipFree(num->send[a]);
        if (num->recv[a]) (void)hipFree(num->recv[a]);
    }
    if (num->pb) (void)hipFree(num->pb);
    if (num->work) (void)hipFree(num->work);
    for (auto& kv : num->ops) free_line_op(kv.second);
    delete num;
    return CS_OK;
}

int cs_set_interface_table(cs_numerics* num, const double* pbco, const double* pbci) {
    if (!num || !pbco || !pbci) return fail(CS_ERR_ARG, "null argument");
    std::memcpy(num->pb_host, pbco, 16 * sizeof(double));
    std::memcpy(num->pb_host + 16, pbci, 16 * sizeof(double));
    CS_HIP(hipMemcpy(num->pb, num->pb_host, sizeof num->pb_host, hipMemcpyHostToDevice));
    num->table_set = true;
    // transposed derivative tables hold the interface coefficients: rebuild on use
    for (auto it = num->ops.begin(); it != num->ops.end();) {
        if (it->first.op == OP_DERIV_T) { free_line_op(it->second); it = num->ops.erase(it); }
        else ++it;
    }
    return CS_OK;
}

int cs_deriv_setup(cs_numerics* num) {
    if (!num) return fail(CS_ERR_ARG, "null context");
    double pbco[16], pbci[16];
    interface_table(pbco, pbci);
    return cs_set_interface_table(num, pbco, pbci);
}

int cs_get_interface_table(const cs_numerics* num, double* pbco, double* pbci) {
    if (!num || !pbco || !pbci) return fail(CS_ERR_ARG, "null argument");
    std::memcpy(pbco, num->pb_host, 16 * sizeof(double));
    std::memcpy(pbci, num->pb_host + 16, 16 * sizeof(double));
    return CS_OK;
}

int cs_numerics_buffers(cs_numerics* num, int axis, double** drva, double** send, double** recv) {
    if (!num || axis < 0 || axis > 2) return fail(CS_ERR_ARG, "bad context or axis");
    if (drva) *drva = num->drva[axis];
    if (send) *send = num->send[axis];
    if (recv) *recv = num->recv[axis];
    return CS_OK;
}

int cs_numerics_ndf(const cs_numerics* num, int* ndf6) {
    if (!num || !ndf6) return fail(CS_ERR_ARG, "null argument");
    for (int ip = 0; ip < 2; ip++)
        for (int nn = 0; nn < 3; nn++) ndf6[ip * 3 + nn] = num->ndf[ip][nn];
    return CS_OK;
}

int cs_set_filter_alpha(cs_numerics* num, double af) {
    if (!num) return fail(CS_ERR_ARG, "null context");
    if (!(af > -0.5 && af < 0.5)) return fail(CS_ERR_ARG, "filter alpha must lie in (-0.5, 0.5)");
    num->alpha_f = af;
    for (auto it = num->ops.begin(); it != num->ops.end();) {
        if (it->first.op == OP_FILTER || it->first.op == OP_FILTER_T) { free_line_op(it->second); it = num->ops.erase(it); }
        else ++it;
    }
    return CS_OK;
}

int cs_set_precision(cs_numerics* num, int mode) {
    if (!num) return fail(CS_ERR_ARG, "null context");
    if (mode < 0 || mode > 4) return fail(CS_ERR_ARG, "precision mode is 0 (FP64) or 1..4 (FP32 + that many FP64 refinement sweeps)");
    num->precision = mode;
    return CS_OK;
}

int cs_fill_buffers(cs_numerics* num, const double* infield, int lo, int hi, int nrt, cs_stream stream) {
    if (!num || !infield) return fail(CS_ERR_ARG, "null argument");
    if (lo < 0 || hi > kNV || lo > hi) return fail(CS_ERR_ARG, "variable range out of 0..5");
    if (!num->table_set) return fail(CS_ERR_STATE, "fill_buffers needs deriv_setup first");
    hipStream_t s = (hipStream_t)stream;
    const Geo& g = num->g;
    for (int m = lo; m < hi; m++)
        for (int nn = 0; nn < 3; nn++) {
            int nzk = (1 - nrt) * nn;
            const double* src = infield + (long)nzk * g.lmx + (long)m * g.lmx * 3;
            long fs = g.face_size(nn);
            double* send = num->send[nn] + (long)m * 4 * fs;
            for (int ip = 0; ip < 2; ip++) {
                if (num->ndf[ip][nn] != 1) continue;
                if (g.n(nn) < kLmd) return fail(CS_ERR_SHAPE, "seam face needs at least 8 points");
                int rc = launch_fill(src, send + (long)ip * 2 * fs, num->pb, nn, ip, g.n(nn), fs, g.lxi,
                                     (long)g.lxi * g.let, g.stride(nn), s);
                if (rc) return rc;
            }
        }
    return CS_OK;
}

int cs_mpigo(cs_numerics* num, int var, cs_stream stream) {
    if (!num || var < 0 || var >= kNV) return fail(CS_ERR_ARG, "bad context or variable");
    hipStream_t s = (hipStream_t)stream;
    for (int a = 0; a < 3; a++) {
        long fs = num->g.face_size(a);
        int sp = seam_pair(num, a);
        if (sp < 0) return fail(CS_ERR_UNSUPPORTED, "axis with one seam face needs a neighbour subdomain");
        if (sp == 0) continue;
        double* sb = num->send[a] + (long)var * 4 * fs;
        double* rb = num->recv[a] + (long)var * 4 * fs;
        for (int ip = 0; ip < 2; ip++)
            CS_HIP(hipMemcpyAsync(rb + (long)ip * 2 * fs, sb + (long)(1 - ip) * 2 * fs, sizeof(double) * 2 * fs,
                                  hipMemcpyDeviceToDevice, s));
    }
    return CS_OK;
}

int cs_exchange(cs_numerics* a, int face_a, cs_numerics* b, int face_b, int axis, int var, cs_stream stream) {
    if (!a || !b || axis < 0 || axis > 2 || var < 0 || var >= kNV) return fail(CS_ERR_ARG, "bad context, axis or variable");
    if ((face_a != 0 && face_a != 1) || (face_b != 0 && face_b != 1)) return fail(CS_ERR_ARG, "face is 0 or 1");
    if (a->ndf[face_a][axis] != 1 || b->ndf[face_b][axis] != 1)
        return fail(CS_ERR_STATE, "both faces must be seam faces (BC 35, 40 or 45)");
    const long fs = a->g.face_size(axis);
    if (b->g.face_size(axis) != fs || (axis != 0 && a->g.lxi != b->g.lxi) || (axis != 1 && a->g.let != b->g.let) ||
        (axis != 2 && a->g.lze != b->g.lze))
        return fail(CS_ERR_SHAPE, "the two faces do not match point for point");
    if (a == b && face_a == face_b) return fail(CS_ERR_ARG, "a face cannot adjoin itself");
    hipStream_t s = (hipStream_t)stream;
    const size_t bytes = sizeof(double) * 2 * fs;
    double* sa = a->send[axis] + ((long)var * 2 + face_a) * 2 * fs;
    double* ra = a->recv[axis] + ((long)var * 2 + face_a) * 2 * fs;
    double* sb = b->send[axis] + ((long)var * 2 + face_b) * 2 * fs;
    double* rb = b->recv[axis] + ((long)var * 2 + face_b) * 2 * fs;
    CS_HIP(hipMemcpyAsync(rb, sa, bytes, hipMemcpyDeviceToDevice, s));
    CS_HIP(hipMemcpyAsync(ra, sb, bytes, hipMemcpyDeviceToDevice, s));
    return CS_OK;
}

int cs_deriv1d(cs_numerics* num, int axis, const double* in, double* out, int nstart, int nend, int var,
               cs_stream stream) {
    if (!num || axis < 0 || axis > 2) return fail(CS_ERR_ARG, "bad context or axis");
    if (var < 0 || var >= kNV) return fail(CS_ERR_ARG, "variable out of range");
    int st = nstart != 0, en = nend != 0;
    const double* recv = num->recv[axis] + (long)var * 4 * num->g.face_size(axis);
    return line_apply(num, OP_DERIV, axis, in, out, st, en, (st || en) ? recv : nullptr, var, true,
                      (hipStream_t)stream);
}

int cs_deriv2d(cs_numerics* num, int axis, const double* in, double* out, int nstart, int nend, int var, int comp,
               cs_stream stream) {
    if (!num || !in || !out) return fail(CS_ERR_ARG, "null argument");
    if (comp < 0 || comp > 2 || axis < 0 || axis > 2) return fail(CS_ERR_ARG, "component/axis out of range");
    return cs_deriv1d(num, axis, in + (long)comp * num->g.lmx, out + (long)axis * num->g.lmx, nstart, nend, var,
                      stream);
}

int cs_deriv(cs_numerics* num, int axis, const double* in, double* out, int var, cs_stream stream) {
    if (!num || axis < 0 || axis > 2) return fail(CS_ERR_ARG, "bad context or axis");
    if (var < 0 || var >= kNV) return fail(CS_ERR_ARG, "variable out of range");
    int sp = seam_pair(num, axis);
    if (sp < 0) return fail(CS_ERR_UNSUPPORTED, "axis with one seam face needs a neighbour subdomain");
    return line_apply(num, OP_DERIV, axis, in, out, sp, sp, nullptr, var, true, (hipStream_t)stream);
}

int cs_filter(cs_numerics* num, int axis, const double* in, double* out, cs_stream stream) {
    if (!num || axis < 0 || axis > 2) return fail(CS_ERR_ARG, "bad context or axis");
    int sp = seam_pair(num, axis);
    if (sp < 0) return fail(CS_ERR_UNSUPPORTED, "axis with one seam face needs a neighbour subdomain");
    return line_apply(num, OP_FILTER, axis, in, out, sp, sp, nullptr, 0, false, (hipStream_t)stream);
}

// ---- physics ---------------------------------------------------------------

int cs_physics_create(const cs_dims* dims, int viscous, cs_point umf, cs_physics** out) {
    if (!out) return fail(CS_ERR_ARG, "null argument");
    *out = nullptr;
    Geo g;
    int rc = check_dims(dims, &g);
    if (rc) return rc;
    cs_physics* ph = new cs_physics();
    ph->g = g;
    ph->viscous = viscous ? 1 : 0;
    ph->umf = umf;
    long lmx = g.lmx;
    ph->red_blocks = time_step_blocks(lmx);
    bool ok = hipMalloc(&ph->buffer, sizeof(double) * 15 * lmx) == hipSuccess &&
              hipMalloc(&ph->buffer_deriv, sizeof(double) * 15 * lmx) == hipSuccess &&
              hipMalloc(&ph->red, sizeof(double) * (2 * ph->red_blocks + 2)) == hipSuccess &&
              hipHostMalloc(&ph->red_host, sizeof(double) * 2) == hipSuccess;
    if (ok && ph->viscous)
        ok = hipMalloc(&ph->ss, sizeof(double) * 3 * lmx) == hipSuccess &&
             hipMalloc(&ph->stress, sizeof(double) * 6 * lmx) == hipSuccess &&
             hipMalloc(&ph->heat, sizeof(double) * 3 * lmx) == hipSuccess;
    if (!ok) {
        cs_physics_destroy(ph);
        return fail(CS_ERR_HIP, "allocation of physics buffers failed");
    }
    if (ph->viscous) {
        (void)hipMemset(ph->stress, 0, sizeof(double) * 6 * lmx);
        (void)hipMemset(ph->heat, 0, sizeof(double) * 3 * lmx);
    }
    *out = ph;
    return CS_OK;
}

int cs_physics_destroy(cs_physics* ph) {
    if (!ph) return CS_OK;
    if (ph->buffer) (void)hipFree(ph->buffer);
    if (ph->buffer_deriv) (void)hipFree(ph->buffer_deriv);
    if (ph->ss) (void)hipFree(ph->ss);
    if (ph->stress) (void)hipFree(ph->stress);
    if (ph->heat) (void)hipFree(ph->heat);
    if (ph->red) (void)hipFree(ph->red);
    if (ph->red_host) (void)hipHostFree(ph->red_host);
    delete ph;
    return CS_OK;
}

int cs_physics_set_frame(cs_physics* ph, cs_point umf, cs_point dudtmf) {
    if (!ph) return fail(CS_ERR_ARG, "null context");
    ph->umf = umf;
    ph->dudtmf = dudtmf;
    return CS_OK;
}

int cs_physics_buffers(cs_physics* ph, double** stress, double** heat, double** buffer, double** buffer_deriv) {
    if (!ph) return fail(CS_ERR_ARG, "null context");
    if (stress) *stress = ph->stress;
    if (heat) *heat = ph->heat;
    if (buffer) *buffer = ph->buffer;
    if (buffer_deriv) *buffer_deriv = ph->buffer_deriv;
    return CS_OK;
}

int cs_primitives(const cs_dims* dims, const double* qa, double* de, double* p, cs_stream stream) {
    Geo g;
    int rc = check_dims(dims, &g);
    if (rc) return rc;
    if (!qa || !de || !p) return fail(CS_ERR_ARG, "null field pointer");
    return launch_primitives(qa, de, p, g.lmx, (hipStream_t)stream);
}

static int same_grid(const cs_physics* ph, const cs_numerics* num) {
    if (!ph || !num) return fail(CS_ERR_ARG, "null context");
    if (ph->g.lxi != num->g.lxi || ph->g.let != num->g.let || ph->g.lze != num->g.lze)
        return fail(CS_ERR_ARG, "physics and numerics contexts have different grids");
    return CS_OK;
}

int cs_calc_fluxes(cs_physics* ph, cs_numerics* num, const double* qa, const double* p, double* de,
                   const double* xim, const double* etm, const double* zem, cs_stream stream) {
    int rc = same_grid(ph, num);
    if (rc) return rc;
    if (!qa || !p || !de || !xim || !etm || !zem) return fail(CS_ERR_ARG, "null field pointer");
    hipStream_t s = (hipStream_t)stream;
    const long lmx = ph->g.lmx;
    rc = launch_fluxes_pre(ph->viscous, ph->buffer, qa, p, de, xim, etm, zem, ph->stress, ph->heat, ph->umf, lmx, s);
    if (rc) return rc;
    for (int m = 0; m < kNV; m++) {
        for (int a = 0; a < 3; a++) {
            int sp = seam_pair(num, a);
            if (sp < 0) return fail(CS_ERR_UNSUPPORTED, "axis with one seam face needs a neighbour subdomain");
            const long off = (long)m * 3 * lmx + (long)a * lmx;
            rc = line_apply(num, OP_DERIV, a, ph->buffer + off, ph->buffer_deriv + off, sp, sp, nullptr, m, true, s);
            if (rc) return rc;
        }
        rc = launch_fluxes_post(de, ph->buffer_deriv, m, lmx, s);
        if (rc) return rc;
    }
    return CS_OK;
}

int cs_calc_viscous_shear_stress(cs_physics* ph, cs_numerics* num, double* de, const double* ssk, const double* xim,
                                 const double* etm, const double* zem, const double* yaco, cs_stream stream) {
    int rc = same_grid(ph, num);
    if (rc) return rc;
    if (!ph->viscous) return fail(CS_ERR_STATE, "physics context is inviscid");
    if (!de || !ssk || !xim || !etm || !zem || !yaco) return fail(CS_ERR_ARG, "null field pointer");
    hipStream_t s = (hipStream_t)stream;
    const long lmx = ph->g.lmx;
    rc = launch_visc_init(ph->buffer, de, lmx, s);
    if (rc) return rc;
    for (int m = 1; m < kNV; m++) {
        for (int a = 0; a < 3; a++) {
            int sp = seam_pair(num, a);
            if (sp < 0) return fail(CS_ERR_UNSUPPORTED, "axis with one seam face needs a neighbour subdomain");
            rc = line_apply(num, OP_DERIV, a, ph->buffer + (long)m * 3 * lmx,
                            ph->buffer_deriv + (long)m * 3 * lmx + (long)a * lmx, sp, sp, nullptr, m, true, s);
            if (rc) return rc;
        }
        rc = launch_visc_post(ph->stress, ph->heat, ph->ss, ph->buffer_deriv, xim, etm, zem, m, lmx, s);
        if (rc) return rc;
    }
    return launch_visc_final(de, ssk, yaco, ph->ss, ph->stress, ph->heat, lmx, s);
}

int cs_calc_time_step(cs_physics* ph, const double* xim, const double* etm, const double* zem, const double* de,
                      const double* yaco, const double* ssk, double cfl, double* dte, double* max2,
                      cs_stream stream) {
    if (!ph || !dte) return fail(CS_ERR_ARG, "null argument");
    if (!xim || !etm || !zem || !de || !yaco || !ssk) return fail(CS_ERR_ARG, "null field pointer");
    hipStream_t s = (hipStream_t)stream;
    int rc = launch_time_step(xim, etm, zem, de, yaco, ssk, ph->umf, ph->red, ph->red_blocks, ph->g.lmx, s);
    if (rc) return rc;
    CS_HIP(hipMemcpyAsync(ph->red_host, ph->red + 2 * ph->red_blocks, 2 * sizeof(double), hipMemcpyDeviceToHost, s));
    CS_HIP(hipStreamSynchronize(s));
    double m0 = ph->red_host[0], m1 = ph->red_host[1];
    if (max2) { max2[0] = m0; max2[1] = m1; }
    *dte = std::fmin(cfl / m0, 0.5 / m1);
    if (!std::isfinite(*dte) || *dte <= 0.0) return fail(CS_ERR_ARG, "time step is not a positive finite number");
    return CS_OK;
}

int cs_gcbc_go(cs_numerics* num, const double* const* cm, const double* const* drvb, const double* qa, double* de,
               const double* p, const double* yaco, const double* sbcc, cs_point umf, cs_point dudtmf, double dtwi,
               cs_stream stream) {
    if (!num || !cm || !qa || !de || !p || !yaco || !sbcc) return fail(CS_ERR_ARG, "null argument");
    hipStream_t s = (hipStream_t)stream;
    for (int nn = 0; nn < 3; nn++)
        for (int ip = 0; ip < 2; ip++) {
            int np = num->nbc[ip][nn];
            if (!is_char_bc(np)) continue;
            if (!cm[nn]) return fail(CS_ERR_ARG, "missing face normals for a characteristic face");
            if (np == CS_BC_INTER_CURV && (!drvb || !drvb[nn]))
                return fail(CS_ERR_UNSUPPORTED, "inter-curv face needs the neighbour's characteristics (drvb)");
        }
    // setup on every characteristic face (src/gcbc.hpp:101-131)
    for (int nn = 0; nn < 3; nn++)
        for (int ip = 0; ip < 2; ip++) {
            int np = num->nbc[ip][nn];
            if (!is_char_bc(np)) continue;
            int flag = (CS_BC_WALL_INVISCID - np) * (CS_BC_WALL_VISCOUS - np) * (CS_BC_INTER_CURV - np) / 3000;
            int rc = launch_gcbc_setup(num, nn, ip, flag, cm[nn], qa, de, p, yaco, umf, s);
            if (rc) return rc;
        }
    // updates (src/gcbc.hpp:160-185)
    for (int nn = 0; nn < 3; nn++)
        for (int ip = 0; ip < 2; ip++) {
            int np = num->nbc[ip][nn];
            if (!is_char_bc(np)) continue;
            int rc = launch_gcbc_update(num, nn, ip, np, cm[nn], drvb ? drvb[nn] : nullptr, qa, de, p, sbcc, yaco,
                                        umf, dudtmf, dtwi, s);
            if (rc) return rc;
        }
    return CS_OK;
}

int cs_wall_condition_update(cs_numerics* num, double* qa, const int* npex, cs_point umf, cs_stream stream) {
    if (!num || !qa || !npex) return fail(CS_ERR_ARG, "null argument");
    for (int nn = 0; nn < 3; nn++)
        for (int ip = 0; ip < 2; ip++) {
            int np = num->nbc[ip][nn];
            if (np != CS_BC_WALL_INVISCID && np != CS_BC_WALL_VISCOUS) continue;
            int rc = launch_wall(num, nn, ip, np, qa, npex, umf, (hipStream_t)stream);
            if (rc) return rc;
        }
    return CS_OK;
}

int cs_rk_stage(const cs_dims* dims, double* q, const double* q0, const double* de, const double* yaco, double dtk,
                cs_stream stream) {
    Geo g;
    int rc = check_dims(dims, &g);
    if (rc) return rc;
    if (!q || !q0 || !de || !yaco) return fail(CS_ERR_ARG, "null field pointer");
    return launch_rk_stage(q, q0, de, yaco, dtk, g.lmx, (hipStream_t)stream);
}

}  // extern "C"
