/* compact_schemes.hpp -- header-only C++ operator classes shaped like the
 * reference's host API, implemented on the C-ABI of compact_schemes.h.
 *
 * A program written against the reference's classes
 *   numerics_base<T> / numerics_pc<T>   (src/numerics_base.hpp, src/numerics_pc.hpp)
 *   physics<Visc, T>                     (src/physics_base.hpp, src/physics_pc.hpp)
 *   gcbc<T, I>, gcbc_go, wall_condition_update   (src/gcbc.hpp)
 * compiles against this header with T = double (the library computes in FP64).
 * Member functions keep the reference's names, argument order and meaning.
 * Arguments that only matter across MPI ranks (mcd, itag, the sync functor)
 * are accepted and ignored: a subdomain's seam faces wrap onto itself.
 * Streams are HIP streams passed as cs_stream (void*); the reference's array
 * of five CUDA streams becomes an array whose first entry is used (NULL: the
 * default stream).
 *
 * Errors are fail-fast as in the reference (src/cuda/check.hpp:36-41): a
 * non-zero status prints cs_last_error() and exits with EXIT_FAILURE.
 */
#ifndef COMPACT_SCHEMES_HPP
#define COMPACT_SCHEMES_HPP

#include <cstdio>
#include <cstdlib>
#include <map>
#include <mutex>
#include <type_traits>
#include <vector>

#include <hip/hip_runtime_api.h>

#include "compact_schemes.h"

/* t_dcomp, src/common/data_types.hpp:33-39 (lmx = lxi*let*lze) */
struct t_dcomp {
    int lmx;
    int lxi;
    int let;
    int lze;
};

/* t_point, src/common/data_types.hpp:41-47 */
template <typename Type>
struct t_point {
    Type x;
    Type y;
    Type z;
};

/* BC codes, src/common/parameters.hpp */
constexpr int BC_NON_REFLECTIVE = CS_BC_NON_REFLECTIVE;
constexpr int BC_WALL_INVISCID = CS_BC_WALL_INVISCID;
constexpr int BC_WALL_VISCOUS = CS_BC_WALL_VISCOUS;
constexpr int BC_INTER_CURV = CS_BC_INTER_CURV;
constexpr int BC_INTER_STRAIGHT = CS_BC_INTER_STRAIGHT;
constexpr int BC_INTER_SUBDOMAINS = CS_BC_INTER_SUBDOMAINS;
constexpr int BC_PERIODIC = CS_BC_PERIODIC;
constexpr unsigned int NumberOfVariables = 5;
constexpr unsigned int NumberOfSpatialDims = 3;
constexpr unsigned int NumberOfFaces = 2;

inline void check_cs(int status) {
    if (status != CS_OK) {
        std::fprintf(stderr, "compact_schemes error %d: %s\n", status, cs_last_error());
        std::exit(EXIT_FAILURE);
    }
}

namespace cs_detail {
inline cs_dims dims_of(const t_dcomp& d) { return cs_dims{d.lxi, d.let, d.lze}; }

/* live numerics contexts by their x-axis drva pointer: the reference's free
 * functions gcbc_go / wall_condition_update receive buffers, not objects */
inline std::map<const void*, cs_numerics*>& registry() {
    static std::map<const void*, cs_numerics*> m;
    return m;
}
inline std::mutex& registry_lock() {
    static std::mutex m;
    return m;
}
inline void enroll(const void* key, cs_numerics* h) {
    std::lock_guard<std::mutex> g(registry_lock());
    registry()[key] = h;
}
inline void withdraw(const void* key) {
    std::lock_guard<std::mutex> g(registry_lock());
    registry().erase(key);
}
inline cs_numerics* lookup(const void* key) {
    std::lock_guard<std::mutex> g(registry_lock());
    auto it = registry().find(key);
    return it == registry().end() ? nullptr : it->second;
}
inline cs_stream first(cs_stream* streams) { return streams ? streams[0] : nullptr; }
template <typename Type>
inline cs_point point_of(const t_point<Type>& p) { return cs_point{(double)p.x, (double)p.y, (double)p.z}; }
}  // namespace cs_detail

/* numerics_base<Type>, src/numerics_base.hpp:40-306 */
template <typename Type>
struct numerics_base {
    static_assert(std::is_same<Type, double>::value, "the HIP library computes in FP64");

    numerics_base(t_dcomp dcomp_info, int nbc[2][3]) {
        cs_dims d = cs_detail::dims_of(dcomp_info);
        check_cs(cs_numerics_create(&d, &nbc[0][0], &handle));
        for (int a = 0; a < 3; a++)
            check_cs(cs_numerics_buffers(handle, a, &drva_buffer[a], &send_buffer[a], &recv_buffer[a]));
        int n6[6];
        check_cs(cs_numerics_ndf(handle, n6));
        for (int ip = 0; ip < 2; ip++)
            for (int nn = 0; nn < 3; nn++) ndf[ip][nn] = n6[ip * 3 + nn];
        cs_detail::enroll(drva_buffer[0], handle);
    }
    ~numerics_base() {
        cs_detail::withdraw(drva_buffer[0]);
        cs_numerics_destroy(handle);
    }
    numerics_base(const numerics_base&) = delete;
    numerics_base& operator=(const numerics_base&) = delete;

    /* numerics_base::deriv_setup, src/numerics_base.hpp:182-297 */
    void deriv_setup() { check_cs(cs_deriv_setup(handle)); }

    /* numerics_base::mpigo (2D), src/numerics_base.hpp:114-180 */
    template <typename StreamType, typename SyncFunction>
    void mpigo(t_dcomp, int[2][3], int, unsigned int variable_id, StreamType* stream, SyncFunction) {
        check_cs(cs_mpigo(handle, (int)variable_id, stream ? (cs_stream)stream[0] : nullptr));
    }

    /* interface coefficients (host copies of pbco/pbci) */
    void interface_table(double pbco16[16], double pbci16[16]) const {
        check_cs(cs_get_interface_table(handle, pbco16, pbci16));
    }

    Type* drva_buffer[3];
    Type* send_buffer[3];
    Type* recv_buffer[3];
    int ndf[2][3];
    cs_numerics* handle = nullptr;
};

/* numerics_pc<Type>, src/numerics_pc.hpp:42-288 */
template <typename Type>
struct numerics_pc : public numerics_base<Type> {
    using Base = numerics_base<Type>;
    using Base::handle;

    numerics_pc(t_dcomp dcomp_info, int nbc[2][3]) : Base(dcomp_info, nbc) {}

    /* numerics_pc::deriv1d<Axis>, src/numerics_pc.hpp:74-119 */
    template <unsigned int Axis>
    void deriv1d(Type* infield, Type* outfield, int nstart, int nend, t_dcomp, unsigned int variable_id) {
        static_assert(Axis < 3, "Axis index must be 0, 1, or 2");
        check_cs(cs_deriv1d(handle, (int)Axis, infield, outfield, nstart, nend, (int)variable_id, nullptr));
    }

    /* numerics_pc::deriv2d<Axis>, src/numerics_pc.hpp:121-168 */
    template <unsigned int Axis>
    void deriv2d(Type* infield, Type* outfield, int nstart, int nend, t_dcomp, unsigned int variable_id,
                 unsigned int component_id, cs_stream* stream) {
        static_assert(Axis < 3, "Axis index must be 0, 1, or 2");
        check_cs(cs_deriv2d(handle, (int)Axis, infield, outfield, nstart, nend, (int)variable_id,
                            (int)component_id, cs_detail::first(stream)));
    }

    /* numerics_pc::fill_buffers<lo, hi>, src/numerics_pc.hpp:231-287 */
    template <unsigned int lo, unsigned int hi>
    void fill_buffers(Type* infield, int nrt, t_dcomp, cs_stream* streams) {
        check_cs(cs_fill_buffers(handle, infield, (int)lo, (int)hi, nrt, cs_detail::first(streams)));
    }

    /* fill + exchange + deriv1d of one field in one launch (no counterpart) */
    template <unsigned int Axis>
    void deriv(Type* infield, Type* outfield, unsigned int variable_id, cs_stream stream = nullptr) {
        check_cs(cs_deriv(handle, (int)Axis, infield, outfield, (int)variable_id, stream));
    }

    /* compact filter along Axis (absent upstream) */
    template <unsigned int Axis>
    void filter(Type* infield, Type* outfield, cs_stream stream = nullptr) {
        check_cs(cs_filter(handle, (int)Axis, infield, outfield, stream));
    }
    void set_filter_alpha(double alpha_f) { check_cs(cs_set_filter_alpha(handle, alpha_f)); }
};

/* physics<EnableViscous, Type>: physics_base (src/physics_base.hpp:42-162)
 * and the precompiled physics (src/physics_pc.hpp:54-257) */
template <bool EnableViscous, typename Type>
struct physics {
    static_assert(std::is_same<Type, double>::value, "the HIP library computes in FP64");

    explicit physics(t_dcomp dcomp_info) {
        check_cs(cs_physics_create_dims(dcomp_info));
        double* st = nullptr;
        double* hf = nullptr;
        check_cs(cs_physics_buffers(handle, &st, &hf, &buffer, &buffer_deriv));
        for (int k = 0; k < 6; k++) stress[k] = st ? st + (long)k * dcomp_info.lmx : nullptr;
        for (int k = 0; k < 3; k++) heat[k] = hf ? hf + (long)k * dcomp_info.lmx : nullptr;
    }
    ~physics() { cs_physics_destroy(handle); }
    physics(const physics&) = delete;
    physics& operator=(const physics&) = delete;

    void init() {}

    /* physics_base::read_config sets umf from amach1..3
     * (src/physics_base.hpp:135-145); without a YAML parser the three Mach
     * numbers are passed directly */
    void set_mach(Type amach1, Type amach2, Type amach3) {
        umf.x = amach1;
        umf.y = amach2;
        umf.z = amach3;
    }

    /* physics_base::movef, src/physics_base.hpp:147-150 (empty upstream) */
    void movef(Type, Type, Type) {}

    /* physics::calc_fluxes, src/physics_pc.hpp:83-149 */
    void calc_fluxes(Type* qa, Type* pressure, Type* de, Type* xim, Type* etm, Type* zem, t_dcomp, int[2][3],
                     numerics_pc<Type>* numerics_instance, cs_stream* streams) {
        frame();
        check_cs(cs_calc_fluxes(handle, numerics_instance->handle, qa, pressure, de, xim, etm, zem,
                                cs_detail::first(streams)));
    }

    /* physics::calc_viscous_shear_stress, src/physics_pc.hpp:151-227 */
    void calc_viscous_shear_stress(Type* de, Type* ssk, Type* xim, Type* etm, Type* zem, Type* yaco, t_dcomp,
                                   int[2][3], numerics_pc<Type>* numerics_instance, cs_stream* streams) {
        static_assert(EnableViscous, "viscous stress needs physics<true, Type>");
        frame();
        check_cs(cs_calc_viscous_shear_stress(handle, numerics_instance->handle, de, ssk, xim, etm, zem, yaco,
                                              cs_detail::first(streams)));
    }

    /* physics::calc_time_step, src/physics_pc.hpp:229-256; dte is a host pointer */
    void calc_time_step(Type* xim, Type* etm, Type* zem, Type* de, Type* yaco, Type* ssk, Type cfl, Type* dte,
                        unsigned int) {
        frame();
        check_cs(cs_calc_time_step(handle, xim, etm, zem, de, yaco, ssk, cfl, dte, nullptr, nullptr));
    }

    t_point<Type> umf{0, 0, 0};
    t_point<Type> dudtmf{0, 0, 0};
    Type* buffer = nullptr;          /* [5 var][3 dir][lmx] */
    Type* buffer_deriv = nullptr;    /* [5 var][3 dir][lmx] */
    Type* stress[6];                 /* xx, yy, zz, xy, yz, zx (physics_base::stress_tensor) */
    Type* heat[3];                   /* xx, yy, zz (physics_base::heat_fluxes) */
    cs_physics* handle = nullptr;

private:
    int cs_physics_create_dims(const t_dcomp& d) {
        cs_dims cd = cs_detail::dims_of(d);
        return cs_physics_create(&cd, EnableViscous ? 1 : 0, cs_point{0.0, 0.0, 0.0}, &handle);
    }
    void frame() { check_cs(cs_physics_set_frame(handle, cs_detail::point_of(umf), cs_detail::point_of(dudtmf))); }
};

inline void check_hip(hipError_t e) {
    if (e != hipSuccess) {
        std::fprintf(stderr, "HIP error: %s\n", hipGetErrorString(e));
        std::exit(EXIT_FAILURE);
    }
}

/* gcbc<Type, TypeIndex>, src/gcbc.hpp:42-84: owns npex [lmx], drvb
 * [2][5][face] per axis and sbcc [lmx].  The reference allocates them and
 * never sets them; here npex = 0 (adiabatic walls) and init() sets sbcc = 1
 * on every boundary point of the subdomain, 0 inside. */
template <typename Type, typename TypeIndex>
struct gcbc {
    static_assert(std::is_same<Type, double>::value, "the HIP library computes in FP64");
    static_assert(std::is_same<TypeIndex, int>::value, "npex is int");

    explicit gcbc(t_dcomp dcomp_info) : d_(dcomp_info) {
        const size_t n = (size_t)dcomp_info.lmx;
        const size_t fs[3] = {(size_t)dcomp_info.let * dcomp_info.lze, (size_t)dcomp_info.lxi * dcomp_info.lze,
                              (size_t)dcomp_info.lxi * dcomp_info.let};
        check_hip(hipMalloc((void**)&npex, n * sizeof(TypeIndex)));
        check_hip(hipMemset(npex, 0, n * sizeof(TypeIndex)));
        for (int a = 0; a < 3; a++) {
            check_hip(hipMalloc((void**)&drvb[a], 2 * NumberOfVariables * fs[a] * sizeof(Type)));
            check_hip(hipMemset(drvb[a], 0, 2 * NumberOfVariables * fs[a] * sizeof(Type)));
        }
        check_hip(hipMalloc((void**)&sbcc, n * sizeof(Type)));
        check_hip(hipMemset(sbcc, 0, n * sizeof(Type)));
    }
    ~gcbc() {
        (void)hipFree(npex);
        for (int a = 0; a < 3; a++) (void)hipFree(drvb[a]);
        (void)hipFree(sbcc);
    }
    gcbc(const gcbc&) = delete;
    gcbc& operator=(const gcbc&) = delete;

    /* gcbc::init, src/gcbc.hpp:63-67 (empty upstream): sbcc = 1 on every
     * boundary point, 0 inside */
    void init() {
        const int lxi = d_.lxi, let = d_.let, lze = d_.lze;
        std::vector<Type> h((size_t)d_.lmx, Type(0));
        for (int k = 0; k < lze; k++)
            for (int j = 0; j < let; j++)
                for (int i = 0; i < lxi; i++) {
                    const bool on = i == 0 || i == lxi - 1 || j == 0 || j == let - 1 || k == 0 || k == lze - 1;
                    if (on) h[((size_t)k * let + j) * lxi + i] = Type(1);
                }
        check_hip(hipMemcpy(sbcc, h.data(), h.size() * sizeof(Type), hipMemcpyHostToDevice));
    }

    TypeIndex* npex = nullptr;
    Type* drvb[3] = {nullptr, nullptr, nullptr};
    Type* sbcc = nullptr;

private:
    t_dcomp d_;
};

/* gcbc_go, src/gcbc.hpp:86-186.  drva_buffer is the drva_buffer member of
 * the numerics object of the same grid (the reference call site passes
 * numerics_instance.drva_buffer); its boundary derivatives and BC codes are
 * the ones used.  The reference's nbc and mcd are accepted for the signature. */
template <typename Type>
void gcbc_go(Type* drva_buffer[3], Type* cm_buffer[3], Type* drvb_buffer[3], Type* qa, Type* de, Type* pressure,
             Type* yaco, Type* sbcc, t_point<Type> umf, t_point<Type> dudtmf, t_dcomp, Type dtwi, int[2][3],
             int[2][3]) {
    cs_numerics* h = cs_detail::lookup(drva_buffer[0]);
    if (!h) {
        std::fprintf(stderr, "gcbc_go: drva_buffer does not belong to a live numerics object\n");
        std::exit(EXIT_FAILURE);
    }
    const double* cm[3] = {cm_buffer[0], cm_buffer[1], cm_buffer[2]};
    const double* drvb[3] = {drvb_buffer ? drvb_buffer[0] : nullptr, drvb_buffer ? drvb_buffer[1] : nullptr,
                             drvb_buffer ? drvb_buffer[2] : nullptr};
    check_cs(cs_gcbc_go(h, cm, drvb, qa, de, pressure, yaco, sbcc, cs_detail::point_of(umf),
                        cs_detail::point_of(dudtmf), dtwi, nullptr));
}

/* wall_condition_update, src/gcbc.hpp:188-228.  Needs a geometry context for
 * (dcomp, nbc): a short-lived one is created per call (walls are updated
 * once per step, outside the line solves). */
template <typename Type, typename TypeIndex>
void wall_condition_update(Type* qa, TypeIndex* npex, t_point<Type> umf, t_dcomp dcomp_info, int nbc[2][3]) {
    static_assert(std::is_same<TypeIndex, int>::value, "npex is int");
    cs_dims d = cs_detail::dims_of(dcomp_info);
    cs_numerics* h = nullptr;
    check_cs(cs_numerics_create(&d, &nbc[0][0], &h));
    const int rc = cs_wall_condition_update(h, qa, npex, cs_detail::point_of(umf), nullptr);
    cs_numerics_destroy(h);
    check_cs(rc);
}

/* RK stage update q = q0 - dtk*yaco*de (absent upstream; driver/canard.cu:196-200) */
template <typename Type>
void rk_stage(t_dcomp dcomp_info, Type* q, Type* q0, Type* de, Type* yaco, Type dtk, cs_stream stream = nullptr) {
    cs_dims d = cs_detail::dims_of(dcomp_info);
    check_cs(cs_rk_stage(&d, q, q0, de, yaco, dtk, stream));
}

#endif /* COMPACT_SCHEMES_HPP */
