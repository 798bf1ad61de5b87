// Reference-ordered right-hand side through the C++ operator classes, the
// call sequence of the reference driver (driver/canard.cu:132-205) with the
// gcbc step enabled:
//   primitives -> physics::calc_viscous_shear_stress -> physics::calc_fluxes
//   -> gcbc_go
// usage: test_rhs_mirror DIR lxi let lze viscous nbc0..nbc5 umfx umfy umfz
// DIR holds the inputs as raw little-endian float64 files (qa, xim, etm,
// zem, yaco, ssk, sbcc, cm0, cm1, cm2); the RHS is written to DIR/de.
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "compact_schemes.hpp"

static std::vector<double> load(const std::string& path, size_t n) {
    std::vector<double> v(n);
    FILE* f = std::fopen(path.c_str(), "rb");
    if (!f || std::fread(v.data(), sizeof(double), n, f) != n) {
        std::fprintf(stderr, "cannot read %zu values from %s\n", n, path.c_str());
        std::exit(2);
    }
    std::fclose(f);
    return v;
}

static double* upload(const std::vector<double>& h) {
    double* d = nullptr;
    check_hip(hipMalloc((void**)&d, h.size() * sizeof(double)));
    check_hip(hipMemcpy(d, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice));
    return d;
}

template <bool Visc>
static void rhs(const std::string& dir, t_dcomp dc, int nbc[2][3], t_point<double> umf) {
    const size_t lmx = (size_t)dc.lmx;
    const size_t fs[3] = {(size_t)dc.let * dc.lze, (size_t)dc.lxi * dc.lze, (size_t)dc.lxi * dc.let};
    double* qa = upload(load(dir + "/qa", 5 * lmx));
    double* xim = upload(load(dir + "/xim", 3 * lmx));
    double* etm = upload(load(dir + "/etm", 3 * lmx));
    double* zem = upload(load(dir + "/zem", 3 * lmx));
    double* yaco = upload(load(dir + "/yaco", lmx));
    double* ssk = upload(load(dir + "/ssk", lmx));
    double* cm[3];
    for (int a = 0; a < 3; a++) cm[a] = upload(load(dir + "/cm" + std::to_string(a), 2 * 3 * fs[a]));
    double *de = nullptr, *pressure = nullptr;
    check_hip(hipMalloc((void**)&de, 5 * lmx * sizeof(double)));
    check_hip(hipMalloc((void**)&pressure, lmx * sizeof(double)));
    int mcd[2][3] = {{0, 0, 0}, {0, 0, 0}};

    numerics_pc<double> numerics_instance(dc, nbc);
    numerics_instance.deriv_setup();
    physics<Visc, double> physics_instance(dc);
    physics_instance.umf = umf;
    gcbc<double, int> gcbc_instance(dc);
    gcbc_instance.init();

    cs_dims d{dc.lxi, dc.let, dc.lze};
    check_cs(cs_primitives(&d, qa, de, pressure, nullptr));
    if constexpr (Visc)
        physics_instance.calc_viscous_shear_stress(de, ssk, xim, etm, zem, yaco, dc, mcd, &numerics_instance,
                                                   nullptr);
    physics_instance.calc_fluxes(qa, pressure, de, xim, etm, zem, dc, mcd, &numerics_instance, nullptr);
    gcbc_go(numerics_instance.drva_buffer, cm, gcbc_instance.drvb, qa, de, pressure, yaco, gcbc_instance.sbcc,
            physics_instance.umf, physics_instance.dudtmf, dc, 0.0, nbc, mcd);
    check_hip(hipDeviceSynchronize());

    std::vector<double> h(5 * lmx);
    check_hip(hipMemcpy(h.data(), de, h.size() * sizeof(double), hipMemcpyDeviceToHost));
    FILE* f = std::fopen((dir + "/de").c_str(), "wb");
    if (!f || std::fwrite(h.data(), sizeof(double), h.size(), f) != h.size()) {
        std::fprintf(stderr, "cannot write %s/de\n", dir.c_str());
        std::exit(2);
    }
    std::fclose(f);
    double* bufs[] = {qa, xim, etm, zem, yaco, ssk, cm[0], cm[1], cm[2], de, pressure};
    for (double* b : bufs) (void)hipFree(b);
}

int main(int argc, char** argv) {
    if (argc != 15) {
        std::fprintf(stderr, "usage: %s DIR lxi let lze viscous nbc0..nbc5 umfx umfy umfz\n", argv[0]);
        return 2;
    }
    const std::string dir = argv[1];
    t_dcomp dc;
    dc.lxi = std::atoi(argv[2]);
    dc.let = std::atoi(argv[3]);
    dc.lze = std::atoi(argv[4]);
    dc.lmx = dc.lxi * dc.let * dc.lze;
    const int visc = std::atoi(argv[5]);
    int nbc[2][3];
    for (int k = 0; k < 6; k++) nbc[k / 3][k % 3] = std::atoi(argv[6 + k]);
    t_point<double> umf{std::atof(argv[12]), std::atof(argv[13]), std::atof(argv[14])};
    if (visc) rhs<true>(dir, dc, nbc, umf);
    else rhs<false>(dir, dc, nbc, umf);
    std::printf("PASS (wrote %s/de)\n", dir.c_str());
    return 0;
}
