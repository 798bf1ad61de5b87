// canard.cpp -- the reference's run driver (driver/canard.cu) on the C-ABI
// library: run configuration -> block grid -> initial state -> RK4 steps of
// the fused solver -> legacy VTK output.
//
//   canard [config.yaml [output_dir]]
//
// Without arguments the configuration is $CANARD_ROOT/config.yaml, as the
// reference reads it (driver/canard.cu:30-37).  One block runs as one
// subdomain on the current GPU (nbpc > 1 is reported and the block is not
// split: halo exchange between GPUs is not part of this build).  The block
// grid is the reference's (src/grid.hpp:89-110): uniform spacings
// (doml0 + doml1)/lximb, 2 domh/letmb, -span/lzemb, i.e. constant diagonal
// metrics, the fused solver's case.  Time loop (driver/canard.cu:131-205):
// the step is the CFL time step of each step's state (the reference leaves
// dt at a constant and its update commented out), output whenever the time
// crosses (ndati + 1) tmax / ndata, and once more at the end.  Output files
// output_data<N>.vtk hold x, y, z and density, u, v, w, p (the velocities
// include the frame velocity).
#include <hip/hip_runtime_api.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "compact_schemes.h"

namespace {

void check(int rc, const char* what) {
    if (rc != CS_OK) {
        std::fprintf(stderr, "canard: %s: %s\n", what, cs_last_error());
        std::exit(EXIT_FAILURE);
    }
}

void check_hip(hipError_t e, const char* what) {
    if (e != hipSuccess) {
        std::fprintf(stderr, "canard: %s: %s\n", what, hipGetErrorString(e));
        std::exit(EXIT_FAILURE);
    }
}

struct Output {
    std::string dir;
    int number = 0;
    const cs_dims* d;
    const std::vector<double>* xyz;
    cs_point umf;
    double *qa, *de, *p;   // device: state, primitives, pressure
    hipStream_t s;

    void write() {
        const long lmx = (long)d->lxi * d->let * d->lze;
        check(cs_primitives(d, qa, de, p, s), "primitives");
        std::vector<double> h(7 * lmx);
        check_hip(hipMemcpyAsync(h.data(), qa, sizeof(double) * lmx, hipMemcpyDeviceToHost, s), "copy rho");
        check_hip(hipMemcpyAsync(h.data() + lmx, de + lmx, sizeof(double) * 3 * lmx, hipMemcpyDeviceToHost, s),
                  "copy velocity");
        check_hip(hipMemcpyAsync(h.data() + 4 * lmx, p, sizeof(double) * lmx, hipMemcpyDeviceToHost, s), "copy p");
        check_hip(hipStreamSynchronize(s), "output synchronisation");
        const double uf[3] = {umf.x, umf.y, umf.z};
        for (int c = 0; c < 3; c++)
            for (long i = 0; i < lmx; i++) h[(1 + c) * lmx + i] += uf[c];
        const char* names[5] = {"density", "u", "v", "w", "p"};
        const double* vars[5] = {h.data(), h.data() + lmx, h.data() + 2 * lmx, h.data() + 3 * lmx, h.data() + 4 * lmx};
        const std::string path = dir + "/output_data" + std::to_string(number) + ".vtk";
        check(cs_vtk_write(path.c_str(), d, xyz->data(), 5, names, vars, 32), "vtk output");
        std::printf("output %s\n", path.c_str());
        number++;
    }
};

}  // namespace

int main(int argc, char** argv) {
    std::string path;
    if (argc > 1) {
        path = argv[1];
    } else {
        const char* root = std::getenv("CANARD_ROOT");
        if (!root) {
            std::fprintf(stderr, "usage: canard [config.yaml [output_dir]] (or set CANARD_ROOT)\n");
            return EXIT_FAILURE;
        }
        path = std::string(root) + "/config.yaml";
    }
    const std::string outdir = argc > 2 ? argv[2] : ".";

    cs_run_config cfg;
    check(cs_config_load(path.c_str(), &cfg), "configuration");
    cs_dims d;
    int nbc[6];
    check(cs_config_dims(&cfg, &d, nbc), "configuration");
    const long lmx = (long)d.lxi * d.let * d.lze;
    if (cfg.nbpc[0] * cfg.nbpc[1] * cfg.nbpc[2] > 1)
        std::printf("note: nbpc = [%d, %d, %d]; the block runs as one subdomain on this GPU\n", cfg.nbpc[0],
                    cfg.nbpc[1], cfg.nbpc[2]);
    std::printf("block %d x %d x %d, faces [%d %d %d] [%d %d %d], Mach (%g, %g, %g)\n", d.lxi, d.let, d.lze, nbc[0],
                nbc[1], nbc[2], nbc[3], nbc[4], nbc[5], cfg.amach.x, cfg.amach.y, cfg.amach.z);

    std::vector<double> xyz(3 * lmx);
    check(cs_config_grid(&cfg, xyz.data()), "grid");
    const double dx = (cfg.doml0 + cfg.doml1) / cfg.lximb;
    const double dy = 2.0 * cfg.domh / cfg.letmb;
    const double dz = -cfg.span / cfg.lzemb;

    cs_solver_config sc = {};
    sc.dims = d;
    for (int i = 0; i < 6; i++) sc.nbc[i] = nbc[i];
    sc.viscous = cfg.viscous;
    sc.umf = cfg.amach;
    sc.mu = 1.0 / cfg.re;
    sc.metric[0] = dy * dz;          // cofactors of the diagonal d x / d xi
    sc.metric[1] = dz * dx;
    sc.metric[2] = dx * dy;
    sc.yaco = 1.0 / (dx * dy * dz);
    sc.cfl = cfg.cfl;
    sc.alpha_f = cfg.alpha_f;
    sc.filter = cfg.filter;
    cs_solver* sv = nullptr;
    check(cs_solver_create(&sc, &sv), "solver");

    double *qa, *sbcc, *de, *p;
    check_hip(hipMalloc(&qa, sizeof(double) * 5 * lmx), "hipMalloc");
    check_hip(hipMalloc(&de, sizeof(double) * 5 * lmx), "hipMalloc");
    check_hip(hipMalloc(&sbcc, sizeof(double) * lmx), "hipMalloc");
    check_hip(hipMalloc(&p, sizeof(double) * lmx), "hipMalloc");
    hipStream_t s;
    check_hip(hipStreamCreate(&s), "hipStreamCreate");
    const cs_point centre = {0.5 * (d.lxi - 1), 0.5 * (d.let - 1), 0.5 * (d.lze - 1)};
    check(cs_init_pulse(&d, centre, cfg.pulse_amp, cfg.pulse_width, qa, s), "initial state");
    check(cs_boundary_weights(&d, sbcc, s), "boundary weights");

    Output out{outdir, 0, &d, &xyz, cfg.amach, qa, de, p, s};
    long n = 0;
    int ndati = -1;
    double timo = 0.0;
    bool written = false;
    while (timo < cfg.tmax && (cfg.nsteps == 0 || n < cfg.nsteps)) {
        double dt = 0.0;
        check(cs_solver_time_step(sv, qa, &dt, s), "time step");
        const double res = (ndati + 1) * cfg.tmax / cfg.ndata;
        written = false;
        if ((timo - res) * (timo + dt - res) <= 0.0) {
            out.write();
            ndati++;
            written = true;
        }
        check(cs_solver_rk4_step(sv, qa, sbcc, dt, nullptr, s), "RK4 step");
        timo += dt;
        n++;
        std::printf("Time step = %ld  time = %.6e  dt = %.6e\n", n, timo, dt);
    }
    check_hip(hipStreamSynchronize(s), "synchronisation");
    if (n > 0 || !written) out.write();
    std::printf("%ld steps, time %.6e, %d output files\n", n, timo, out.number);

    cs_solver_destroy(sv);
    (void)hipFree(qa);
    (void)hipFree(de);
    (void)hipFree(sbcc);
    (void)hipFree(p);
    (void)hipStreamDestroy(s);
    return 0;
}
