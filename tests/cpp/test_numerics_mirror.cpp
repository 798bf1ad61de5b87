// Ramp known-answer test of the derivative through the C++ operator classes
// (the reference's tests/test_numerics_serial.cu:41-146, in FP64): a field
// equal to its linear index has the constant derivative 1, lxi, lxi*let along
// x, y, z, and the closure rows are exact for it.  Exit status 0 = pass.
#include <cmath>
#include <cstdio>
#include <vector>

#include "compact_schemes.hpp"

template <unsigned int Axis>
static int test_deriv1d_pc() {
    t_dcomp dcomp_info;
    dcomp_info.lxi = 256;
    dcomp_info.let = 128;
    dcomp_info.lze = 64;
    dcomp_info.lmx = dcomp_info.lxi * dcomp_info.let * dcomp_info.lze;

    std::vector<double> infield(dcomp_info.lmx), outfield(dcomp_info.lmx);
    for (int i = 0; i < dcomp_info.lmx; ++i) infield[i] = i;
    double *d_infield = nullptr, *d_outfield = nullptr;
    check_hip(hipMalloc((void**)&d_infield, dcomp_info.lmx * sizeof(double)));
    check_hip(hipMalloc((void**)&d_outfield, dcomp_info.lmx * sizeof(double)));
    check_hip(hipMemcpy(d_infield, infield.data(), dcomp_info.lmx * sizeof(double), hipMemcpyHostToDevice));

    int nstart = 0, nend = 0;
    int nbc[2][3];
    for (unsigned int nn = 0; nn < NumberOfSpatialDims; ++nn)
        for (unsigned int ip = 0; ip < NumberOfFaces; ++ip) nbc[ip][nn] = BC_NON_REFLECTIVE;

    numerics_pc<double> numerics_instance(dcomp_info, nbc);
    numerics_instance.deriv_setup();
    for (unsigned int i = 0; i < 10; ++i)
        numerics_instance.template deriv1d<Axis>(d_infield, d_outfield, nstart, nend, dcomp_info, 0);
    check_hip(hipMemcpy(outfield.data(), d_outfield, dcomp_info.lmx * sizeof(double), hipMemcpyDeviceToHost));

    double solution;
    if constexpr (Axis == 0) solution = infield[1] - infield[0];
    else if constexpr (Axis == 1) solution = infield[dcomp_info.lxi] - infield[0];
    else solution = infield[dcomp_info.lxi * dcomp_info.let] - infield[0];

    double worst = 0.0;
    for (int i = 0; i < dcomp_info.lmx; ++i) worst = std::fmax(worst, std::abs(outfield[i] - solution) / solution);
    (void)hipFree(d_infield);
    (void)hipFree(d_outfield);
    std::printf("deriv1d<%u> ramp: max relative error %.3e\n", Axis, worst);
    return worst < 1e-12 ? 0 : 1;
}

int main() {
    int fails = 0;
    fails += test_deriv1d_pc<0>();
    fails += test_deriv1d_pc<1>();
    fails += test_deriv1d_pc<2>();
    std::printf("%s\n", fails ? "FAIL" : "PASS");
    return fails ? 1 : 0;
}
