// Host program that drives the C++ adapter (include/heliocubed_boxop.hpp) the
// way HelioCubed's StageLevel drives its per-box operator: one operator for
// one box, a_Rhs accumulated, a_fluxes written, a_scale applied. It reads the
// box's ghosted input (raw doubles) from a file, runs the operator on the
// device and writes Rhs, F0, F1, F2 (raw doubles) to a file; it then checks
// that a view on the wrong box is rejected.
//
// usage: hc_boxop_adapter_check order n_rad n_ang spatial_operator limiter
//            riemann precision block lo0 lo1 lo2 hi0 hi1 hi2 scale in.bin out.bin
// order 4 reads native conserved averages (ghost 5), order 2 Cartesian
// primitives (ghost 2, the a_cartesianPrimitives call).
#include <hip/hip_runtime_api.h>

#include <array>
#include <cstdio>
#include <cstdlib>
#include <stdexcept>
#include <vector>

#include "../../include/heliocubed_boxop.hpp"

static void check(hipError_t e, const char* what)
{
    if (e != hipSuccess) { std::fprintf(stderr, "%s: %s\n", what, hipGetErrorString(e)); std::exit(2); }
}

int main(int argc, char** argv)
{
    if (argc != 18) { std::fprintf(stderr, "usage: see the source header\n"); return 2; }
    ++argv;   // argv[0] is the order from here on; the rest as before
    hc_options opt{};
    opt.spatial_order = std::atoi(argv[0]);
    hc_shell shell{std::atoi(argv[1]), std::atoi(argv[2]), 21.5, 215.0, 0.0};
    opt.spatial_operator = std::atoi(argv[3]);
    opt.limiter = std::atoi(argv[4]);
    opt.riemann_solver = std::atoi(argv[5]);
    opt.precision = std::atoi(argv[6]);
    opt.entropy_fix = 0.3;
    opt.gamma = 5.0/3.0;
    opt.berta_eta_d = 0.02;
    opt.berta_fallback = 0;
    opt.divergence_cleaning = 0;
    const unsigned block = (unsigned)std::atoi(argv[7]);
    const int lo[3] = {std::atoi(argv[8]), std::atoi(argv[9]), std::atoi(argv[10])};
    const int hi[3] = {std::atoi(argv[11]), std::atoi(argv[12]), std::atoi(argv[13])};
    const double scale = std::atof(argv[14]);
    const int g = opt.spatial_order == 2 ? 2 : 5;
    long long n[3], e[3];
    for (int d = 0; d < 3; ++d) { n[d] = hi[d]-lo[d]+1; e[d] = n[d]+2*g; }
    const size_t nu = 8*e[0]*e[1]*e[2], nr = 8*n[0]*n[1]*n[2];
    size_t nf[3];
    for (int d = 0; d < 3; ++d) nf[d] = 8*(n[0]+(d == 0))*(n[1]+(d == 1))*(n[2]+(d == 2));

    std::vector<double> u(nu);
    FILE* f = std::fopen(argv[15], "rb");
    if (!f || std::fread(u.data(), sizeof(double), nu, f) != nu) { std::fprintf(stderr, "cannot read %s\n", argv[15]); return 2; }
    std::fclose(f);

    double *dU = nullptr, *dR = nullptr, *dF[3] = {nullptr, nullptr, nullptr};
    check(hipMalloc(&dU, nu*sizeof(double)), "hipMalloc");
    check(hipMalloc(&dR, nr*sizeof(double)), "hipMalloc");
    for (int d = 0; d < 3; ++d) check(hipMalloc(&dF[d], nf[d]*sizeof(double)), "hipMalloc");
    check(hipMemcpy(dU, u.data(), nu*sizeof(double), hipMemcpyHostToDevice), "hipMemcpy");
    check(hipMemset(dR, 0, nr*sizeof(double)), "hipMemset");

    heliocubed::DeviceBoxData rhs, usph;
    std::array<heliocubed::DeviceBoxData, 3> fluxes;
    rhs.data = dR; usph.data = dU;
    for (int d = 0; d < 3; ++d)
    {
        rhs.lo[d] = lo[d]; rhs.hi[d] = hi[d];
        usph.lo[d] = lo[d]-g; usph.hi[d] = hi[d]+g;
    }
    for (int d = 0; d < 3; ++d)
    {
        fluxes[d].data = dF[d];
        for (int q = 0; q < 3; ++q) { fluxes[d].lo[q] = lo[q]; fluxes[d].hi[q] = hi[q]+(q == d); }
    }
    try
    {
        heliocubed::BoxOp_EulerCubedSphere op(shell, opt, block, lo, hi);
        if (op.ghost() != g) { std::fprintf(stderr, "ghost %d, expected %d\n", op.ghost(), g); return 1; }
        op(rhs, fluxes, usph, 0.1, scale, -1, opt.spatial_order == 2);
        check(hipDeviceSynchronize(), "hipDeviceSynchronize");
        // a view on the wrong box is rejected before any device work
        heliocubed::DeviceBoxData bad = rhs;
        bad.hi[0] += 1;
        bool rejected = false;
        try { op(bad, fluxes, usph, 0.1, scale, -1, opt.spatial_order == 2); }
        catch (const std::runtime_error&) { rejected = true; }
        if (!rejected) { std::fprintf(stderr, "a Rhs view on the wrong box was accepted\n"); return 1; }
    }
    catch (const std::exception& ex) { std::fprintf(stderr, "adapter: %s\n", ex.what()); return 1; }

    std::vector<double> out(nr + nf[0] + nf[1] + nf[2]);
    check(hipMemcpy(out.data(), dR, nr*sizeof(double), hipMemcpyDeviceToHost), "hipMemcpy");
    size_t o = nr;
    for (int d = 0; d < 3; ++d)
    {
        check(hipMemcpy(out.data()+o, dF[d], nf[d]*sizeof(double), hipMemcpyDeviceToHost), "hipMemcpy");
        o += nf[d];
    }
    f = std::fopen(argv[16], "wb");
    if (!f || std::fwrite(out.data(), sizeof(double), out.size(), f) != out.size()) { std::fprintf(stderr, "cannot write\n"); return 2; }
    std::fclose(f);
    for (int d = 0; d < 3; ++d) (void)hipFree(dF[d]);
    (void)hipFree(dR); (void)hipFree(dU);
    std::printf("adapter ok\n");
    return 0;
}
