// Reference-style user program against the C++ drop-in (include/viltrum_amd/viltrum.hpp):
// the calls of main/doc/integrators.cc, with device-evaluable integrands.  Prints one
// line per result for tests/test_gpu_dropin.py.
#include <cstdio>
#include <vector>

#include "viltrum_amd/viltrum.hpp"

// a logger (src/log.h:10-19 shape) that keeps the region list an integrator logs
struct KeepRegions {
    viltrum::RegionList regions;
    void log(const viltrum::RegionLog& r) { regions = r.fetch(); }
    template <typename Data>
    void log(const Data&) {}
};

int main() {
    // x^2 + y^2 over [0,1]^2 with integrator_crespo2021(16, 64): 2/3
    vmx_poly_params pp{};
    pp.nterms = 2;
    pp.coef[0] = 1.0;
    pp.expo[0][0] = 2;
    pp.coef[1] = 1.0;
    pp.expo[1][1] = 2;
    auto f = viltrum::poly(pp, 2);
    const double v = viltrum::integrate(viltrum::integrator_crespo2021(16, 64), f, viltrum::range_primary<2, double>());
    std::printf("scalar %.17g\n", v);

    // config-1 integrand into 32x32 float bins (vector<vector<float>>)
    vmx_blobs_params bp{};
    bp.nblobs = 1;
    const double blob[7] = {0.3, 0.6, 0.0, 0.0, 50.0, 50.0, 1.0};
    for (int k = 0; k < 7; ++k) bp.blob[0][k] = blob[k];
    bp.step_mode = 1;
    bp.step[0] = 1.0;
    bp.step[1] = 1.0;
    bp.step[3] = 1.0;
    bp.step[4] = 0.5;
    bp.wave[0] = 0.1;
    bp.wave[1] = 20.0;
    bp.wave[2] = 13.0;
    auto g = viltrum::blobs(bp, 2);
    std::vector<std::vector<float>> bins(32, std::vector<float>(32, 0.f));
    viltrum::integrate(viltrum::integrator_crespo2021(255, 16, 1), bins, g, viltrum::range_primary<2, float>());
    double s = 0;
    for (auto& row : bins)
        for (float x : row) s += x;
    std::printf("bins %.9g %.9g %.9g\n", double(bins[3][7]), double(bins[20][11]), s);

    // Fubini form: one blob moving in t, regions over (x, y), t integrated by Monte Carlo
    vmx_blobs_params mp{};
    mp.nblobs = 1;
    const double moving[7] = {0.4, 0.5, 0.2, -0.1, 40.0, 60.0, 1.0};
    for (int k = 0; k < 7; ++k) mp.blob[0][k] = moving[k];
    auto h = viltrum::blobs(mp, 3);
    std::vector<std::vector<float>> fb(16, std::vector<float>(16, 0.f));
    viltrum::integrate(viltrum::integrator_crespo2021_infinite<2>(200, 2, 8, 4), fb, h,
                       viltrum::range_primary<3, float>());
    double sf = 0;
    for (auto& row : fb)
        for (float x : row) sf += x;
    std::printf("fubini %.9g %.9g %.9g\n", double(fb[5][9]), double(fb[12][3]), sf);

    // three bin dims (src/integrate.h:150-167): the moving blob into 8x8x4 float bins
    std::vector<std::vector<std::vector<float>>> b3(8, std::vector<std::vector<float>>(8, std::vector<float>(4, 0.f)));
    viltrum::integrate(viltrum::integrator_crespo2021(300, 4, 2), b3, h, viltrum::range_primary<3, float>());
    double s3 = 0;
    for (auto& plane : b3)
        for (auto& row : plane)
            for (float x : row) s3 += x;
    std::printf("bins3 %.9g %.9g %.9g\n", double(b3[1][2][3]), double(b3[6][5][0]), s3);

    // a logger that keeps the generated regions, then the per-bin phase replayed over
    // them with more samples and another seed
    KeepRegions keep;
    std::vector<std::vector<float>> b1(32, std::vector<float>(32, 0.f));
    viltrum::integrate(viltrum::integrator_crespo2021(255, 0, 1), b1, g, viltrum::range_primary<2, float>(), keep);
    std::vector<std::vector<float>> b2(32, std::vector<float>(32, 0.f));
    auto acc = [&b2](const std::array<std::size_t, 2>& i) -> float& { return b2[i[0]][i[1]]; };
    viltrum::LoggerNull quiet;
    viltrum::integrator_crespo2021(0, 64, 9).integrate_regions(acc, std::array<std::size_t, 2>{32, 32}, keep.regions, g,
                                                               viltrum::range_primary<2, float>(), quiet);
    double s2 = 0;
    for (auto& row : b2)
        for (float x : row) s2 += x;
    std::printf("replay %zu %.9g %.9g %.9g\n", keep.regions.count, double(b2[3][7]), double(b2[20][11]), s2);

    // the Dyson series over an infinite range (main/compilation-tests/crespo21.cc:27-35):
    // integrator_crespo2021_infinite<4> into 10 bins, and the optimised Fubini integrator
    auto dy = viltrum::dyson(0.75);
    std::vector<float> sol_bins(10, 0.0f);
    viltrum::integrate(viltrum::integrator_crespo2021_infinite<4>(1280, 8, 256), sol_bins, dy,
                       viltrum::range_primary_infinite<float>());
    double si = 0;
    for (float x : sol_bins) si += x;
    std::printf("infinite %.9g %.9g %.9g\n", double(sol_bins[0]), double(sol_bins[9]), si);
    std::vector<float> opt_bins(10, 0.0f);
    viltrum::integrate(viltrum::integrator_adaptive_fubini_variance_reduction_parallel_optimized<2>(300, 4, 32, 7),
                       opt_bins, dy, viltrum::range_infinite<float>({0.0f, 0.0f, 0.1f}, {1.0f, 1.0f, 0.9f}));
    double so = 0;
    for (float x : opt_bins) so += x;
    std::printf("optimized %.9g %.9g %.9g\n", double(opt_bins[0]), double(opt_bins[9]), so);
    return 0;
}
