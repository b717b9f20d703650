// Reference-style user program against the C++ drop-in (include/viltrum_amd/viltrum.hpp):
// the calls of main/doc/integrators.cc, with device-evaluable integrands.  Prints one
// line per result for tests/test_gpu_dropin.py.
#include <cstdio>
#include <vector>

#include "viltrum_amd/viltrum.hpp"

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
    return 0;
}
