// The branch-free Cartesian monomial of the HIP kernels (slm_decode, slm_value,
// slm_all in champ_amd/csrc/device_common.hpp) against the case-by-case form of
// slm.f90:7-200 (slm), for every l in 0..21 at random points: value, gradient
// and Laplacian to rounding, and exact zeros outside 1..20.
#include "../../champ_amd/csrc/device_common.hpp"

#include <cmath>
#include <cstdio>
#include <random>

using namespace champ;

int main()
{
    std::mt19937 gen(7);
    std::uniform_real_distribution<double> U(-3.0, 3.0);
    double worst = 0.0;
    int zero_mismatch = 0;
    auto cmp = [&](double a, double b) {
        if (b == 0.0) { if (a != 0.0) zero_mismatch++; return; }
        worst = std::fmax(worst, std::fabs(a - b) / std::fabs(b));
    };
    for (int l = 0; l <= 21; l++)
        for (int t = 0; t < 2000; t++) {
            const double x = U(gen), y = U(gen), z = U(gen);
            double v, dy[3], lap, w, ey[3], elap;
            slm<double>(l, x, y, z, v, dy, lap);
            const SlmMono m = slm_decode(l);
            slm_all<double>(m, x, y, z, w, ey, elap);
            cmp(slm_value<double>(m, x, y, z), v);
            cmp(w, v);
            for (int k = 0; k < 3; k++) cmp(ey[k], dy[k]);
            cmp(elap, lap);
        }
    std::printf("slm: largest relative difference %.3e, zero mismatches %d\n", worst, zero_mismatch);
    return (zero_mismatch == 0 && worst < 4e-15) ? 0 : 1;
}
