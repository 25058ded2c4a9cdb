// Host check of een_poly.hpp: the table + Horner form of the e-e-n polynomial
// against the direct term sum of jastrow4.f90:203-262 (value and the first and
// second derivatives in u, x and y), for nordc = 2..5 at random points.
// Prints the worst relative deviation per order; exit status 1 if any exceeds
// 1e-12.
#include "../../champ_amd/csrc/een_poly.hpp"

#include <cstdio>
#include <random>

using namespace champ;

struct Ref { double f, fu, fx, fy, fuu, fxx, fyy, fux, fuy; };

static double pw(double v, int p) { double r = 1; for (int i = 0; i < p; i++) r *= v; return r; }
// d^d/dv^d of v^p
static double dpw(double v, int p, int d)
{
    double c = 1;
    for (int i = 0; i < d; i++) { if (p - i <= 0) return 0; c *= (p - i); }
    return c * pw(v, p - d);
}

template <int N>
static Ref direct(const double *c, double u, double x, double y)
{
    EenTable<N> T{};
    Ref r{0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int t = 0; t < T.n; t++) {
        const int k = T.k[t], l = T.l[t], m = T.m[t];
        // c u^k (x^l + y^l)(xy)^m = c u^k (x^(l+m) y^m + x^m y^(l+m))
        auto g = [&](int du, int dx, int dy) {
            return c[t] * dpw(u, k, du) * (dpw(x, l + m, dx) * dpw(y, m, dy) + dpw(x, m, dx) * dpw(y, l + m, dy));
        };
        r.f += g(0, 0, 0); r.fu += g(1, 0, 0); r.fx += g(0, 1, 0); r.fy += g(0, 0, 1);
        r.fuu += g(2, 0, 0); r.fxx += g(0, 2, 0); r.fyy += g(0, 0, 2); r.fux += g(1, 1, 0); r.fuy += g(1, 0, 1);
    }
    return r;
}

static double relerr(double a, double b) { return std::fabs(a - b) / std::max(1.0, std::fabs(b)); }

template <int N>
static double check(std::mt19937_64 &g)
{
    std::uniform_real_distribution<double> U(-1, 1), P(0.0, 1.0);
    constexpr int NA = een_naz<N>();
    double worst = 0;
    for (int trial = 0; trial < 2000; trial++) {
        double c[40];
        for (double &v : c) v = U(g);
        const double u = P(g), x = P(g), y = P(g);
        double tab[3 * NA];
        een_tables<N, 2>(c, x, tab);
        const EenOut<double> o = een_eval_tab<N, 2>(tab, u, y);
        const EenOut<double> o1 = een_eval_tab<N, 1>(tab, u, y);
        const EenOut<double> o0 = een_eval_tab<N, 0>(tab, u, y);
        const Ref r = direct<N>(c, u, x, y);
        const double e[] = {relerr(o.f, r.f), relerr(o.fu, r.fu), relerr(o.fx, r.fx), relerr(o.fy, r.fy),
                            relerr(o.fuu, r.fuu), relerr(o.fxx, r.fxx), relerr(o.fyy, r.fyy), relerr(o.fux, r.fux),
                            relerr(o.fuy, r.fuy), relerr(o1.f, r.f), relerr(o1.fu, r.fu), relerr(o1.fx, r.fx),
                            relerr(o1.fy, r.fy), relerr(o0.f, r.f)};
        for (double v : e) worst = std::max(worst, v);
    }
    return worst;
}

int main()
{
    std::mt19937_64 g(12345);
    const double w[] = {check<2>(g), check<3>(g), check<4>(g), check<5>(g)};
    int bad = 0;
    for (int i = 0; i < 4; i++) {
        std::printf("nordc=%d worst=%.3e\n", i + 2, w[i]);
        if (!(w[i] < 1e-12)) bad = 1;
    }
    // structure at nordc = 4: 13 terms, 12 non-zero coefficients
    EenTable<4> T{};
    std::printf("nordc=4 terms=%d naz=%d\n", T.n, T.naz);
    if (T.n != 13 || T.naz != 12) bad = 1;
    return bad;
}
