// dmc_ae.hpp -- the all-electron DMC mover of CHAMP (dmc/dmc_ae_mov1.f90,
// used by dmc.f90:98-102 when nloc = 0): the drift is split into the parts
// along and across the direction from the nearest nucleus, with a drift
// factor that grows near the nucleus, and the diffusion is a Gaussian around
// the drifted point or, with the probability the drift has of overshooting
// the nucleus, an exponential centred on the nucleus itself (the UNR93
// algorithm).  Forward proposal, reverse Green function and the velocity
// ratio of dmc_eloc_cutoff for nloc = 0 (branching.f90:240-278).
#pragma once

#include "metrop6.hpp"

namespace champ {

struct AeFrame {
    int ic;                  // nearest nucleus
    double rmin, vr, va, vp; // distance, radial, total and perpendicular drift
    double v2, zeta, zeff;
    double xaxis[3], zaxis[3];
};

// The frame of a position x with drift v (dmc_ae_mov1.f90:226-275 and
// :388-444): z axis from the nearest nucleus, x axis along the drift's
// perpendicular part.
__device__ inline void ae_frame(const double *cent, const int *iwctype, const double *znuc, int NC, double tau,
                                const double x[3], const double v[3], AeFrame &f)
{
#pragma clang fp contract(off)
    const double eps = 1.e-10;
    double d2;
    f.ic = m6_nearest(cent, NC, x[0], x[1], x[2], d2);
    double rv[3];
    double rmin = 0, vr = 0, v2 = 0;
    for (int k = 0; k < 3; k++) {
        rv[k] = x[k] - cent[3 * f.ic + k];
        rmin = rmin + rv[k] * rv[k];
        vr = vr + v[k] * rv[k];
        v2 = v2 + v[k] * v[k];
    }
    rmin = sqrt(rmin);
    vr = vr / rmin;
    f.rmin = rmin;
    f.vr = vr;
    f.v2 = v2;
    f.va = sqrt(v2);
    f.zeff = znuc[iwctype[f.ic]];
    f.zeta = sqrt(1.0 / tau + f.zeff * f.zeff);
    double vp = 0;
    for (int k = 0; k < 3; k++) {
        f.zaxis[k] = rv[k] / rmin;
        f.xaxis[k] = v[k] - vr * f.zaxis[k];
        vp = vp + f.xaxis[k] * f.xaxis[k];
    }
    vp = sqrt(vp);
    if (vp < eps) {
        f.xaxis[0] = eps * (1.0 - f.zaxis[0] * f.zaxis[0]);
        f.xaxis[1] = eps * (-f.zaxis[0] * f.zaxis[1]);
        f.xaxis[2] = eps * (-f.zaxis[0] * f.zaxis[2]);
        vp = eps * sqrt(1.0 + eps - f.zaxis[0] * f.zaxis[0]);
    }
    for (int k = 0; k < 3; k++) f.xaxis[k] = f.xaxis[k] / vp;
    f.vp = vp;
}

// adrift of the frame and the drift factor vavvt (dmc_ae_mov1.f90:273-277)
__device__ inline double ae_vavvt(const AeFrame &f, double tau)
{
#pragma clang fp contract(off)
    const double eps = 1.e-10, adrift0 = 0.1;
    const double hz = 0.5 * f.zeff * f.rmin;
    const double hafzr2 = hz * hz;
    const double adrift = (0.5 * (1 + eps + f.vr / f.va)) + adrift0 * hafzr2 / (1 + hafzr2);
    return (sqrt(1.0 + 2.0 * adrift * f.v2 * tau) - 1.0) / (adrift * f.v2);
}

// (sqrt(2 pi tau))^3 / pi, the exponential's normalisation (dmc_ae_mov1.f90:147-148)
__device__ inline double ae_term(double tau)
{
#pragma clang fp contract(off)
    const double pi = 3.141592653589793;
    const double t = sqrt(2.0 * pi * tau);
    return t * t * t / pi;
}

// Forward move of the electron at xo with drift vo (dmc_ae_mov1.f90:226-353):
// xn, dfus2o and the normalisation fnormo of the Green function's mixture.
// Draws one uniform, then three Gaussians or five uniforms, from s.
__device__ inline void ae_propose(const double *cent, const int *iwctype, const double *znuc, int NC, double tau,
                                  const double xo[3], const double vo[3], uint64_t s[4], double xn[3],
                                  double &dfus2o, double &fnormo)
{
#pragma clang fp contract(off)
    const double pi = 3.141592653589793;
    const double rttau = sqrt(tau);
    AeFrame f;
    ae_frame(cent, iwctype, znuc, NC, tau, xo, vo, f);
    const double *c = cent + 3 * f.ic;
    const double vavvt = ae_vavvt(f, tau);
    const double driftr = vavvt * f.vr;
    double rtry = f.rmin + driftr, qgaus;
    if (rtry > 0) {
        qgaus = 0.5 * erfc(rtry / sqrt(2.0 * tau));
        const double xprime = vavvt * f.vp * rtry / (0.5 * (f.rmin + rtry)), zprime = rtry;
        for (int k = 0; k < 3; k++) xn[k] = c[k] + f.xaxis[k] * xprime + f.zaxis[k] * zprime;
    } else {
        qgaus = 0.5 * (1.0 + erf(-rtry / sqrt(2.0 * tau)));
        for (int k = 0; k < 3; k++) xn[k] = c[k];
    }
    const double pgaus = 1.0 - qgaus;
    double dfus2a = 0, dfusb;
    if (xo_uniform(s) < pgaus) {
        double dfus2b = 0;
        for (int k = 0; k < 3; k++) {
            const double dfus = xo_gauss(s) * rttau;
            dfus2a = dfus2a + dfus * dfus;
            xn[k] = xn[k] + dfus;
            const double dc = xn[k] - c[k];
            dfus2b = dfus2b + dc * dc;
        }
        dfusb = sqrt(dfus2b);
    } else {
        const double u1 = xo_uniform(s), u2 = xo_uniform(s), u3 = xo_uniform(s);
        dfusb = (-0.5 / f.zeta) * log(u1 * u2 * u3);
        const double costht = 2.0 * (xo_uniform(s) - 0.5);
        const double sintht = sqrt(1.0 - costht * costht);
        const double phi = 2.0 * pi * xo_uniform(s);
        for (int k = 0; k < 3; k++) {
            const double dfus = k == 0 ? dfusb * sintht * cos(phi) : (k == 1 ? dfusb * sintht * sin(phi) : dfusb * costht);
            const double dd = c[k] + dfus - xn[k];
            dfus2a = dfus2a + dd * dd;
            xn[k] = c[k] + dfus;
        }
    }
    dfus2o = dfus2a;
    const double z3 = f.zeta * f.zeta * f.zeta;
    fnormo = (pgaus + qgaus * ae_term(tau) * z3 * exp(-2.0 * f.zeta * dfusb + 0.5 * dfus2a / tau));
}

// Reverse Green function from xn (drift vn) back to xo (dmc_ae_mov1.f90:388-480):
// dfus2n and fnormn.
__device__ inline void ae_reverse(const double *cent, const int *iwctype, const double *znuc, int NC, double tau,
                                  const double xo[3], const double xn[3], const double vn[3], double &dfus2n,
                                  double &fnormn)
{
#pragma clang fp contract(off)
    AeFrame f;
    ae_frame(cent, iwctype, znuc, NC, tau, xn, vn, f);
    const double *c = cent + 3 * f.ic;
    const double vavvt = ae_vavvt(f, tau);
    const double driftr = vavvt * f.vr;
    double rtry = f.rmin + driftr, qgaus, dfus2a = 0, dfus2b = 0;
    if (rtry > 0) {
        qgaus = 0.5 * erfc(rtry / sqrt(2.0 * tau));
        const double xprime = vavvt * f.vp * rtry / (0.5 * (f.rmin + rtry)), zprime = rtry;
        for (int k = 0; k < 3; k++) {
            const double xbac = c[k] + f.xaxis[k] * xprime + f.zaxis[k] * zprime;
            const double dc = c[k] - xo[k];
            dfus2b = dfus2b + dc * dc;
            const double dfus = xbac - xo[k];
            dfus2a = dfus2a + dfus * dfus;
        }
    } else {
        qgaus = 0.5 * (1.0 + erf(-rtry / sqrt(2.0 * tau)));
        for (int k = 0; k < 3; k++) {
            const double dfus = c[k] - xo[k];
            dfus2b = dfus2b + dfus * dfus;
        }
        dfus2a = dfus2b;
    }
    dfus2n = dfus2a;
    const double pgaus = 1.0 - qgaus, dfusb = sqrt(dfus2b);
    const double z3 = f.zeta * f.zeta * f.zeta;
    fnormn = pgaus + qgaus * ae_term(tau) * z3 * exp(-2.0 * f.zeta * dfusb + 0.5 * dfus2a / tau);
}

// One electron's terms of the velocity ratio for nloc = 0 (branching.f90:240-278,
// adrift = adrift0 = 0.1, tratio = 1): adds vavvn^2 v^2 and v^2.
__device__ inline void ae_fratio_terms(const double *cent, const int *iwctype, const double *znuc, int NC,
                                       double tau, const double x[3], const double v[3], double &vav2,
                                       double &v2sum)
{
#pragma clang fp contract(off)
    AeFrame f;
    ae_frame(cent, iwctype, znuc, NC, tau, x, v, f);
    const double vavvt = ae_vavvt(f, tau);
    const double vavvn = vavvt / tau;
    vav2 = vav2 + vavvn * vavvn * f.v2;
    v2sum = v2sum + f.v2;
}

}  // namespace champ
