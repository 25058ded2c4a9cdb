// metrop6.hpp -- CHAMP's accelerated Metropolis move (imetro = 6,
// metrop_mov1_slat.f90:3-738; C.J. Umrigar, PRL 71, 408 (1993)) for one
// walker, evaluated by one lane: the forward proposal in local spherical
// coordinates about the electron's nearest nucleus (k_move_a) and, once psi
// and the drift at the proposed position are known, the reverse transition
// density and the acceptance probability (k_move_b).
//
// The nearest nucleus of an electron, rvmino and rmino, are functions of its
// position: CHAMP sets them from xold at start (dumper_more.f90:271-286) and
// takes nearestn/rvminn over on acceptance, which are the same functions of
// the accepted position.  So nothing is carried between moves here: each move
// recomputes them, and the proposal hands the decision a short per-walker
// record (M6_*).  Contraction of products into FMAs is off in these scalar
// routines, so that their rounding follows the plain expressions.
#pragma once
#include "device_common.hpp"

namespace champ {

// per-walker record from the proposal to the decision ([W][M6_REC] doubles)
enum {
    M6_FXOP = 0,  // forward density factors of r and phi at the proposal
    M6_AREAO,     // forward normalisation (radial area times theta range)
    M6_RMINN,     // distance of the proposal to its nearest nucleus
    M6_RMINON,    // distance of the old position to that nucleus
    M6_COSTHT,    // cosine of the angle between the two at that nucleus
    M6_RAVERN,    // mean of the two distances
    M6_DELTTN,    // theta range of the reverse move
    M6_RRATIO,    // rminno / rminon (density of the angular move)
    M6_LIVE,      // 1: psi at the proposal decides; 0: geometric rejection (p = 0)
    M6_NEARN,     // nearest nucleus of the proposal
    M6_REC = 12
};

// metropolis parameters of metrop6 (parser.f90:402-404, checked at :774-787)
struct M6Params {
    double deltar, deltat, fbias;
};

// gammai (gammai.f90:3-108) for a = 3/2 or 5/2 with xae = x^a e^-x: a fixed
// number of series terms below a + 4 (lower incomplete gamma, iflag 1), else a
// fixed number of continued-fraction steps (minus the upper one, iflag -1)
__device__ inline double m6_gammai(double a, double x, double xae, int &iflag)
{
#pragma clang fp contract(off)
    if (x < a + 4) {
        iflag = 1;
        double ap = a, sum = 1 / a, del = sum;
        const int iter = (int)(21 + 3 * (x - a));
        for (int n = 1; n <= iter; n++) {
            ap = ap + 1;
            del = del * x / ap;
            sum = sum + del;
        }
        return sum * xae;
    }
    iflag = -1;
    double a0 = 1, a1 = x, b0 = 0, b1 = 1, fac = 1, g = 0;
    const double t = 20 + round(a - x);
    const int iter = t > 7 ? (int)t : 7;  // x >= a + 4 here, so iter <= 16
    for (int n = 1; n <= iter; n++) {
        const double an = (double)n, ana = an - a;
        a0 = (a1 + a0 * ana) * fac;
        b0 = (b1 + b0 * ana) * fac;
        const double anf = an * fac;
        a1 = x * a0 + anf * a1;
        b1 = x * b0 + anf * b1;
        fac = 1 / a1;
        g = b1 * fac;
    }
    return -(g * xae);
}

__device__ inline double m6_cube(double x) { return x * (x * x); }  // x**3 as __powidf2 forms it

// integral of sqrt(r) |1 + co r| exp(-zeta r) over [rbot, rtop] divided by bot,
// split at the zero of 1 + co r (metrop_mov1_slat.f90:256-293, 490-527)
__device__ inline double m6_area(double zeta, double co, double rbot, double rtop, double bot)
{
#pragma clang fp contract(off)
    const double d5b2 = 2.5, d2b3 = .666666666666667, g5b2 = 1.329340388179137;
    int ib, it, iz;
    const double rzero = -1 / co;
    const double zrbot = zeta * rbot, zrtop = zeta * rtop;
    const double zebot = sqrt(m6_cube(zrbot)) * exp(-zrbot);
    const double zetop = sqrt(m6_cube(zrtop)) * exp(-zrtop);
    const double g52bot = m6_gammai(d5b2, zrbot, zrbot * zebot, ib);
    const double g52top = m6_gammai(d5b2, zrtop, zrtop * zetop, it);
    if (rzero < rbot || rzero > rtop) {
        const double g52dif = ib * it == 1 ? g52top - g52bot : g52top - g52bot + g5b2;
        const double g32dif = d2b3 * (g52dif + zetop - zebot);
        return fabs(g32dif + co * g52dif / zeta) / (bot * sqrt(m6_cube(zeta)));
    }
    const double zrzer = zeta * rzero;
    const double zezer = sqrt(m6_cube(zrzer)) * exp(-zrzer);
    const double g52zer = m6_gammai(d5b2, zrzer, zrzer * zezer, iz);
    const double g52dif1 = ib * iz == 1 ? g52zer - g52bot : g52zer - g52bot + g5b2;
    const double g32dif1 = d2b3 * (g52dif1 + zezer - zebot);
    const double g52dif2 = it * iz == 1 ? g52top - g52zer : g52top - g52zer + g5b2;
    const double g32dif2 = d2b3 * (g52dif2 + zetop - zezer);
    return (fabs(g32dif1 + co * g52dif1 / zeta) + fabs(g32dif2 + co * g52dif2 / zeta)) / (bot * sqrt(m6_cube(zeta)));
}

// exponent zeta and linear coefficient co of the Slater-type radial density
// from the radial velocity (metrop_mov1_slat.f90:204-218, 473-486)
__device__ inline void m6_zeta_co(double zcusp, double vr, double rmin, double &zeta, double &co)
{
#pragma clang fp contract(off)
    const double eps = 1.e-10;
    double root = (zcusp + vr) * (zcusp + vr - 4 / rmin), z;
    if (root >= 0) {
        root = sqrt(root);
        z = 0.5 * (zcusp - vr - root);
        if (z <= 0) z = 0.5 * (zcusp - vr + root);
        if (z <= 0) z = eps;
    } else {
        z = vr < 0 ? -vr - eps : 1;
    }
    zeta = z;
    co = (z + vr) / (1 - (z + vr) * rmin);
}

// thetamx (metrop_mov1_slat.f90:740-752)
__device__ inline double m6_thetamx(double r, double z, double deltat)
{
#pragma clang fp contract(off)
    const double zr = z * r;
    return deltat + (2 - deltat) / (1 + zr * zr);
}

// nearest nucleus of a point (first index on ties) and its squared distance
__device__ inline int m6_nearest(const double *cent, int NC, double x0, double x1, double x2, double &d2)
{
#pragma clang fp contract(off)
    int best = 0;
    double dmin = 99.e9;
    for (int j = 0; j < NC; j++) {
        const double e0 = x0 - cent[3 * j], e1 = x1 - cent[3 * j + 1], e2 = x2 - cent[3 * j + 2];
        double dist = 0;
        dist = dist + e0 * e0;
        dist = dist + e1 * e1;
        dist = dist + e2 * e2;
        if (dist < dmin) { dmin = dist; best = j; }
    }
    d2 = dmin;
    return best;
}

// Forward proposal of electron i (metrop_mov1_slat.f90:129-375) from its old
// position xo, its drift vo and the walker's stream s: writes the proposal to
// xn and the record rec.  The radial and azimuthal rejection loops are capped
// at 10^6 trials so that every lane finishes; a capped walker keeps its
// position (rec[M6_LIVE] = 0 and p = 0, the one extra uniform still drawn).
__device__ inline void m6_propose(const double *cent, const int *iwctype, const double *znuc, int nloc, int NC,
                                  const M6Params &mp, const double xo[3], const double vo[3], uint64_t s[4],
                                  double xn[3], double *rec)
{
#pragma clang fp contract(off)
    const double eps = 1.e-10, pi = 3.141592653589793;  // 4 atan(1) in double
    const double deltar = mp.deltar, deltri = 1 / deltar;
    double d2;
    const int nearo = m6_nearest(cent, NC, xo[0], xo[1], xo[2], d2);
    double rvmino[3], xaxis[3], yaxis[3], zaxis[3];
    for (int k = 0; k < 3; k++) rvmino[k] = xo[k] - cent[3 * nearo + k];
    const double rmino = sqrt(rvmino[0] * rvmino[0] + rvmino[1] * rvmino[1] + rvmino[2] * rvmino[2]);
    const double znear = znuc[iwctype[nearo]];
    const double zcusp = nloc == 0 ? znear : 0.0;
    const double rbot = rmino * deltri, rtop = rmino * deltar;
    double voldr = 0;
    for (int k = 0; k < 3; k++) voldr = voldr + vo[k] * rvmino[k];
    voldr = voldr / rmino;
    double voldp = 0;
    for (int k = 0; k < 3; k++) {
        zaxis[k] = rvmino[k] / rmino;
        xaxis[k] = vo[k] - voldr * zaxis[k];
        voldp = voldp + xaxis[k] * xaxis[k];
    }
    voldp = sqrt(voldp);
    if (voldp < eps) {
        xaxis[0] = eps * (1 - zaxis[0] * zaxis[0]);
        xaxis[1] = eps * (-zaxis[0] * zaxis[1]);
        xaxis[2] = eps * (-zaxis[0] * zaxis[2]);
        voldp = eps * sqrt(1 - zaxis[0] * zaxis[0]);
    }
    for (int k = 0; k < 3; k++) xaxis[k] = xaxis[k] / voldp;
    voldr = fmax(voldr, -2 * znear);
    yaxis[0] = zaxis[1] * xaxis[2] - zaxis[2] * xaxis[1];
    yaxis[1] = zaxis[2] * xaxis[0] - zaxis[0] * xaxis[2];
    yaxis[2] = zaxis[0] * xaxis[1] - zaxis[1] * xaxis[0];
    voldp = voldp * mp.fbias;
    double zeta, co;
    m6_zeta_co(zcusp, voldr, rmino, zeta, co);
    // the density's maximum on [rbot, rtop] bounds the rejection sampling of r
    const double t15 = 1.5 * co - zeta;
    const double root = sqrt(t15 * t15 + 2 * zeta * co);
    double rmax1 = (t15 + root) / (2 * zeta * co);
    rmax1 = rmax1 < rbot ? rbot : rmax1;
    rmax1 = rmax1 > rtop ? rtop : rmax1;
    double fmx = sqrt(rmax1) * fabs(1 + co * rmax1) * exp(-zeta * rmax1);
    if (co < 0) {
        double rmax2 = (t15 - root) / (2 * zeta * co);
        rmax2 = rmax2 < rbot ? rbot : rmax2;
        rmax2 = rmax2 > rtop ? rtop : rmax2;
        const double f2 = sqrt(rmax2) * fabs(1 + co * rmax2) * exp(-zeta * rmax2);
        fmx = fmx > f2 ? fmx : f2;
    }
    const double bot = sqrt(rmino) * fabs(1 + co * rmino) * exp(-zeta * rmino);
    double rtry = 0, top = 0;
    bool ok = false;
    for (int trial = 0; trial < 1000000 && !ok; trial++) {
        rtry = ((deltar - deltri) * xo_uniform(s) + deltri) * rmino;
        top = sqrt(rtry) * fabs(1 + co * rtry) * exp(-zeta * rtry);
        ok = top / fmx > xo_uniform(s);
    }
    double fxop = top / bot;
    double areao = m6_area(zeta, co, rbot, rtop, bot);
    const double raver = 0.5 * (rmino + rtry);
    const double deltt = m6_thetamx(raver, znear, mp.deltat);
    double costht = 1 - deltt * xo_uniform(s);
    const double zprime = rtry * costht;
    double sintht = sqrt(1 - costht * costht);
    areao = areao * deltt;
    const double term = fmin(voldp * raver * sintht, 1.0);
    fmx = 1 + term;
    double phitry = 0;
    bool okp = false;
    for (int trial = 0; trial < 1000000 && !okp; trial++) {
        phitry = pi * xo_uniform(s);
        top = 1 + term * cos(phitry);
        okp = top > xo_uniform(s) * fmx;
    }
    fxop = fxop * top;
    const double xprime = rtry * sintht * cos(phitry);
    double yprime = sqrt(fmax(0.0, rtry * rtry - xprime * xprime - zprime * zprime));
    if (xo_uniform(s) < 0.5) yprime = -yprime;
    for (int k = 0; k < 3; k++)
        xn[k] = xaxis[k] * xprime + yaxis[k] * yprime + zaxis[k] * zprime + cent[3 * nearo + k];
    // geometric rejection: the reverse move must be one the forward rule can make
    const int nearn = m6_nearest(cent, NC, xn[0], xn[1], xn[2], d2);
    const double rminn = sqrt(d2);
    double rminon = 0, dot = 0;
    for (int k = 0; k < 3; k++) {
        const double rn = xn[k] - cent[3 * nearn + k], ro = xo[k] - cent[3 * nearn + k];
        rminon = rminon + ro * ro;
        dot = dot + rn * ro;
    }
    rminon = sqrt(rminon);
    dot = dot / (rminn * rminon);
    const double ravern = 0.5 * (rminn + rminon);
    const double delttn = m6_thetamx(ravern, znuc[iwctype[nearn]], mp.deltat);
    const bool geo = rminon > rminn * deltar || dot < 1 - delttn;
    rec[M6_FXOP] = fxop;
    rec[M6_AREAO] = areao;
    rec[M6_RMINN] = rminn;
    rec[M6_RMINON] = rminon;
    rec[M6_COSTHT] = dot;
    rec[M6_RAVERN] = ravern;
    rec[M6_DELTTN] = delttn;
    rec[M6_RRATIO] = rtry / rminon;
    rec[M6_LIVE] = (ok && okp && !geo) ? 1.0 : 0.0;
    rec[M6_NEARN] = (double)nearn;
}

// Acceptance probability of the proposal xn from xo (metrop_mov1_slat.f90:
// 420-563): the reverse transition density from the drift vn at xn, then
// p = rratio^2 psi2n/psi2o |fxnp areao / (fxop arean)|, at most 1; 0 for a
// geometric rejection, a zero psi, or a NaN (as the reference's dmin1 keeps
// a NaN, which then never passes the uniform test).
__device__ inline double m6_accept_prob(const double *cent, const int *iwctype, const double *znuc, int nloc,
                                        const M6Params &mp, const double xo[3], const double xn[3],
                                        const double vn[3], double psi2n, double psi2o, bool psi_zero,
                                        const double *rec)
{
#pragma clang fp contract(off)
    if (rec[M6_LIVE] == 0.0 || psi_zero) return 0.0;
    const double eps = 1.e-10;
    const double deltar = mp.deltar, deltri = 1 / deltar;
    const int nearn = (int)rec[M6_NEARN];
    const double rminn = rec[M6_RMINN], rminon = rec[M6_RMINON], costht = rec[M6_COSTHT];
    double rvminn[3], rvminon[3], xaxis[3];
    for (int k = 0; k < 3; k++) {
        rvminn[k] = xn[k] - cent[3 * nearn + k];
        rvminon[k] = xo[k] - cent[3 * nearn + k];
    }
    const double sintht = sqrt(1 - costht * costht);
    const double rbot = rminn * deltri, rtop = rminn * deltar;
    double vnewr = 0;
    for (int k = 0; k < 3; k++) vnewr = vnewr + vn[k] * rvminn[k];
    vnewr = vnewr / rminn;
    double vnewp = 0;
    for (int k = 0; k < 3; k++) {
        xaxis[k] = vn[k] - vnewr * rvminn[k] / rminn;
        vnewp = vnewp + xaxis[k] * xaxis[k];
    }
    vnewp = sqrt(vnewp);
    if (vnewp < eps) {
        xaxis[0] = eps * (1 - rvminn[0] * rvminn[0]);
        xaxis[1] = eps * (-rvminn[0] * rvminn[1]);
        xaxis[2] = eps * (-rvminn[0] * rvminn[2]);
        vnewp = eps * sqrt(1 - rvminn[0] * rvminn[0] * (2 - rminn * rminn));
    }
    for (int k = 0; k < 3; k++) xaxis[k] = xaxis[k] / vnewp;
    const double znearn = znuc[iwctype[nearn]];
    const double zcusp = nloc == 0 ? znearn : 0.0;
    vnewr = fmax(vnewr, -2 * znearn);
    vnewp = vnewp * mp.fbias;
    double zeta, co;
    m6_zeta_co(zcusp, vnewr, rminn, zeta, co);
    const double bot = sqrt(rminn) * fabs(1 + co * rminn) * exp(-zeta * rminn);
    const double top = sqrt(rminon) * fabs(1 + co * rminon) * exp(-zeta * rminon);
    double fxnp = top / bot;
    double arean = m6_area(zeta, co, rbot, rtop, bot);
    arean = arean * rec[M6_DELTTN];
    const double term = fmin(vnewp * rec[M6_RAVERN] * sintht, 1.0);
    double cosphi = 0, rnorm = 0;
    for (int k = 0; k < 3; k++) {
        const double term2 = rvminon[k] / rminon - costht * rvminn[k] / rminn;
        rnorm = rnorm + term2 * term2;
        cosphi = cosphi + term2 * xaxis[k];
    }
    cosphi = cosphi / sqrt(rnorm);
    fxnp = fxnp * (1 + term * cosphi);
    const double rr = rec[M6_RRATIO];
    const double p = rr * rr * exp(psi2n - psi2o) * fabs((fxnp * rec[M6_AREAO]) / (rec[M6_FXOP] * arean));
    return p < 1.0 ? p : (p >= 1.0 ? 1.0 : 0.0);
}

}  // namespace champ
