/*
 * champ_oracle.c -- CPU restatement (FP64) of CHAMP's per-move wavefunction
 * path, one walker per call, following the reference routines cited per
 * function (paths relative to the CHAMP source tree, src/vmc unless noted).
 *
 * TEST INFRASTRUCTURE ONLY: this file is the parity checker of the GPU engine
 * and the CPU baseline timed beside it.  Nothing in the engine links it.
 *
 * How each function is pinned (tests/test_oracle_*.py):
 *   orc_xoshiro_*, orc_random_dp  -- known-answer streams of the reference's own
 *                                    C file (xoroshiro256starstar.c) compiled
 *                                    from the reference tree into oracle/_ref.
 *   orc_gesqua/orc_rotqua         -- icosahedral rule integrates Y_lm exactly to
 *                                    l=5; rotation is orthogonal; weights 1/12.
 *   orc_matinv                    -- the two reference unit-test matrices
 *                                    (tests/unit_test/test_matinv.f90:10-39),
 *                                    A*A^-1 = I and det vs. brute force.
 *   orc_pot_nn                    -- reference unit test value for water
 *                                    (tests/unit_test/test_electron_repulsion.f90).
 *   orc_splfit + table build      -- spline reproduces nodes, matches the
 *                                    analytic Gaussian contraction to ~1e-8,
 *                                    derivatives vs the analytic ones.
 *   orc_slm, orc_basis_fns_point  -- closed-form Cartesian Gaussians (analytic
 *                                    value, gradient, Laplacian) and central
 *                                    finite differences.
 *   jastrow4 / jastrow4e          -- incremental update equals full recompute at
 *                                    the moved configuration; gradient and
 *                                    Laplacian vs finite differences; cusp
 *                                    d f_ee/dr = sspinn*b1 at r=0; J -> 0 as all
 *                                    distances grow (asymptote subtraction).
 *   determinante (SM)             -- ratio equals det(M_new)/det(M_old) by fresh
 *                                    LU; SM-updated inverse equals fresh inverse.
 *   nonloc quadrature             -- with psi replaced by a polynomial of degree
 *                                    <= 5 the 12-point rule reproduces the exact
 *                                    angular projection; yl0 = (2L+1) P_L.
 *   hpsi E_L                      -- kinetic part vs finite-difference Laplacian
 *                                    of ln psi built from the oracle's own psi;
 *                                    drift vs finite-difference gradient.
 *   multideterminant psi_D        -- every determinant built and inverted on its
 *                                    own (no excitation tables); drift, E_L and
 *                                    the ECP term vs finite differences and full
 *                                    psi at the quadrature points; per-move psi
 *                                    and drift vs a fresh evaluation; a reordered
 *                                    copy of the reference determinant gives
 *                                    (c0 +- c1) D_ref by the parity of the order.
 *   metrop6 (m6_*)                -- gammai against scipy's incomplete gamma;
 *                                    the radial normalisation against quadrature
 *                                    of its density; VMC energies of the
 *                                    reference's H2 and butadiene CI cases.
 *   DMC (orc_dmc_*, dmc_tmove,     -- through the GPU engine, which follows it step
 *        ae_propose/ae_reverse)      by step: the reference's DMC CI energy (with
 *                                    and without T-moves) and, for the
 *                                    all-electron mover, the nodeless H2 ground
 *                                    state's DMC energy being the same for two
 *                                    Jastrows.  The mixture normalisation of
 *                                    ae_propose has no check of its own.
 * The end-to-end per-walker values of the Fortran reference itself are not
 * available (the Fortran build needs an MPI module this image lacks), so the
 * composition of the pinned pieces is what fixes hpsi/psie.
 */
#include "champ_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define HB 0.5

struct oracle_walker {
    int nelec, nup, ndn, ncent, norb, nbasis, nmax, npair, nquad;
    double *x, *xnew_e;
    double *r_en, *rvec_en, *r_ee, *rvec_ee;
    double *r_en_sav, *rvec_en_sav, *r_ee_sav, *rvec_ee_sav;
    double fsumo, d2o, fsumn, d2n;
    double *fjo, *fjn, *fso, *fsn, *fijo, *fijn, *d2ijo, *d2ijn, *vj, *vjn;
    double *orb, *dorb, *ddorb, *orbn, *dorbn, *ddorbn;
    double *slmi, *slmin, *fp, *fpp, detiab[2], detn, *ddx, *d2dx2;
    double *phin, *dphin, *d2phin;
    double *vps;
    double xq[CHAMP_MPS_QUAD], yq[CHAMP_MPS_QUAD], zq[CHAMP_MPS_QUAD];
    double *asymp_jasa, asymp_jasb[2], pecent;
    double psid, psij, ekin, energy, pe_local, d2j, vpsp_det[2];
    int nxquad;
    double psido, psijo, psi2o, eold;
    double ratio_kref;
    uint64_t rng[4];
    /* multideterminant state (ndet > 1): every determinant's own LU inverse
     * and value per spin, the proposal's, and the orbital-space coefficients
     * vmd[i][o] = d ln psi_D / d phi_o(r_i) */
    int ndet;
    double *md_inv, *md_det, *md_invn, *md_detn, *vmd, md_psidn;
    /* t_vpsp (nonloc.f90:297): [i][ic][iq] det_ratio * term_radial_jas, zero
     * for pairs beyond the cut-off; the T-move weights of DMC */
    double *tvpsp;
};

/* ------------------------------------------------------------------------ */
/* integer power as gfortran's real**integer (libgcc __powidf2)             */
double orc_powi(double x, int m)
{
    unsigned n = m < 0 ? (unsigned)(-m) : (unsigned)m;
    double y = (n % 2) ? x : 1.0;
    while (n >>= 1) {
        x = x * x;
        if (n % 2) y = y * x;
    }
    return m < 0 ? 1.0 / y : y;
}

/* ------------------------------------------------------------------------ */
/* splfit.f90:3-94, grid type 3 only (the TREXIO table build uses gridtype=3, */
/* read_data_trexio.f90:991-1010).                                          */
void orc_splfit(const champ_system_t *s, int it, int irb, double r, double wfv[4], int ider)
{
    const int nr = s->nr;
    const double r0 = s->grid_r0, arg = s->grid_arg;
    const size_t base = ((size_t)it * s->nrbas_max + irb);
    const double *rwf = s->rwf + base * nr;
    const double *d2rwf = s->d2rwf + base * nr;
    const double *ce = s->ce + base * CHAMP_NCOEF;
    const double *ae = s->ae + base * 2;
    const double sixth = 1.0 / 6.0;

    double dlogag = log(arg);
    double xr = (log((r + r0) / r0)) / dlogag + 1.0;
    int jx = (int)xr;
    double ref0 = r0 * orc_powi(arg, jx - 1) - r0;
    double ref1 = (ref0 + r0) * arg - r0;
    double delh = ref1 - ref0;

    if (jx < 1) {
        wfv[0] = ce[0];
        for (int ii = 2; ii <= CHAMP_NCOEF; ii++) wfv[0] += ce[ii - 1] * orc_powi(r, ii - 1);
        if (ider >= 1) {
            wfv[1] = wfv[2] = wfv[3] = 0.0;
            for (int ii = 2; ii <= CHAMP_NCOEF; ii++) {
                wfv[1] += (ii - 1) * ce[ii - 1] * orc_powi(r, ii - 2);
                wfv[2] += (ii - 1) * (ii - 2) * ce[ii - 1] * orc_powi(r, ii - 3);
                wfv[3] += (ii - 1) * (ii - 2) * (ii - 3) * ce[ii - 1] * orc_powi(r, ii - 4);
            }
        }
    } else if (jx >= nr) {
        wfv[0] = ae[0] * exp(-ae[1] * r);
        if (ider >= 1) {
            wfv[1] = -ae[1] * wfv[0];
            wfv[2] = -ae[1] * wfv[1];
            wfv[3] = -ae[1] * wfv[2];
        }
    } else {
        /* Fortran rwf(jx) with 1-based jx is element jx-1 here */
        const double y0 = rwf[jx - 1], y1 = rwf[jx], s0 = d2rwf[jx - 1], s1 = d2rwf[jx];
        double delhi = 1.0 / delh;
        double bb = (r - ref0) * delhi;
        double aa = (ref1 - r) * delhi;
        double cc = aa * (aa * aa - 1.0) * delh * delh * sixth;
        double dd = bb * (bb * bb - 1.0) * delh * delh * sixth;
        wfv[0] = aa * y0 + bb * y1 + cc * s0 + dd * s1;
        if (ider >= 1) {
            wfv[1] = (y1 - y0) * delhi + (-(3.0 * aa * aa - 1.0) * s0 + (3.0 * bb * bb - 1.0) * s1) * delh * sixth;
            wfv[2] = aa * s0 + bb * s1;
            wfv[3] = (-s0 + s1) * delhi;
        }
    }
}

/* ------------------------------------------------------------------------ */
/* slm.f90:7-200 (ider < 3 branch), Cartesian monomials 1..20 (s,p,d,f).     */
void orc_slm(int l, const double v[3], double r2, double *y, double dy[3], double *lap, int ider)
{
    (void)r2; (void)ider;
    const double cs = 0.28209479177387814, cp = 0.48860251190291992;
    const double cd1 = 0.630783130505040, cd2 = 1.092548430592079;
    const double cf = 0.74635266518023078, cf2 = 1.66889529453113635, cf3 = 2.89061144264055405;
    const double x = v[0], yy = v[1], z = v[2];
    dy[0] = dy[1] = dy[2] = 0.0;
    *lap = 0.0;
    switch (l) {
    case 1: *y = cs; break;
    case 2: *y = cp * x; dy[0] = cp; break;
    case 3: *y = cp * yy; dy[1] = cp; break;
    case 4: *y = cp * z; dy[2] = cp; break;
    case 5: *y = cd1 * x * x; dy[0] = 2.0 * cd1 * x; *lap = 2.0 * cd1; break;
    case 6: *y = cd2 * x * yy; dy[0] = cd2 * yy; dy[1] = cd2 * x; break;
    case 7: *y = cd2 * x * z; dy[0] = cd2 * z; dy[2] = cd2 * x; break;
    case 8: *y = cd1 * yy * yy; dy[1] = 2.0 * cd1 * yy; *lap = 2.0 * cd1; break;
    case 9: *y = cd2 * yy * z; dy[1] = cd2 * z; dy[2] = cd2 * yy; break;
    case 10: *y = cd1 * z * z; dy[2] = 2.0 * cd1 * z; *lap = 2.0 * cd1; break;
    case 11: *y = cf * x * x * x; dy[0] = cf * 3.0 * x * x; *lap = cf * 6.0 * x; break;
    case 12: *y = cf2 * x * x * yy; dy[0] = cf2 * 2.0 * x * yy; dy[1] = cf2 * x * x; *lap = cf2 * 2.0 * yy; break;
    case 13: *y = cf2 * x * x * z; dy[0] = cf2 * 2.0 * x * z; dy[2] = cf2 * x * x; *lap = cf2 * 2.0 * z; break;
    case 14: *y = cf2 * yy * yy * x; dy[1] = cf2 * 2.0 * yy * x; dy[0] = cf2 * yy * yy; *lap = cf2 * 2.0 * x; break;
    case 15: *y = cf3 * x * yy * z; dy[0] = cf3 * yy * z; dy[1] = cf3 * x * z; dy[2] = cf3 * x * yy; break;
    case 16: *y = cf2 * z * z * x; dy[0] = cf2 * z * z; dy[2] = cf2 * 2.0 * z * x; *lap = cf2 * 2.0 * x; break;
    case 17: *y = cf * yy * yy * yy; dy[1] = cf * 3.0 * yy * yy; *lap = cf * 6.0 * yy; break;
    case 18: *y = cf2 * yy * yy * z; dy[1] = cf2 * 2.0 * yy * z; dy[2] = cf2 * yy * yy; *lap = cf2 * 2.0 * z; break;
    case 19: *y = cf2 * z * z * yy; dy[1] = cf2 * z * z; dy[2] = cf2 * 2.0 * z * yy; *lap = cf2 * 2.0 * yy; break;
    case 20: *y = cf * z * z * z; dy[2] = cf * 3.0 * z * z; *lap = cf * 6.0 * z; break;
    default: *y = 0.0; break;
    }
}

/* phi_combine, basis_fns.f90:217-296 (ider 0..2) */
static void phi_combine(const double xc[3], double ri, const double wfv[4], double y, const double dy[3],
                        double ddy_lap, double *phi, double dphi[3], double *d2phi, int ider)
{
    *phi = y * wfv[0];
    if (ider == 1) {
        double xcri[3] = {xc[0] * ri, xc[1] * ri, xc[2] * ri};
        for (int jj = 0; jj < 3; jj++) dphi[jj] = y * xcri[jj] * wfv[1] + dy[jj] * wfv[0];
    } else if (ider >= 2) {
        double xcri[3] = {xc[0] * ri, xc[1] * ri, xc[2] * ri};
        double d2 = y * wfv[2] + y * 2.0 * ri * wfv[1] + ddy_lap * wfv[0];
        double dum = 0.0;
        for (int jj = 0; jj < 3; jj++) {
            dphi[jj] = y * xcri[jj] * wfv[1] + dy[jj] * wfv[0];
            dum = dum + dy[jj] * xcri[jj];
        }
        *d2phi = d2 + 2.0 * dum * wfv[1];
    }
}

/* basis_fns.f90:8-214 for one point (electron or quadrature point).        */
void orc_basis_fns_point(const champ_system_t *s, const double *rvec, const double *rr, int ider,
                         double *phi, double *dphi, double *d2phi)
{
    const int nb = s->nbasis;
    int lilm = 0;
    for (int ic = 0; ic < s->ncent; ic++) {
        const int it = s->iwctype[ic];
        const int nrb = s->nrbas[it], nbt = s->nbastyp[it];
        const double *xc = rvec + 3 * ic;
        const double r = rr[ic], r2 = r * r, ri = 1.0 / r;
        double wfv[64][4];
        for (int irb = 0; irb < nrb; irb++) {
            if (r <= s->rmaxwf[it * s->nrbas_max + irb]) orc_splfit(s, it, irb, r, wfv[irb], ider);
            else wfv[irb][0] = wfv[irb][1] = wfv[irb][2] = wfv[irb][3] = 0.0;
        }
        for (int i = 0; i < nbt; i++) {
            const int l = s->iwlbas[it * s->nbastyp_max + i];
            const int irad = s->iwrwf[it * s->nbastyp_max + i];
            double y, dy[3], lap, p, dp[3] = {0, 0, 0}, d2 = 0.0;
            orc_slm(l, xc, r2, &y, dy, &lap, ider);
            phi_combine(xc, ri, wfv[irad], y, dy, lap, &p, dp, &d2, ider);
            const int m = lilm + i;
            phi[m] = p;
            if (ider >= 1 && dphi) { dphi[m] = dp[0]; dphi[nb + m] = dp[1]; dphi[2 * nb + m] = dp[2]; }
            if (ider >= 2 && d2phi) d2phi[m] = d2;
        }
        lilm += nbt;
    }
}

/* ------------------------------------------------------------------------ */
/* matinv.f90:22-92.  a is column-major n x n and is replaced by its inverse.
 * LU with partial pivoting (dgetrf semantics: first maximal |pivot|, row
 * interchanges), determinant by the reference's mantissa/exponent loop,
 * inverse by solving A X = I with the factors.                              */
int orc_matinv(double *a, int n, double *determinant)
{
    if (n == 0) return 0;
    if (n == 1) { *determinant = a[0]; a[0] = 1.0 / a[0]; return 0; }
    if (n == 2) {
        double det = a[0] * a[3] - a[2] * a[1];
        *determinant = det;
        double deti = 1.0 / det, aux = a[0];
        a[0] = a[3] * deti; a[3] = aux * deti;
        a[1] = -a[1] * deti; a[2] = -a[2] * deti;
        return 0;
    }
#define A(i, j) a[(size_t)(j) * n + (i)]
    int *ipvt = (int *)malloc(sizeof(int) * n);
    int info = 0;
    for (int k = 0; k < n; k++) {
        int p = k; double amax = fabs(A(k, k));
        for (int i = k + 1; i < n; i++) if (fabs(A(i, k)) > amax) { amax = fabs(A(i, k)); p = i; }
        ipvt[k] = p;
        if (A(p, k) == 0.0) { if (!info) info = k + 1; continue; }
        if (p != k) for (int j = 0; j < n; j++) { double t = A(k, j); A(k, j) = A(p, j); A(p, j) = t; }
        double piv = 1.0 / A(k, k);
        for (int i = k + 1; i < n; i++) A(i, k) *= piv;
        for (int j = k + 1; j < n; j++) {
            double akj = A(k, j);
            if (akj != 0.0) for (int i = k + 1; i < n; i++) A(i, j) -= A(i, k) * akj;
        }
    }
    double det1 = 1.0, det2 = 0.0;
    int nonfinite = 0;
    for (int i = 0; i < n; i++) {
        if (ipvt[i] != i) det1 = -det1;
        det1 = A(i, i) * det1;
        if (det1 == 0.0) break;
        if (!isfinite(det1)) { nonfinite = 1; break; }   /* the loops below never end on NaN/Inf */
        for (int it = 0; it < 800 && fabs(det1) < 1.0; it++) { det1 = 10.0 * det1; det2 = det2 - 1.0; }
        for (int it = 0; it < 800 && fabs(det1) >= 10.0; it++) { det1 = det1 / 10.0; det2 = det2 + 1.0; }
    }
    *determinant = nonfinite ? det1 : det1 * pow(10.0, det2);
    if (info) { free(ipvt); return info; }
    /* inverse: X = U^-1 L^-1 P, column by column of P^T e_j */
    double *inv = (double *)malloc(sizeof(double) * n * n);
    double *col = (double *)malloc(sizeof(double) * n);
    for (int j = 0; j < n; j++) {
        for (int i = 0; i < n; i++) col[i] = (i == j) ? 1.0 : 0.0;
        for (int k = 0; k < n; k++) if (ipvt[k] != k) { double t = col[k]; col[k] = col[ipvt[k]]; col[ipvt[k]] = t; }
        for (int i = 1; i < n; i++) { double sum = col[i]; for (int k = 0; k < i; k++) sum -= A(i, k) * col[k]; col[i] = sum; }
        for (int i = n - 1; i >= 0; i--) { double sum = col[i]; for (int k = i + 1; k < n; k++) sum -= A(i, k) * col[k]; col[i] = sum / A(i, i); }
        for (int i = 0; i < n; i++) inv[(size_t)j * n + i] = col[i];
    }
    memcpy(a, inv, sizeof(double) * n * n);
    free(inv); free(col); free(ipvt);
#undef A
    return 0;
}

/* ------------------------------------------------------------------------ */
/* nterms4, jastrow4.f90:418-445 */
int orc_nterms4(int nord)
{
    int i = 0;
    for (int n = 2; n <= nord; n++)
        for (int k = n - 1; k >= 0; k--) {
            int l_hi = (k == 0) ? n - k - 2 : n - k;
            for (int l = l_hi; l >= 0; l--) { int m = (n - k - l) / 2; if (2 * m == n - k - l) i++; }
        }
    return i;
}

/* set_scale_dist, scale_dist.f90:3-60 */
void orc_asymp(const champ_system_t *s, double *ja, double jb[2])
{
    const double ar = 1.0 / s->scalek;
    const int na = s->norda + 1, nb = s->nordb + 1;
    for (int it = 0; it < s->nctype; it++) {
        const double *a4 = s->a4 + it * na;
        ja[it] = a4[0] * ar / (1 + a4[1] * ar);
        for (int iord = 2; iord <= s->norda; iord++) ja[it] += a4[iord] * orc_powi(ar, iord);
    }
    for (int i = 0; i < 2; i++) {
        double sspinn; int isp;
        if (i == 0) { sspinn = 1; isp = 0; }
        else { sspinn = (s->nspin2b == 1 && s->nocuspb == 0) ? 0.5 : 1.0; isp = s->nspin2b - 1; }
        const double *b = s->b + isp * nb;
        jb[i] = sspinn * b[0] * ar / (1 + b[1] * ar);
        for (int iord = 2; iord <= s->nordb; iord++) jb[i] += b[iord] * orc_powi(ar, iord);
    }
}

/* scale_dist1/2 + switch_scale1/2 for isc=2, scale_dist.f90:112-331 */
static inline void scale2(double kap, double r, double *rr, double *dd1, double *dd2)
{
    *rr = (1.0 - exp(-kap * r)) / kap;
    *dd1 = 1.0 - kap * (*rr);
    *dd2 = -kap * (*dd1);
}
static inline void switch2(double kap, double *rr, double *dd1, double *dd2)
{
    *rr = 1 - kap * (*rr);
    *dd1 = -kap * (*dd1);
    if (dd2) *dd2 = -kap * (*dd2);
}

/* ------------------------------------------------------------------------ */
/* pot_nn, pot.f90:3-50 */
double orc_pot_nn(const champ_system_t *s)
{
    double pecent = 0;
    for (int i = 1; i < s->ncent; i++)
        for (int j = 0; j < i; j++) {
            double r2 = 0;
            for (int k = 0; k < 3; k++) { double d = s->cent[3 * i + k] - s->cent[3 * j + k]; r2 = r2 + d * d; }
            double r = sqrt(r2);
            pecent = pecent + s->znuc[s->iwctype[i]] * s->znuc[s->iwctype[j]] / r;
        }
    return pecent;
}

/* gauss_pot, readps_gauss.f90:57-92 (l is the 1-based component index) */
void orc_gauss_pot(const champ_system_t *s, double r, int l, int ict, double *vpot, double *dvpot)
{
    double v = 0, dv = 0, rsq = r * r;
    const size_t base = ((size_t)ict * CHAMP_MPS_L + (l - 1));
    const int nt = s->necp_term[base];
    for (int i = 0; i < nt; i++) {
        const int pw = s->necp_power[base * CHAMP_MGAUSS + i];
        const double co = s->ecp_coef[base * CHAMP_MGAUSS + i], ex = s->ecp_exponent[base * CHAMP_MGAUSS + i];
        double p, dpv, e;
        if (pw != 2) { p = orc_powi(r, pw - 2); dpv = (pw - 2) * p / r; }
        else { p = 1.0; dpv = 0.0; }
        if (ex * rsq > 7.0e2) e = 0.0;
        else e = co * exp(-ex * rsq);
        v = v + p * e;
        dv = dv + (dpv - 2 * p * ex * r) * e;
    }
    *vpot = v;
    *dvpot = dv;
}

/* gesqua (rotqua.f90:3-353), icosahedral branch nq=12 (label 130) and the
 * octahedral nq=6 branch; weights from the ww table.                      */
void orc_gesqua(int nq, double *xq, double *yq, double *zq, double *wq)
{
    const double pi = 3.14159265359;
    for (int j = 0; j < CHAMP_MPS_QUAD; j++) xq[j] = yq[j] = zq[j] = wq[j] = 0.0;
    if (nq == 6) {
        for (int j = 0; j < 6; j++) wq[j] = 1.0 / 6.0;
        xq[0] = 1; xq[1] = -1; yq[2] = 1; yq[3] = -1; zq[4] = 1; zq[5] = -1;
        return;
    }
    /* nq == 12 */
    for (int j = 0; j < 12; j++) wq[j] = 1.0 / 12.0;
    zq[0] = 1.0; zq[1] = -1.0;
    double fi0 = pi / 5.0, s5 = sqrt(5.0), cstha = 1.0 / s5, sntha = 2.0 / s5;
    for (int j = 1; j <= 5; j++) {
        double rk2 = 2 * (double)(j - 1);
        double crk2 = cos(rk2 * fi0), srk2 = sin(rk2 * fi0);
        xq[j + 1] = sntha * crk2; yq[j + 1] = sntha * srk2; zq[j + 1] = cstha;
        double rk21 = rk2 + 1.0;
        double crk21 = cos(rk21 * fi0), srk21 = sin(rk21 * fi0);
        xq[j + 6] = sntha * crk21; yq[j + 6] = sntha * srk21; zq[j + 6] = -cstha;
    }
}

/* ------------------------------------------------------------------------ */
/* xoshiro256** 1.0 (Blackman & Vigna, 2018), as bound by random.f90:3-55   */
static inline uint64_t rotl64(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
uint64_t orc_xoshiro_next(uint64_t s[4])
{
    const uint64_t result = rotl64(s[1] * 5, 7) * 9;
    const uint64_t t = s[1] << 17;
    s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl64(s[3], 45);
    return result;
}
void orc_xoshiro_jump(uint64_t s[4])
{
    static const uint64_t J[4] = {0x180ec6d33cfd0abaULL, 0xd5a61266f0c9392cULL,
                                  0xa9582618e03fc9aaULL, 0x39abdc4529b1661cULL};
    uint64_t t[4] = {0, 0, 0, 0};
    for (int i = 0; i < 4; i++)
        for (int b = 0; b < 64; b++) {
            if (J[i] & ((uint64_t)1 << b)) { t[0] ^= s[0]; t[1] ^= s[1]; t[2] ^= s[2]; t[3] ^= s[3]; }
            orc_xoshiro_next(s);
        }
    s[0] = t[0]; s[1] = t[1]; s[2] = t[2]; s[3] = t[3];
}
double orc_random_dp(uint64_t s[4]) { return (double)(orc_xoshiro_next(s) >> 11) * 0x1.0p-53; }
/* gauss.f90:3-22 */
double orc_gauss(uint64_t s[4])
{
    const double pi = 4.0 * atan(1.0);
    double g = cos(2 * pi * orc_random_dp(s));
    return g * sqrt(-2 * log(orc_random_dp(s)));
}
/* setrn: 8 x int32 reinterpreted as 4 x uint64 (little endian), random.f90:27-36 */
void orc_setrn(uint64_t s[4], const int32_t seed[8])
{
    for (int i = 0; i < 4; i++)
        s[i] = (uint64_t)(uint32_t)seed[2 * i] | ((uint64_t)(uint32_t)seed[2 * i + 1] << 32);
}
/* rotqua.f90:357-430 */
void orc_rotqua(const champ_system_t *s, uint64_t rng[4], double *xq, double *yq, double *zq)
{
    double x1, x2, x3, xsum, cfi, sfi, sthet;
    do {
        x1 = 1.0 - 2.0 * orc_random_dp(rng);
        x2 = 1.0 - 2.0 * orc_random_dp(rng);
        xsum = x1 * x1 + x2 * x2;
    } while (xsum >= 1.0);
    double xsum2 = 2.0 * sqrt(fabs(1.0 - xsum));
    x1 = x1 * xsum2; x2 = x2 * xsum2; x3 = 1.0 - 2.0 * xsum;
    double theta = acos(x3);
    sthet = sin(theta);
    if (sthet < 1.0e-05) { cfi = 1.0; sfi = 0.0; x1 = 0.0; x2 = 0.0; x3 = 1.0; sthet = 0.0; }
    else { cfi = x1 / sthet; sfi = x2 / sthet; }
    double yy1 = x3 * cfi, yy2 = x3 * sfi, yy3 = -sthet, zz1 = -sfi, zz2 = cfi, zz3 = 0.0;
    double u1, u2, usum;
    do {
        u1 = orc_random_dp(rng) * 2.0 - 1.0;
        u2 = orc_random_dp(rng) * 2.0 - 1.0;
        usum = u1 * u1 + u2 * u2;
    } while (usum >= 1.0);
    double uu = sqrt(usum);
    u1 = u1 / uu; u2 = u2 / uu;
    double y1 = yy1 * u1 + zz1 * u2, y2 = yy2 * u1 + zz2 * u2, y3 = yy3 * u1 + zz3 * u2;
    double z1 = yy1 * u2 - zz1 * u1, z2 = yy2 * u2 - zz2 * u1, z3 = yy3 * u2 - zz3 * u1;
    for (int iq = 0; iq < s->nquad; iq++) {
        xq[iq] = s->xq0[iq] * x1 + s->yq0[iq] * y1 + s->zq0[iq] * z1;
        yq[iq] = s->xq0[iq] * x2 + s->yq0[iq] * y2 + s->zq0[iq] * z2;
        zq[iq] = s->xq0[iq] * x3 + s->yq0[iq] * y3 + s->zq0[iq] * z3;
    }
}

/* ------------------------------------------------------------------------ */
static void *xcalloc(size_t n, size_t sz) { void *p = calloc(n ? n : 1, sz); return p; }

oracle_walker_t *orc_walker_new(const champ_system_t *s)
{
    oracle_walker_t *w = (oracle_walker_t *)xcalloc(1, sizeof(*w));
    const int ne = s->nelec, nc = s->ncent;
    w->nelec = ne; w->nup = s->nup; w->ndn = ne - s->nup; w->ncent = nc;
    w->norb = s->norb; w->nbasis = s->nbasis; w->nquad = s->nquad;
    w->nmax = w->nup > w->ndn ? w->nup : w->ndn;
    w->npair = ne * (ne - 1) / 2;
    const int nm = w->nmax, nb = s->nbasis, no = s->norb;
    w->x = xcalloc(3 * ne, 8); w->xnew_e = xcalloc(3, 8);
    w->tvpsp = xcalloc((size_t)ne * nc * (s->nquad > 0 ? s->nquad : 1), 8);
    w->r_en = xcalloc((size_t)ne * nc, 8); w->rvec_en = xcalloc((size_t)3 * ne * nc, 8);
    w->r_ee = xcalloc(w->npair, 8); w->rvec_ee = xcalloc((size_t)3 * w->npair, 8);
    w->r_en_sav = xcalloc(nc, 8); w->rvec_en_sav = xcalloc(3 * nc, 8);
    w->r_ee_sav = xcalloc(ne, 8); w->rvec_ee_sav = xcalloc(3 * ne, 8);
    w->fjo = xcalloc(3 * ne, 8); w->fjn = xcalloc(3 * ne, 8);
    w->fso = xcalloc((size_t)ne * ne, 8); w->fsn = xcalloc((size_t)ne * ne, 8);
    w->fijo = xcalloc((size_t)3 * ne * ne, 8); w->fijn = xcalloc((size_t)3 * ne * ne, 8);
    w->d2ijo = xcalloc((size_t)ne * ne, 8); w->d2ijn = xcalloc((size_t)ne * ne, 8);
    w->vj = xcalloc(3 * ne, 8); w->vjn = xcalloc(3 * ne, 8);
    w->orb = xcalloc((size_t)ne * no, 8); w->dorb = xcalloc((size_t)3 * ne * no, 8);
    w->ddorb = xcalloc((size_t)ne * no, 8);
    w->orbn = xcalloc(no, 8); w->dorbn = xcalloc(3 * no, 8); w->ddorbn = xcalloc(no, 8);
    w->slmi = xcalloc((size_t)2 * nm * nm, 8); w->slmin = xcalloc((size_t)nm * nm, 8);
    w->fp = xcalloc((size_t)3 * nm * nm, 8); w->fpp = xcalloc((size_t)nm * nm, 8);
    w->ddx = xcalloc(3 * ne, 8); w->d2dx2 = xcalloc(ne, 8);
    w->phin = xcalloc(nb, 8); w->dphin = xcalloc(3 * nb, 8); w->d2phin = xcalloc(nb, 8);
    w->vps = xcalloc((size_t)ne * nc * CHAMP_MPS_L, 8);
    w->asymp_jasa = xcalloc(s->nctype, 8);
    w->ndet = s->ndet;
    if (s->ndet > 1) {
        w->md_inv = xcalloc((size_t)s->ndet * 2 * nm * nm, 8); w->md_det = xcalloc((size_t)s->ndet * 2, 8);
        w->md_invn = xcalloc((size_t)s->ndet * nm * nm, 8); w->md_detn = xcalloc(s->ndet, 8);
        w->vmd = xcalloc((size_t)ne * no, 8);
    }
    orc_asymp(s, w->asymp_jasa, w->asymp_jasb);
    w->pecent = orc_pot_nn(s);
    for (int iq = 0; iq < s->nquad; iq++) { w->xq[iq] = s->xq0[iq]; w->yq[iq] = s->yq0[iq]; w->zq[iq] = s->zq0[iq]; }
    return w;
}

void orc_walker_free(oracle_walker_t *w)
{
    if (!w) return;
    double *ps[] = {w->x, w->xnew_e, w->r_en, w->rvec_en, w->r_ee, w->rvec_ee, w->r_en_sav, w->rvec_en_sav,
                    w->r_ee_sav, w->rvec_ee_sav, w->fjo, w->fjn, w->fso, w->fsn, w->fijo, w->fijn,
                    w->d2ijo, w->d2ijn, w->vj, w->vjn, w->orb, w->dorb, w->ddorb, w->orbn, w->dorbn,
                    w->ddorbn, w->slmi, w->slmin, w->fp, w->fpp, w->ddx, w->d2dx2, w->phin, w->dphin,
                    w->d2phin, w->vps, w->asymp_jasa, w->md_inv, w->md_det, w->md_invn, w->md_detn, w->vmd,
                    w->tvpsp};
    for (size_t i = 0; i < sizeof(ps) / sizeof(ps[0]); i++) free(ps[i]);
    free(w);
}
void orc_walker_set_x(oracle_walker_t *w, const double *x) { memcpy(w->x, x, sizeof(double) * 3 * w->nelec); }
void orc_walker_get_x(const oracle_walker_t *w, double *x) { memcpy(x, w->x, sizeof(double) * 3 * w->nelec); }
void orc_walker_set_quad(oracle_walker_t *w, const double *xq, const double *yq, const double *zq)
{
    memcpy(w->xq, xq, sizeof(double) * w->nquad); memcpy(w->yq, yq, sizeof(double) * w->nquad);
    memcpy(w->zq, zq, sizeof(double) * w->nquad);
}
void orc_walker_set_rng(oracle_walker_t *w, const uint64_t st[4]) { memcpy(w->rng, st, 32); }
void orc_walker_get_rng(const oracle_walker_t *w, uint64_t st[4]) { memcpy(st, w->rng, 32); }

#define PAIR(i, j) ((size_t)(i) * ((i) - 1) / 2 + (j)) /* i > j, 0-based */
#define IJ(i, j) ((size_t)(i) * ne + (j))

/* ------------------------------------------------------------------------ */
/* distances.f90:3-167 (iel < 0 means all electrons; else one electron with save) */
static void distances(const champ_system_t *s, oracle_walker_t *w, int iel, const double *x)
{
    const int ne = w->nelec, nc = w->ncent;
    int i1, i2;
    if (iel < 0) { i1 = 0; i2 = ne - 1; }
    else {
        i1 = i2 = iel;
        for (int ic = 0; ic < nc; ic++) {
            w->r_en_sav[ic] = w->r_en[(size_t)iel * nc + ic];
            for (int m = 0; m < 3; m++) w->rvec_en_sav[3 * ic + m] = w->rvec_en[((size_t)iel * nc + ic) * 3 + m];
        }
        for (int jj = 0; jj < ne; jj++) {
            if (jj == iel) continue;
            size_t ij = jj < iel ? PAIR(iel, jj) : PAIR(jj, iel);
            w->r_ee_sav[jj] = w->r_ee[ij];
            for (int m = 0; m < 3; m++) w->rvec_ee_sav[3 * jj + m] = w->rvec_ee[3 * ij + m];
        }
    }
    for (int i = i1; i <= i2; i++)
        for (int ic = 0; ic < nc; ic++) {
            double *rv = w->rvec_en + ((size_t)i * nc + ic) * 3;
            double r = 0;
            for (int m = 0; m < 3; m++) rv[m] = x[3 * i + m] - s->cent[3 * ic + m];
            for (int m = 0; m < 3; m++) r = r + rv[m] * rv[m];
            w->r_en[(size_t)i * nc + ic] = sqrt(r);
        }
    if (iel < 0) {
        for (int i = 1; i < ne; i++)
            for (int j = 0; j < i; j++) {
                size_t ij = PAIR(i, j);
                double r = 0;
                for (int m = 0; m < 3; m++) w->rvec_ee[3 * ij + m] = x[3 * i + m] - x[3 * j + m];
                for (int m = 0; m < 3; m++) r = r + w->rvec_ee[3 * ij + m] * w->rvec_ee[3 * ij + m];
                w->r_ee[ij] = sqrt(r);
            }
    } else {
        for (int i = 0; i < iel; i++) {
            size_t ij = PAIR(iel, i);
            double r = 0;
            for (int m = 0; m < 3; m++) w->rvec_ee[3 * ij + m] = x[3 * iel + m] - x[3 * i + m];
            for (int m = 0; m < 3; m++) r = r + w->rvec_ee[3 * ij + m] * w->rvec_ee[3 * ij + m];
            w->r_ee[ij] = sqrt(r);
        }
        for (int i = iel + 1; i < ne; i++) {
            size_t ij = PAIR(i, iel);
            double r = 0;
            for (int m = 0; m < 3; m++) w->rvec_ee[3 * ij + m] = x[3 * i + m] - x[3 * iel + m];
            for (int m = 0; m < 3; m++) r = r + w->rvec_ee[3 * ij + m] * w->rvec_ee[3 * ij + m];
            w->r_ee[ij] = sqrt(r);
        }
    }
}

/* distancese_restore, distances.f90:169-211 */
static void distancese_restore(oracle_walker_t *w, int iel)
{
    const int ne = w->nelec, nc = w->ncent;
    for (int ic = 0; ic < nc; ic++) {
        w->r_en[(size_t)iel * nc + ic] = w->r_en_sav[ic];
        for (int m = 0; m < 3; m++) w->rvec_en[((size_t)iel * nc + ic) * 3 + m] = w->rvec_en_sav[3 * ic + m];
    }
    for (int jj = 0; jj < ne; jj++) {
        if (jj == iel) continue;
        size_t ij = jj < iel ? PAIR(iel, jj) : PAIR(jj, iel);
        w->r_ee[ij] = w->r_ee_sav[jj];
        for (int m = 0; m < 3; m++) w->rvec_ee[3 * ij + m] = w->rvec_ee_sav[3 * jj + m];
    }
}

/* ------------------------------------------------------------------------ */
/* e-e-n polynomial: the inner term loop shared by jastrow4 (jastrow4.f90:203-262)
 * and jastrow4e (jastrow4e.f90:204-239).  uu, rri, rrj are offset arrays with
 * valid indices -3..nordc and zeros below 0.                                 */
typedef struct { double fc, fu, fuu, fi, fii, fj, fjj, fui, fuj; } een_t;
static void een_terms(int nordc, const double *c, const double *uu, const double *rri, const double *rrj,
                      const double *ss, const double *tt, int want2, een_t *o)
{
    double fc = 0, fu = 0, fuu = 0, fi = 0, fii = 0, fj = 0, fjj = 0, fui = 0, fuj = 0;
    int ll = 0;
    for (int n = 2; n <= nordc; n++)
        for (int k = n - 1; k >= 0; k--) {
            int l_hi = (k == 0) ? n - k - 2 : n - k;
            for (int l = l_hi; l >= 0; l--) {
                int m = (n - k - l) / 2;
                if (2 * m != n - k - l) continue;
                const double cl = c[ll++];
                fc = fc + cl * uu[k] * ss[l] * tt[m];
                fu = fu + cl * k * uu[k - 1] * ss[l] * tt[m];
                if (want2) fuu = fuu + cl * k * (k - 1) * uu[k - 2] * ss[l] * tt[m];
                fi = fi + cl * uu[k] * ((l + m) * rri[l + m - 1] * rrj[m] + m * rri[m - 1] * rrj[l + m]);
                if (want2) fii = fii + cl * uu[k] * ((l + m) * (l + m - 1) * rri[l + m - 2] * rrj[m] + m * (m - 1) * rri[m - 2] * rrj[l + m]);
                fj = fj + cl * uu[k] * ((l + m) * rrj[l + m - 1] * rri[m] + m * rrj[m - 1] * rri[l + m]);
                if (want2) {
                    fjj = fjj + cl * uu[k] * ((l + m) * (l + m - 1) * rrj[l + m - 2] * rri[m] + m * (m - 1) * rrj[m - 2] * rri[l + m]);
                    fui = fui + cl * k * uu[k - 1] * ((l + m) * rri[l + m - 1] * rrj[m] + m * rri[m - 1] * rrj[l + m]);
                    fuj = fuj + cl * k * uu[k - 1] * ((l + m) * rrj[l + m - 1] * rri[m] + m * rrj[m - 1] * rri[l + m]);
                }
            }
        }
    o->fc = fc; o->fu = fu; o->fuu = fuu; o->fi = fi; o->fii = fii; o->fj = fj; o->fjj = fjj; o->fui = fui; o->fuj = fuj;
}

/* pair (i>j) e-e + e-e-n contribution: value, gradients fij (on i) and fji
 * (on j), Laplacian (want2).  Follows jastrow4.f90:79-290 / jastrow4e.f90:81-281. */
static void jas_pair(const champ_system_t *s, const oracle_walker_t *w, int i, int j, int want2,
                     double *fs, double fij[3], double fji[3], double *d2)
{
    const int ne = w->nelec, nc = w->ncent, nordc = s->nordc;
    const double kap = s->scalek, asymp_r = 1.0 / kap;
    const int nbp = s->nordb + 1, ncp = orc_nterms4(nordc);
    double uu_[CHAMP_MAX_NORD + 4] = {0}, ri_[CHAMP_MAX_NORD + 4] = {0}, rj_[CHAMP_MAX_NORD + 4] = {0};
    double ss_[CHAMP_MAX_NORD + 4] = {0}, tt_[CHAMP_MAX_NORD + 4] = {0};
    double *uu = uu_ + 3, *rri = ri_ + 3, *rrj = rj_ + 3, *ss = ss_ + 3, *tt = tt_ + 3;
    uu[0] = 1; ss[0] = 2; tt[0] = 1; rri[0] = 1; rrj[0] = 1;
    *fs = 0; *d2 = 0;
    for (int k = 0; k < 3; k++) fij[k] = fji[k] = 0;

    double sspinn = 1; int isb = 0, ipar = 0;
    if (i < s->nup || j >= s->nup) {
        if (s->nspin2b == 2) isb = 1;
        else if (s->nocuspb == 0) sspinn = 0.5;
        ipar = 1;
    }
    const size_t ij = PAIR(i, j);
    const double rij = w->r_ee[ij];
    const double *rv = w->rvec_ee + 3 * ij;
    double dd1, dd2;
    scale2(kap, rij, &uu[1], &dd1, &dd2);
    if (uu[1] != asymp_r) {
        const double *b = s->b + isb * nbp;
        double top = sspinn * b[0] * uu[1], topu = sspinn * b[0], topuu = 0;
        double bot = 1 + b[1] * uu[1], botu = b[1], botuu = 0, bot2 = bot * bot;
        double fee = top / bot - w->asymp_jasb[ipar];
        double feeu = topu / bot - botu * top / bot2;
        double feeuu = topuu - (botuu * top + 2 * botu * topu) / bot + 2 * botu * botu * top / bot2;
        feeuu = feeuu / bot;
        for (int iord = 2; iord <= s->nordb; iord++) {
            uu[iord] = uu[1] * uu[iord - 1];
            fee = fee + b[iord] * uu[iord];
            feeu = feeu + b[iord] * iord * uu[iord - 1];
            feeuu = feeuu + b[iord] * iord * (iord - 1) * uu[iord - 2];
        }
        feeuu = feeuu * dd1 * dd1 + feeu * dd2;
        feeu = feeu * dd1 / rij;
        *fs += fee;
        for (int k = 0; k < 3; k++) { fij[k] += feeu * rv[k]; fji[k] -= feeu * rv[k]; }
        if (want2) *d2 += 2 * (feeuu + 2 * feeu);
    }
    if (nordc <= 1) return;
    switch2(kap, &uu[1], &dd1, &dd2);
    for (int iord = 2; iord <= nordc; iord++) uu[iord] = uu[1] * uu[iord - 1];
    for (int ic = 0; ic < nc; ic++) {
        const int it = s->iwctype[ic];
        const double ri = w->r_en[(size_t)i * nc + ic], rj = w->r_en[(size_t)j * nc + ic];
        double dd7, dd8, dd9, dd10;
        scale2(kap, ri, &rri[1], &dd7, &dd9);
        scale2(kap, rj, &rrj[1], &dd8, &dd10);
        if (rri[1] == asymp_r && rrj[1] == asymp_r) continue;
        switch2(kap, &rri[1], &dd7, &dd9);
        switch2(kap, &rrj[1], &dd8, &dd10);
        for (int iord = 1; iord <= nordc; iord++) {
            rri[iord] = rri[1] * rri[iord - 1];
            rrj[iord] = rrj[1] * rrj[iord - 1];
            ss[iord] = rri[iord] + rrj[iord];
            tt[iord] = rri[iord] * rrj[iord];
        }
        een_t e;
        een_terms(nordc, s->c + (size_t)it * ncp, uu, rri, rrj, ss, tt, want2, &e);
        double fu = e.fu, fi = e.fi, fj = e.fj;
        if (want2) {
            double sv = ri + rj, t = ri - rj;
            double u2pst = rij * rij + sv * t, u2mst = rij * rij - sv * t;
            double fuu = e.fuu * dd1 * dd1 + fu * dd2;
            double fui = e.fui * dd1 * dd7, fuj = e.fuj * dd1 * dd8;
            double fii = e.fii * dd7 * dd7 + fi * dd9, fjj = e.fjj * dd8 * dd8 + fj * dd10;
            fu = fu * dd1 / rij; fi = fi * dd7 / ri; fj = fj * dd8 / rj;
            *d2 += 2 * (fuu + 2 * fu) + fui * u2pst / (ri * rij) + fuj * u2mst / (rj * rij) + fii + 2 * fi + fjj + 2 * fj;
        } else {
            fu = fu * dd1 / rij; fi = fi * dd7 / ri; fj = fj * dd8 / rj;
        }
        *fs += e.fc;
        const double *rvi = w->rvec_en + ((size_t)i * nc + ic) * 3, *rvj = w->rvec_en + ((size_t)j * nc + ic) * 3;
        for (int k = 0; k < 3; k++) {
            fij[k] = fij[k] + fi * rvi[k] + fu * rv[k];
            fji[k] = fji[k] + fj * rvj[k] - fu * rv[k];
        }
    }
    (void)ne;
}

/* e-n term for electron i: jastrow4.f90:300-370 / jastrow4e.f90:288-343 */
static void jas_en(const champ_system_t *s, const oracle_walker_t *w, int i, int want2, double *fs,
                   double fii_[3], double *d2)
{
    const int nc = w->ncent, nap = s->norda + 1;
    const double kap = s->scalek, asymp_r = 1.0 / kap;
    double rri[CHAMP_MAX_NORD + 2];
    *fs = 0; *d2 = 0; fii_[0] = fii_[1] = fii_[2] = 0;
    for (int ic = 0; ic < nc; ic++) {
        const int it = s->iwctype[ic];
        const double ri = w->r_en[(size_t)i * nc + ic];
        double dd7, dd9;
        scale2(kap, ri, &rri[1], &dd7, &dd9);
        if (rri[1] == asymp_r) continue;
        rri[0] = 1;
        const double *a4 = s->a4 + it * nap;
        double top = a4[0] * rri[1], topi = a4[0], topii = 0;
        double bot = 1 + a4[1] * rri[1], boti = a4[1], botii = 0, bot2 = bot * bot;
        double fen = top / bot - w->asymp_jasa[it];
        double feni = topi / bot - boti * top / bot2;
        double fenii = topii - (botii * top + 2 * boti * topi) / bot + 2 * boti * boti * top / bot2;
        fenii = fenii / bot;
        for (int iord = 2; iord <= s->norda; iord++) {
            rri[iord] = orc_powi(rri[1], iord);
            fen = fen + a4[iord] * rri[iord];
            feni = feni + a4[iord] * iord * rri[iord - 1];
            fenii = fenii + a4[iord] * iord * (iord - 1) * rri[iord - 2];
        }
        fenii = fenii * dd7 * dd7 + feni * dd9;
        feni = feni * dd7 / ri;
        *fs = *fs + fen;
        const double *rv = w->rvec_en + ((size_t)i * nc + ic) * 3;
        for (int k = 0; k < 3; k++) fii_[k] = fii_[k] + feni * rv[k];
        if (want2) *d2 = *d2 + fenii + 2 * feni;
    }
}

/* jastrow_factor4, jastrow4.f90:3-415 (iforce_analy = 0) */
static void jastrow_factor4(const champ_system_t *s, oracle_walker_t *w)
{
    const int ne = w->nelec;
    w->fsumo = 0; w->d2o = 0;
    for (int k = 0; k < 3 * ne; k++) w->fjo[k] = 0;
    for (int i = 1; i < ne; i++)
        for (int j = 0; j < i; j++) {
            double fs, fij[3], fji[3], d2;
            jas_pair(s, w, i, j, 1, &fs, fij, fji, &d2);
            w->fso[IJ(i, j)] = fs; w->d2ijo[IJ(i, j)] = d2;
            for (int k = 0; k < 3; k++) { w->fijo[IJ(i, j) * 3 + k] = fij[k]; w->fijo[IJ(j, i) * 3 + k] = fji[k]; }
            w->fsumo = w->fsumo + fs;
            for (int k = 0; k < 3; k++) { w->fjo[3 * i + k] += fij[k]; w->fjo[3 * j + k] += fji[k]; }
            w->d2o = w->d2o + d2;
        }
    for (int i = 0; i < ne; i++) {
        double fs, fii[3], d2;
        jas_en(s, w, i, 1, &fs, fii, &d2);
        w->fso[IJ(i, i)] = fs; w->d2ijo[IJ(i, i)] = d2;
        for (int k = 0; k < 3; k++) w->fijo[IJ(i, i) * 3 + k] = fii[k];
        w->fsumo = w->fsumo + fs;
        for (int k = 0; k < 3; k++) w->fjo[3 * i + k] += fii[k];
        w->d2o = w->d2o + d2;
    }
}

/* jastrow4e, jastrow4e.f90:3-358 */
static void jastrow4e(const champ_system_t *s, oracle_walker_t *w, int iel, int iflag)
{
    const int ne = w->nelec;
    for (int k = 0; k < 3 * ne; k++) w->fjn[k] = w->fjo[k];
    w->fsumn = w->fsumo;
    w->d2n = w->d2o;
    for (int jj = 0; jj < ne; jj++) {
        if (jj == iel) continue;
        int i = jj < iel ? iel : jj, j = jj < iel ? jj : iel;
        double fs, fij[3], fji[3], d2;
        jas_pair(s, w, i, j, iflag > 0, &fs, fij, fji, &d2);
        w->fsn[IJ(i, j)] = fs;
        w->d2ijn[IJ(i, j)] = iflag > 0 ? d2 : 0.0;
        for (int k = 0; k < 3; k++) { w->fijn[IJ(i, j) * 3 + k] = fij[k]; w->fijn[IJ(j, i) * 3 + k] = fji[k]; }
        w->fsumn = w->fsumn + w->fsn[IJ(i, j)] - w->fso[IJ(i, j)];
        for (int k = 0; k < 3; k++) {
            w->fjn[3 * i + k] = w->fjn[3 * i + k] + w->fijn[IJ(i, j) * 3 + k] - w->fijo[IJ(i, j) * 3 + k];
            w->fjn[3 * j + k] = w->fjn[3 * j + k] + w->fijn[IJ(j, i) * 3 + k] - w->fijo[IJ(j, i) * 3 + k];
        }
        w->d2n = w->d2n + w->d2ijn[IJ(i, j)] - w->d2ijo[IJ(i, j)];
    }
    double fs, fii[3], d2;
    jas_en(s, w, iel, iflag > 0, &fs, fii, &d2);
    w->fsn[IJ(iel, iel)] = fs;
    w->d2ijn[IJ(iel, iel)] = iflag > 0 ? d2 : 0.0;
    for (int k = 0; k < 3; k++) w->fijn[IJ(iel, iel) * 3 + k] = fii[k];
    w->fsumn = w->fsumn + w->fsn[IJ(iel, iel)] - w->fso[IJ(iel, iel)];
    for (int k = 0; k < 3; k++)
        w->fjn[3 * iel + k] = w->fjn[3 * iel + k] + w->fijn[IJ(iel, iel) * 3 + k] - w->fijo[IJ(iel, iel) * 3 + k];
    w->d2n = w->d2n + w->d2ijn[IJ(iel, iel)] - w->d2ijo[IJ(iel, iel)];
}

/* jassav, jassav.f90:3-78 (iflag = 0: no Laplacian caches) */
static void jassav(oracle_walker_t *w, int iel, int iflag)
{
    const int ne = w->nelec;
    w->fsumo = w->fsumn;
    for (int k = 0; k < 3 * ne; k++) w->fjo[k] = w->fjn[k];
    for (int j = 0; j <= iel; j++) w->fso[IJ(iel, j)] = w->fsn[IJ(iel, j)];
    for (int j = iel + 1; j < ne; j++) w->fso[IJ(j, iel)] = w->fsn[IJ(j, iel)];
    for (int j = 0; j < ne; j++)
        for (int k = 0; k < 3; k++) {
            w->fijo[IJ(iel, j) * 3 + k] = w->fijn[IJ(iel, j) * 3 + k];
            w->fijo[IJ(j, iel) * 3 + k] = w->fijn[IJ(j, iel) * 3 + k];
        }
    if (iflag > 0) {
        w->d2o = w->d2n;
        for (int j = 0; j <= iel; j++) w->d2ijo[IJ(iel, j)] = w->d2ijn[IJ(iel, j)];
        for (int j = iel + 1; j < ne; j++) w->d2ijo[IJ(j, iel)] = w->d2ijn[IJ(j, iel)];
    }
    for (int k = 0; k < 3 * ne; k++) w->vj[k] = w->vjn[k];
}

void orc_jastrow_full(const champ_system_t *s, oracle_walker_t *w, double *fsum, double *d2, double *fj)
{
    distances(s, w, -1, w->x);
    jastrow_factor4(s, w);
    *fsum = w->fsumo; *d2 = w->d2o;
    if (fj) memcpy(fj, w->fjo, sizeof(double) * 3 * w->nelec);
}

/* ------------------------------------------------------------------------ */
/* orbitals_no_qmckl (dense VECTORIZATION path, orbitals_no_qmckl.f90:73-128) */
static void orbitals_all(const champ_system_t *s, oracle_walker_t *w)
{
    const int ne = w->nelec, nc = w->ncent, nb = s->nbasis, no = s->norb;
    for (int i = 0; i < ne; i++) {
        orc_basis_fns_point(s, w->rvec_en + (size_t)i * nc * 3, w->r_en + (size_t)i * nc, 2, w->phin, w->dphin, w->d2phin);
        for (int iorb = 0; iorb < no; iorb++) {
            const double *cf = s->coef + (size_t)iorb * nb;
            double ao = 0, a1 = 0, a2 = 0, a3 = 0, add = 0;
            for (int m = 0; m < nb; m++) {
                ao = ao + cf[m] * w->phin[m];
                a1 = a1 + cf[m] * w->dphin[m];
                a2 = a2 + cf[m] * w->dphin[nb + m];
                a3 = a3 + cf[m] * w->dphin[2 * nb + m];
                add = add + cf[m] * w->d2phin[m];
            }
            w->orb[(size_t)i * no + iorb] = ao;
            w->dorb[((size_t)i * 3 + 0) * no + iorb] = a1;
            w->dorb[((size_t)i * 3 + 1) * no + iorb] = a2;
            w->dorb[((size_t)i * 3 + 2) * no + iorb] = a3;
            w->ddorb[(size_t)i * no + iorb] = add;
        }
    }
}

/* orbitalse_no_qmckl, orbitals_no_qmckl.f90:195-385 (dense) */
static void orbitalse(const champ_system_t *s, oracle_walker_t *w, int iel, int iflag)
{
    const int nc = w->ncent, nb = s->nbasis, no = s->norb;
    const int ider = iflag > 0 ? 2 : 1;
    orc_basis_fns_point(s, w->rvec_en + (size_t)iel * nc * 3, w->r_en + (size_t)iel * nc, ider, w->phin, w->dphin, w->d2phin);
    for (int iorb = 0; iorb < no; iorb++) {
        const double *cf = s->coef + (size_t)iorb * nb;
        double ao = 0, a1 = 0, a2 = 0, a3 = 0, add = 0;
        for (int m = 0; m < nb; m++) {
            ao = ao + cf[m] * w->phin[m];
            a1 = a1 + cf[m] * w->dphin[m];
            a2 = a2 + cf[m] * w->dphin[nb + m];
            a3 = a3 + cf[m] * w->dphin[2 * nb + m];
            if (iflag > 0) add = add + cf[m] * w->d2phin[m];
        }
        w->orbn[iorb] = ao; w->dorbn[iorb] = a1; w->dorbn[no + iorb] = a2; w->dorbn[2 * no + iorb] = a3;
        w->ddorbn[iorb] = add;
    }
}

/* determinant, determinant.f90:9-140 (single determinant, kref fixed) */
static int determinant(const champ_system_t *s, oracle_walker_t *w)
{
    const int no = s->norb, nup = w->nup;
    int bad = 0;
    orbitals_all(s, w);
    for (int iab = 0; iab < 2; iab++) {
        const int ish = iab ? nup : 0, nel = iab ? w->ndn : nup;
        double *slmi = w->slmi + (size_t)iab * w->nmax * w->nmax;
        w->detiab[iab] = 1.0;
        for (int j = 0; j < nel; j++) {
            const int jorb = s->iworbd[ish + j];
            for (int i = 0; i < nel; i++) {
                slmi[(size_t)j * nel + i] = w->orb[(size_t)(ish + i) * no + jorb];
                for (int k = 0; k < 3; k++) w->fp[((size_t)i * nel + j) * 3 + k] = w->dorb[((size_t)(ish + i) * 3 + k) * no + jorb];
                w->fpp[(size_t)i * nel + j] = w->ddorb[(size_t)(ish + i) * no + jorb];
            }
        }
        if (nel > 0 && orc_matinv(slmi, nel, &w->detiab[iab]) != 0) bad = 1;
        for (int i = 0; i < nel; i++) {
            double d[3] = {0, 0, 0}, dd = 0;
            for (int j = 0; j < nel; j++) {
                const double sv = slmi[(size_t)i * nel + j];
                for (int k = 0; k < 3; k++) d[k] = d[k] + sv * w->fp[((size_t)i * nel + j) * 3 + k];
                dd = dd + sv * w->fpp[(size_t)i * nel + j];
            }
            for (int k = 0; k < 3; k++) w->ddx[3 * (i + ish) + k] = d[k];
            w->d2dx2[i + ish] = dd;
        }
    }
    return bad;
}

/* determinante, determinante.f90:3-67 */
static void determinante(const champ_system_t *s, oracle_walker_t *w, int iel, int iflag)
{
    orbitalse(s, w, iel, iflag);
    const int nup = w->nup;
    const int iab = iel < nup ? 0 : 1, nel = iab ? w->ndn : nup, ish = iab ? nup : 0;
    const double *slmi = w->slmi + (size_t)iab * w->nmax * w->nmax;
    const size_t ikel = (size_t)nel * (iel - ish);
    double ratio = 0.0;
    for (int j = 0; j < nel; j++) ratio = ratio + slmi[j + ikel] * w->orbn[s->iworbd[j + ish]];
    w->ratio_kref = ratio;
    w->detn = w->detiab[iab] * ratio;
    if (ratio == 0.0) return;
    for (int i = 0; i < nel; i++) {
        if (i + ish == iel) continue;
        const size_t ik = (size_t)nel * i;
        double sum = 0;
        for (int j = 0; j < nel; j++) sum = sum + slmi[j + ik] * w->orbn[s->iworbd[j + ish]];
        sum = sum / ratio;
        for (int j = 0; j < nel; j++) w->slmin[j + ik] = slmi[j + ik] - slmi[j + ikel] * sum;
    }
    for (int j = 0; j < nel; j++) w->slmin[j + ikel] = slmi[j + ikel] / ratio;
}

static void md_grad(const champ_system_t *s, const oracle_walker_t *w, int iel, int iflag_move, double g[3]);

/* compute_determinante_grad, determinante.f90:69-283, 285-328 */
void orc_determinante_grad(const champ_system_t *s, oracle_walker_t *w, int iel, int iflag_move, double vd[3])
{
    const int nup = w->nup, no = s->norb;
    const int iab = iel < nup ? 0 : 1, nel = iab ? w->ndn : nup, ish = iab ? nup : 0;
    const size_t ik = (size_t)(iel - ish) * nel;
    double ref[3] = {0, 0, 0};
    if (s->ndet > 1) {
        md_grad(s, w, iel, iflag_move, ref);
        const double *vj = iflag_move == 1 ? w->vj : w->vjn;
        for (int k = 0; k < 3; k++) vd[k] = 0.0 + ref[k] + vj[3 * iel + k];
        return;
    }
    if (iflag_move == 1) {
        const double *slmi = w->slmi + (size_t)iab * w->nmax * w->nmax;
        for (int j = 0; j < nel; j++) {
            const int jo = s->iworbd[j + ish];
            for (int k = 0; k < 3; k++) ref[k] = ref[k] + slmi[j + ik] * w->dorb[((size_t)iel * 3 + k) * no + jo];
        }
        for (int k = 0; k < 3; k++) vd[k] = 0.0 + ref[k] + w->vj[3 * iel + k];
    } else {
        for (int j = 0; j < nel; j++) {
            const int jo = s->iworbd[j + ish];
            for (int k = 0; k < 3; k++) ref[k] = ref[k] + w->slmin[j + ik] * w->dorbn[k * no + jo];
        }
        for (int k = 0; k < 3; k++) vd[k] = 0.0 + ref[k] + w->vjn[3 * iel + k];
    }
}

/* detsav, detsav.f90:3-94 (single det) */
static void detsav(const champ_system_t *s, oracle_walker_t *w, int iel)
{
    const int nup = w->nup, no = s->norb;
    const int iab = iel < nup ? 0 : 1, nel = iab ? w->ndn : nup;
    double *slmi = w->slmi + (size_t)iab * w->nmax * w->nmax;
    memcpy(slmi, w->slmin, sizeof(double) * nel * nel);
    w->detiab[iab] = w->detn;
    for (int iorb = 0; iorb < no; iorb++) {
        w->orb[(size_t)iel * no + iorb] = w->orbn[iorb];
        for (int k = 0; k < 3; k++) w->dorb[((size_t)iel * 3 + k) * no + iorb] = w->dorbn[k * no + iorb];
    }
}

/* ------------------------------------------------------------------------ */
/* Multideterminant psi_D = sum_k cdet_k D_up,k D_dn,k (determinant_psit.f90:
 * 3-26) by brute force: every determinant k is built in its own orbital order
 * iworbd(:,k) and inverted on its own, so neither the excitation tables nor
 * their phase convention (set_input_data.f90:46-414) enter.  The reference
 * reaches the same psi, drift and local energy through aa = S^-1 Phi and the
 * small excitation determinants (multideterminant.f90:3-255,
 * multideterminante.f90:3-222).  Row i of a determinant is linear in
 * phi(r_i), so every one-electron quantity is d ln psi_D / d phi_o(r_i):
 *   vmd[i][o] = sum_k wt_k sum_{j: iworbd(j,k) = o} inv_k[j][i],
 *   wt_k = cdet_k D_up,k D_dn,k / psi_D,
 * and grad_i psi_D / psi_D = sum_o vmd[i][o] grad phi_o(r_i), likewise the
 * Laplacian and the ECP quadrature ratio sum_o vmd[i][o] phi_o(x_q).        */
static double md_build_spin(const champ_system_t *s, const oracle_walker_t *w, int k, int iab, int iel,
                            const double *orbrow, double *inv)
{
    const int no = s->norb, nup = w->nup, nel = iab ? w->ndn : nup, ish = iab ? nup : 0;
    const int32_t *iw = s->iworbd + (size_t)k * w->nelec + ish;
    for (int j = 0; j < nel; j++)
        for (int i = 0; i < nel; i++)
            inv[(size_t)j * nel + i] = (i + ish == iel) ? orbrow[iw[j]] : w->orb[(size_t)(ish + i) * no + iw[j]];
    double det = 1.0;
    if (nel > 0 && orc_matinv(inv, nel, &det) != 0) det = 0.0;
    return det;
}

static double md_weights_sum(const champ_system_t *s, const double *dup, int sup, const double *ddn, int sdn)
{
    double psid = 0.0;
    for (int k = 0; k < s->ndet; k++) psid = psid + dup[(size_t)k * sup] * ddn[(size_t)k * sdn] * s->cdet[k];
    return psid;
}

/* all determinants at the current configuration: psi_D, ddx, d2dx2, vmd */
static double md_hpsi(const champ_system_t *s, oracle_walker_t *w)
{
    const int ne = w->nelec, no = s->norb, nup = w->nup, nm = w->nmax;
    for (int k = 0; k < s->ndet; k++)
        for (int iab = 0; iab < 2; iab++)
            w->md_det[2 * k + iab] = md_build_spin(s, w, k, iab, -1, NULL, w->md_inv + ((size_t)k * 2 + iab) * nm * nm);
    const double psid = md_weights_sum(s, w->md_det, 2, w->md_det + 1, 2);
    memset(w->vmd, 0, sizeof(double) * ne * no);
    for (int k = 0; k < s->ndet; k++) {
        const double wt = s->cdet[k] * w->md_det[2 * k] * w->md_det[2 * k + 1] / psid;
        for (int iab = 0; iab < 2; iab++) {
            const int nel = iab ? w->ndn : nup, ish = iab ? nup : 0;
            const double *inv = w->md_inv + ((size_t)k * 2 + iab) * nm * nm;
            const int32_t *iw = s->iworbd + (size_t)k * ne + ish;
            for (int i = 0; i < nel; i++)
                for (int j = 0; j < nel; j++) w->vmd[(size_t)(ish + i) * no + iw[j]] += wt * inv[(size_t)i * nel + j];
        }
    }
    for (int i = 0; i < ne; i++) {
        double d[3] = {0, 0, 0}, dd = 0;
        for (int o = 0; o < no; o++) {
            const double v = w->vmd[(size_t)i * no + o];
            for (int k = 0; k < 3; k++) d[k] = d[k] + v * w->dorb[((size_t)i * 3 + k) * no + o];
            dd = dd + v * w->ddorb[(size_t)i * no + o];
        }
        for (int k = 0; k < 3; k++) w->ddx[3 * i + k] = d[k];
        w->d2dx2[i] = dd;
    }
    return psid;
}

/* electron iel moved (orbitals orbn/dorbn at the new position): every
 * determinant of its spin rebuilt; returns the new psi_D */
static double md_psie(const champ_system_t *s, oracle_walker_t *w, int iel)
{
    const int nm = w->nmax, iab = iel < w->nup ? 0 : 1;
    for (int k = 0; k < s->ndet; k++)
        w->md_detn[k] = md_build_spin(s, w, k, iab, iel, w->orbn, w->md_invn + (size_t)k * nm * nm);
    w->md_psidn = iab == 0 ? md_weights_sum(s, w->md_detn, 1, w->md_det + 1, 2)
                           : md_weights_sum(s, w->md_det, 2, w->md_detn, 1);
    return w->md_psidn;
}

/* grad ln psi_D of electron iel: current (iflag_move=1) or proposed (0) */
static void md_grad(const champ_system_t *s, const oracle_walker_t *w, int iel, int iflag_move, double g[3])
{
    const int ne = w->nelec, no = s->norb, nm = w->nmax, nup = w->nup;
    const int iab = iel < nup ? 0 : 1, nel = iab ? w->ndn : nup, ish = iab ? nup : 0, row = iel - ish;
    const double psid = iflag_move ? md_weights_sum(s, w->md_det, 2, w->md_det + 1, 2) : w->md_psidn;
    g[0] = g[1] = g[2] = 0.0;
    for (int k = 0; k < s->ndet; k++) {
        const double dme = iflag_move ? w->md_det[2 * k + iab] : w->md_detn[k];
        const double wt = s->cdet[k] * dme * w->md_det[2 * k + 1 - iab] / psid;
        const double *inv = iflag_move ? w->md_inv + ((size_t)k * 2 + iab) * nm * nm : w->md_invn + (size_t)k * nm * nm;
        const int32_t *iw = s->iworbd + (size_t)k * ne + ish;
        for (int j = 0; j < nel; j++) {
            const double v = wt * inv[(size_t)row * nel + j];
            for (int c = 0; c < 3; c++)
                g[c] = g[c] + v * (iflag_move ? w->dorb[((size_t)iel * 3 + c) * no + iw[j]] : w->dorbn[c * no + iw[j]]);
        }
    }
}

static void md_accept(const champ_system_t *s, oracle_walker_t *w, int iel)
{
    const int nm = w->nmax, iab = iel < w->nup ? 0 : 1;
    for (int k = 0; k < s->ndet; k++) {
        memcpy(w->md_inv + ((size_t)k * 2 + iab) * nm * nm, w->md_invn + (size_t)k * nm * nm, sizeof(double) * nm * nm);
        w->md_det[2 * k + iab] = w->md_detn[k];
    }
}

/* ------------------------------------------------------------------------ */
/* getvps_gauss, readps_gauss.f90:6-55 */
static void getvps_gauss(const champ_system_t *s, oracle_walker_t *w, int iel)
{
    const int nc = w->ncent;
    for (int ic = 0; ic < nc; ic++) {
        const int ict = s->iwctype[ic];
        double r = w->r_en[(size_t)iel * nc + ic];
        if (r < 1.0e-10) r = 1.0e-10;
        double ri = 1.0 / r, vpot = 0, dvpot = 0;
        const int lp = s->lpot[ict];
        orc_gauss_pot(s, r, lp, ict, &vpot, &dvpot);
        vpot = vpot - s->znuc[ict] * ri;
        w->vps[((size_t)iel * nc + ic) * CHAMP_MPS_L + (lp - 1)] = vpot;
        for (int l = 1; l <= lp - 1; l++) {
            vpot = 0; dvpot = 0;
            orc_gauss_pot(s, r, l, ict, &vpot, &dvpot);
            w->vps[((size_t)iel * nc + ic) * CHAMP_MPS_L + (l - 1)] = vpot;
        }
    }
}

/* yl0, nonloc.f90:317-340 (l is the 1-based component index) */
double orc_yl0(int l, double costh);
static double yl0(int l, double costh) { return orc_yl0(l, costh); }
double orc_yl0(int l, double costh)
{
    if (l == 1) return 1.0;
    if (l == 2) return 3.0 * costh;
    if (l == 3) return 2.5 * (3 * costh * costh - 1);
    if (l == 4) return 3.5 * (5 * costh * costh * costh - 3 * costh);
    return 0.0;
}

/* psinl / psianl / psibnl, nonlpsi.f90:4-182 (ijas = 4) */
static double psinl(const champ_system_t *s, double u, double rri, double rrj, int it)
{
    const double kap = s->scalek, asymp_r = 1.0 / kap;
    const int nordc = s->nordc, ncp = orc_nterms4(nordc);
    double uu[CHAMP_MAX_NORD + 1], ri[CHAMP_MAX_NORD + 1], rj[CHAMP_MAX_NORD + 1], ss[CHAMP_MAX_NORD + 1], tt[CHAMP_MAX_NORD + 1];
    if (rri == asymp_r && rrj == asymp_r) return 0.0;
    uu[1] = 1 - kap * u; ri[1] = 1 - kap * rri; rj[1] = 1 - kap * rrj;
    uu[0] = 1; ri[0] = 1; rj[0] = 1; ss[0] = 2; tt[0] = 1;
    for (int jp = 1; jp <= nordc; jp++) {
        uu[jp] = uu[1] * uu[jp - 1];
        ri[jp] = ri[1] * ri[jp - 1];
        rj[jp] = rj[1] * rj[jp - 1];
        ss[jp] = ri[jp] + rj[jp];
        tt[jp] = ri[jp] * rj[jp];
    }
    const double *c = s->c + (size_t)it * ncp;
    double psi = 0.0;
    int ll = 0;
    for (int n = 2; n <= nordc; n++)
        for (int k = n - 1; k >= 0; k--) {
            int l_hi = (k == 0) ? n - k - 2 : n - k;
            for (int l = l_hi; l >= 0; l--) {
                int m = (n - k - l) / 2;
                if (2 * m == n - k - l) { psi = psi + c[ll] * uu[k] * ss[l] * tt[m]; ll++; }
            }
        }
    return psi;
}
static double psianl(const champ_system_t *s, const double *ja, double rri, int it)
{
    const double asymp_r = 1.0 / s->scalek;
    const double *a4 = s->a4 + it * (s->norda + 1);
    if (rri == asymp_r) return 0.0;
    double psi = a4[0] * rri / (1.0 + a4[1] * rri) - ja[it];
    double rrip = rri;
    for (int i = 2; i <= s->norda; i++) { rrip = rrip * rri; psi = psi + a4[i] * rrip; }
    return psi;
}
static double psibnl(const champ_system_t *s, const double *jb, double u, int isb, int ipar, double sspinn)
{
    const double asymp_r = 1.0 / s->scalek;
    const double *b = s->b + isb * (s->nordb + 1);
    if (u == asymp_r) return 0.0;
    double fee = b[0] * u / (1 + b[1] * u);
    double psi = sspinn * fee - jb[ipar];
    double uu = u;
    for (int i = 2; i <= s->nordb; i++) { uu = u * uu; psi = psi + b[i] * uu; }
    return psi;
}

/* nonloc_pot (nonloc_pot.f90:3-108) + nonloc (nonloc.f90:5-315) with
 * nonlocd_quad (:462-506) and nonlocj_quad4 (:656-899); single determinant. */
static void nonloc_pot(const champ_system_t *s, oracle_walker_t *w, double *pe)
{
    const int ne = w->nelec, nc = w->ncent, nb = s->nbasis, no = s->norb, nup = w->nup;
    const double kap = s->scalek;
    for (int i = 0; i < ne; i++) getvps_gauss(s, w, i);
    for (int ic = 0; ic < nc; ic++)
        for (int i = 0; i < ne; i++)
            *pe = *pe + w->vps[((size_t)i * nc + ic) * CHAMP_MPS_L + (s->lpot[s->iwctype[ic]] - 1)];

    w->vpsp_det[0] = w->vpsp_det[1] = 0.0;
    w->nxquad = 0;
    memset(w->tvpsp, 0, sizeof(double) * (size_t)ne * nc * (s->nquad > 0 ? s->nquad : 1));
    double *rr_en = (double *)malloc(sizeof(double) * ne * nc);   /* scale_dist of r_en */
    for (int ic = 0; ic < nc; ic++)
        for (int i = 0; i < ne; i++) rr_en[(size_t)i * nc + ic] = (1.0 - exp(-kap * w->r_en[(size_t)i * nc + ic])) / kap;
    double *rvq = (double *)malloc(sizeof(double) * 3 * nc), *rq = (double *)malloc(sizeof(double) * nc);
    double *rrq = (double *)malloc(sizeof(double) * nc);
    double *phi = (double *)malloc(sizeof(double) * nb), *orbq = (double *)malloc(sizeof(double) * no);

    for (int i = 0; i < ne; i++) {
        const int iab = i < nup ? 0 : 1, nel = iab ? w->ndn : nup, ish = iab ? nup : 0;
        const double *slmi = w->slmi + (size_t)iab * w->nmax * w->nmax;
        for (int ic = 0; ic < nc; ic++) {
            const int ict = s->iwctype[ic], lp = s->lpot[ict];
            int iskip = 1;
            for (int l = 1; l <= lp - 1; l++)
                if (fabs(w->vps[((size_t)i * nc + ic) * CHAMP_MPS_L + (l - 1)]) > 1.0e-4) iskip = 0;
            if (iskip) continue;
            const double ren = w->r_en[(size_t)i * nc + ic], ri = 1.0 / ren;
            const double *rv = w->rvec_en + ((size_t)i * nc + ic) * 3;
            for (int iq = 0; iq < s->nquad; iq++) {
                w->nxquad++;
                double costh = rv[0] * w->xq[iq] + rv[1] * w->yq[iq] + rv[2] * w->zq[iq];
                costh = costh * ri;
                double xquad[3] = {ren * w->xq[iq] + s->cent[3 * ic + 0], ren * w->yq[iq] + s->cent[3 * ic + 1],
                                   ren * w->zq[iq] + s->cent[3 * ic + 2]};
                for (int jc = 0; jc < nc; jc++) {
                    double r2 = 0;
                    for (int k = 0; k < 3; k++) rvq[3 * jc + k] = xquad[k] - s->cent[3 * jc + k];
                    for (int k = 0; k < 3; k++) r2 = r2 + rvq[3 * jc + k] * rvq[3 * jc + k];
                    rq[jc] = sqrt(r2);
                }
                rq[ic] = ren;
                /* orbitals_quad (ider=0) + nonlocd_quad */
                orc_basis_fns_point(s, rvq, rq, 0, phi, NULL, NULL);
                for (int iorb = 0; iorb < no; iorb++) {
                    const double *cf = s->coef + (size_t)iorb * nb;
                    double ao = 0;
                    for (int m = 0; m < nb; m++) ao = ao + cf[m] * phi[m];
                    orbq[iorb] = ao;
                }
                const size_t ikel = (size_t)nel * (i - ish);
                double det_ratio = 0.0;
                if (s->ndet > 1)
                    for (int o = 0; o < no; o++) det_ratio = det_ratio + w->vmd[(size_t)i * no + o] * orbq[o];
                else
                    for (int j = 0; j < nel; j++) det_ratio = det_ratio + slmi[j + ikel] * orbq[s->iworbd[j + ish]];
                /* nonlocj_quad4 */
                for (int jc = 0; jc < nc; jc++) rrq[jc] = (1.0 - exp(-kap * rq[jc])) / kap;
                double fsumn = 0;
                for (int jj = 0; jj < ne; jj++) {
                    if (jj == i) continue;
                    int ii = jj < i ? i : jj, j = jj < i ? jj : i;
                    double sspinn = 1; int ipar = 0, isb = 0;
                    if (ii < nup || j >= nup) {
                        if (s->nspin2b == 2) isb = 1;
                        else if (s->nocuspb == 0) sspinn = 0.5;
                        ipar = 1;
                    }
                    double dx[3], rij = 0;
                    for (int k = 0; k < 3; k++) dx[k] = w->x[3 * jj + k] - xquad[k];
                    for (int k = 0; k < 3; k++) rij = rij + dx[k] * dx[k];
                    rij = sqrt(rij);
                    double u = (1.0 - exp(-kap * rij)) / kap;
                    double fsn = psibnl(s, w->asymp_jasb, u, isb, ipar, sspinn);
                    if (s->nordc > 1)
                        for (int jc = 0; jc < nc; jc++)
                            fsn = fsn + psinl(s, u, rrq[jc], rr_en[(size_t)jj * nc + jc], s->iwctype[jc]);
                    fsumn = fsumn + fsn - w->fso[(size_t)ii * ne + j];
                }
                double fsd = 0;
                for (int jc = 0; jc < nc; jc++) fsd = fsd + psianl(s, w->asymp_jasa, rrq[jc], s->iwctype[jc]);
                fsumn = fsumn + fsd - w->fso[(size_t)i * ne + i];
                /* accumulate, nonloc.f90:248-274 */
                double term_radial = 0.0;
                for (int l = 1; l <= lp - 1; l++)
                    term_radial = term_radial + yl0(l, costh) * w->vps[((size_t)i * nc + ic) * CHAMP_MPS_L + (l - 1)];
                double trj = term_radial * s->wq[iq] * exp(fsumn);
                w->vpsp_det[iab] = w->vpsp_det[iab] + trj * det_ratio;
                w->tvpsp[((size_t)i * nc + ic) * s->nquad + iq] = trj * det_ratio;
            }
        }
    }
    free(rr_en); free(rvq); free(rq); free(rrq); free(phi); free(orbq);
}

/* pot_local, pot_local.f90:3-78 */
static double pot_local(const champ_system_t *s, const oracle_walker_t *w)
{
    const int ne = w->nelec, nc = w->ncent;
    double pe = w->pecent;
    if (s->nloc == 0)
        for (int i = 0; i < ne; i++)
            for (int ic = 0; ic < nc; ic++) pe = pe + (-s->znuc[s->iwctype[ic]] / w->r_en[(size_t)i * nc + ic]);
    for (int i = 1; i < ne; i++)
        for (int j = 0; j < i; j++) pe = pe + 1 / w->r_ee[PAIR(i, j)];
    return pe;
}

/* hpsi, hpsi.f90:3-273 (one state) */
void orc_hpsi(const champ_system_t *s, oracle_walker_t *w, double *psid, double *psij, double *ekin, double *energy)
{
    const int ne = w->nelec, nup = w->nup;
    distances(s, w, -1, w->x);
    double pe_local = pot_local(s, w);
    jastrow_factor4(s, w);
    for (int k = 0; k < 3 * ne; k++) w->vj[k] = w->fjo[k];
    const double d2j = w->d2o, pj = w->fsumo;
    int bad = determinant(s, w);
    const double psid_md = s->ndet > 1 ? md_hpsi(s, w) : 0.0;
    if (s->nloc > 0) nonloc_pot(s, w, &pe_local);
    else { w->vpsp_det[0] = w->vpsp_det[1] = 0.0; w->nxquad = 0; }
    double ekin_det[2], eloc_det[2];
    for (int iab = 0; iab < 2; iab++) {
        const int nel = iab ? w->ndn : nup, ish = iab ? nup : 0;
        ekin_det[iab] = 0.0;
        for (int i = 0; i < nel; i++) {
            const double *v = w->vj + 3 * (i + ish), *d = w->ddx + 3 * (i + ish);
            double tmp = -HB * (w->d2dx2[i + ish] + 2.0 * (v[0] * d[0] + v[1] * d[1] + v[2] * d[2]));
            ekin_det[iab] = ekin_det[iab] + tmp;
        }
        eloc_det[iab] = ekin_det[iab] + w->vpsp_det[iab];
    }
    double ekin_other = -HB * d2j;
    for (int i = 0; i < ne; i++) {
        const double *v = w->vj + 3 * i;
        ekin_other = ekin_other + (-HB * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]));
    }
    double e_other = ekin_other + pe_local;
    double pd = s->ndet > 1 ? psid_md : 0.0 + w->detiab[0] * w->detiab[1] * s->cdet[0];
    double en = 0.0 + eloc_det[0] + eloc_det[1] + e_other;
    double ek = 0.0 + ekin_det[0] + ekin_det[1] + ekin_other;
    w->psid = pd; w->psij = pj; w->ekin = ek; w->energy = en; w->pe_local = pe_local; w->d2j = d2j;
    w->psido = pd; w->psijo = pj; w->eold = en;
    w->psi2o = 2 * (log(fabs(pd)) + pj);
    (void)bad;
    if (psid) *psid = pd;
    if (psij) *psij = pj;
    if (ekin) *ekin = ek;
    if (energy) *energy = en;
}

/* psie, hpsie.f90:3-79 */
void orc_psie(const champ_system_t *s, oracle_walker_t *w, int iel, const double xnew[3], double *psidn,
              double *psijn, int iflag)
{
    const int ne = w->nelec;
    double *xx = (double *)malloc(sizeof(double) * 3 * ne);
    memcpy(xx, w->x, sizeof(double) * 3 * ne);
    for (int k = 0; k < 3; k++) xx[3 * iel + k] = xnew[k];
    memcpy(w->xnew_e, xnew, sizeof(double) * 3);
    distances(s, w, iel, xx);
    free(xx);
    jastrow4e(s, w, iel, iflag);
    for (int k = 0; k < 3 * ne; k++) w->vjn[k] = w->fjn[k];
    *psijn = w->fsumn;
    determinante(s, w, iel, iflag);
    if (s->ndet > 1) { *psidn = md_psie(s, w, iel); return; }
    if (w->detn == 0.0) { *psidn = 0.0; return; }
    const int iab = iel < w->nup ? 0 : 1;
    *psidn = 0.0 + w->detn * w->detiab[1 - iab] * s->cdet[0];
}

void orc_accept(const champ_system_t *s, oracle_walker_t *w, int iel)
{
    for (int k = 0; k < 3; k++) w->x[3 * iel + k] = w->xnew_e[k];
    jassav(w, iel, 0);
    if (s->ndet > 1) md_accept(s, w, iel);
    detsav(s, w, iel);
}
void orc_reject(const champ_system_t *s, oracle_walker_t *w, int iel) { (void)s; distancese_restore(w, iel); }

/* metrop1, metrop_mov1_driftdif.f90:4-363 (iguiding=0, node_cutoff=0,
 * nforce=1, one state) preceded by rotqua when nloc > 0 (vmc.f90:113).      */
int orc_metrop1_sweep(const champ_system_t *s, oracle_walker_t *w, double tau)
{
    const int ne = w->nelec;
    const double adrift = 0.5, rttau = sqrt(tau);
    int acc = 0;
    if (s->nloc > 0) orc_rotqua(s, w->rng, w->xq, w->yq, w->zq);
    for (int i = 0; i < ne; i++) {
        double vold[3], vnew[3], xnew[3], xold[3];
        for (int k = 0; k < 3; k++) xold[k] = w->x[3 * i + k];
        orc_determinante_grad(s, w, i, 1, vold);
        double v2old = vold[0] * vold[0] + vold[1] * vold[1] + vold[2] * vold[2];
        double vavvt = (sqrt(1.0 + 2.0 * adrift * v2old * tau) - 1.0) / (adrift * v2old);
        double dfus2o = 0.0;
        for (int k = 0; k < 3; k++) {
            double drift = vavvt * vold[k] * tau;
            double dfus = orc_gauss(w->rng) * rttau;
            double dx = drift + dfus;
            dfus2o = dfus2o + dfus * dfus;
            xnew[k] = xold[k] + dx;
        }
        double psidn, psijn;
        orc_psie(s, w, i, xnew, &psidn, &psijn, 0);
        orc_determinante_grad(s, w, i, 0, vnew);
        double psi2n = 2 * (log(fabs(psidn)) + psijn);
        double v2new = vnew[0] * vnew[0] + vnew[1] * vnew[1] + vnew[2] * vnew[2];
        vavvt = (sqrt(1.0 + 2.0 * adrift * v2new * tau) - 1.0) / (adrift * v2new);
        double dfus2n = 0.0;
        for (int k = 0; k < 3; k++) {
            double drift = vavvt * vnew[k] * tau;
            double xbac = xnew[k] + drift;
            double dfus = xbac - xold[k];
            dfus2n = dfus2n + dfus * dfus;
        }
        double p = exp(psi2n - w->psi2o) * exp(-0.5 * (dfus2n - dfus2o) / tau);
        if (psidn == 0.0) p = 0.0;
        p = p < 1.0 ? p : 1.0;
        if (orc_random_dp(w->rng) < p) {
            w->psi2o = psi2n;
            w->psido = psidn; w->psijo = psijn;
            orc_accept(s, w, i);
            acc++;
        } else {
            orc_reject(s, w, i);
        }
    }
    double psid, psij, ekin, energy;
    orc_hpsi(s, w, &psid, &psij, &ekin, &energy);
    return acc;
}

/* ------------------------------------------------------------------------ */
/* gammai (gammai.f90:3-108): incomplete gamma function of a = 3/2 or 5/2 with */
/* xae = x^a e^-x given; a fixed number of series terms for x < a+4 (returns */
/* gamma(a,x), iflag 1), else a fixed number of continued-fraction steps     */
/* (returns -Gamma(a,x), the complement, iflag -1).                         */
double orc_gammai(double a, double x, double xae, int *iflag)
{
    if (x < a + 4) {
        *iflag = 1;
        double ap = a, sum = 1 / a, del = sum;
        const int iter = (int)(21 + 3 * (x - a));
        for (int n = 1; n <= iter; n++) {
            ap = ap + 1;
            del = del * x / ap;
            sum = sum + del;
        }
        return sum * xae;
    }
    *iflag = -1;
    double a0 = 1, a1 = x, b0 = 0, b1 = 1, fac = 1, g = 0;
    const double t = 20 + round(a - x);  /* nint: half away from zero */
    const int iter = t > 7 ? (int)t : 7;
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

/* Normalisation of metrop6's radial transition density (metrop_mov1_slat.f90:
 * 256-293 forward, 490-527 reverse): the integral of
 * sqrt(r) |1 + co r| exp(-zeta r) over [rbot, rtop] divided by bot, from the
 * incomplete gamma functions, split at the zero r = -1/co of 1 + co r.       */
double orc_m6_area(double zeta, double co, double rbot, double rtop, double bot)
{
    const double d5b2 = 2.5, d2b3 = .666666666666667, g5b2 = 1.329340388179137;
    int iflagb, iflagt, iflagz;
    const double rzero = -1 / co;
    const double zrbot = zeta * rbot, zrtop = zeta * rtop;
    const double zebot = sqrt(orc_powi(zrbot, 3)) * exp(-zrbot);
    const double zetop = sqrt(orc_powi(zrtop, 3)) * exp(-zrtop);
    const double g52bot = orc_gammai(d5b2, zrbot, zrbot * zebot, &iflagb);
    const double g52top = orc_gammai(d5b2, zrtop, zrtop * zetop, &iflagt);
    if (rzero < rbot || rzero > rtop) {
        const double g52dif = iflagb * iflagt == 1 ? g52top - g52bot : g52top - g52bot + g5b2;
        const double g32dif = d2b3 * (g52dif + zetop - zebot);
        return fabs(g32dif + co * g52dif / zeta) / (bot * sqrt(orc_powi(zeta, 3)));
    }
    const double zrzer = zeta * rzero;
    const double zezer = sqrt(orc_powi(zrzer, 3)) * exp(-zrzer);
    const double g52zer = orc_gammai(d5b2, zrzer, zrzer * zezer, &iflagz);
    const double g52dif1 = iflagb * iflagz == 1 ? g52zer - g52bot : g52zer - g52bot + g5b2;
    const double g32dif1 = d2b3 * (g52dif1 + zezer - zebot);
    const double g52dif2 = iflagt * iflagz == 1 ? g52top - g52zer : g52top - g52zer + g5b2;
    const double g32dif2 = d2b3 * (g52dif2 + zetop - zezer);
    return (fabs(g32dif1 + co * g52dif1 / zeta) + fabs(g32dif2 + co * g52dif2 / zeta)) /
           (bot * sqrt(orc_powi(zeta, 3)));
}

/* zeta and co of the Slater-type radial density from the radial velocity   */
/* (metrop_mov1_slat.f90:204-218 forward, 473-486 reverse)                  */
static void m6_zeta_co(double zcusp, double vr, double rmin, double *zeta, double *co)
{
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
    *zeta = z;
    *co = (z + vr) / (1 - (z + vr) * rmin);
}

static double m6_thetamx(double r, double z, double deltat)
{
    return deltat + (2 - deltat) / (1 + orc_powi(z * r, 2));
}

/* nearest centre to point x (first index on ties) and its squared distance */
static int m6_nearest(const champ_system_t *s, const double x[3], double *d2)
{
    int best = 0;
    double dmin = 99.e9;
    for (int j = 0; j < s->ncent; j++) {
        double dist = 0;
        for (int k = 0; k < 3; k++) dist = dist + orc_powi(x[k] - s->cent[3 * j + k], 2);
        if (dist < dmin) { dmin = dist; best = j; }
    }
    *d2 = dmin;
    return best;
}

/* One metrop6 sweep (metrop_mov1_slat.f90:3-738): the accelerated Metropolis
 * move of every electron in local spherical coordinates about its nearest
 * nucleus, iguiding = 0, node_cutoff = 0, one state; then hpsi.  The nearest
 * nucleus, rvmino and rmino of electron i are functions of its position
 * (dumper_more.f90:271-286 at start; nearestn/rvminn taken over on
 * acceptance), so they are recomputed here rather than carried.  The
 * rejection loops of the radial and azimuthal samplings are capped at 10^6
 * trials (never reached in practice); the walker then stays put.  Returns
 * the accepted moves. */
int orc_metrop6_sweep(const champ_system_t *s, oracle_walker_t *w, double deltar, double deltat, double fbias)
{
    const double eps = 1.e-10, pi = 4.0 * atan(1.0);
    const double deltri = 1 / deltar;
    const int ne = w->nelec;
    int acc = 0;
    if (s->nloc > 0) orc_rotqua(s, w->rng, w->xq, w->yq, w->zq);
    for (int i = 0; i < ne; i++) {
        double xold[3], xnew[3], vold[3], vnew[3], xaxis[3], yaxis[3], zaxis[3], rvmino[3], rvminn[3], rvminon[3];
        for (int k = 0; k < 3; k++) xold[k] = w->x[3 * i + k];
        orc_determinante_grad(s, w, i, 1, vold);
        double d2;
        const int nearo = m6_nearest(s, xold, &d2);
        for (int k = 0; k < 3; k++) rvmino[k] = xold[k] - s->cent[3 * nearo + k];
        const double rmino = sqrt(orc_powi(rvmino[0], 2) + orc_powi(rvmino[1], 2) + orc_powi(rvmino[2], 2));
        const double znear = s->znuc[s->iwctype[nearo]];
        double zcusp = s->nloc == 0 ? znear : 0;
        double fxop = 1;
        double rbot = rmino * deltri, rtop = rmino * deltar;
        double voldr = 0;
        for (int k = 0; k < 3; k++) voldr = voldr + vold[k] * rvmino[k];
        voldr = voldr / rmino;
        double voldp = 0;
        for (int k = 0; k < 3; k++) {
            zaxis[k] = rvmino[k] / rmino;
            xaxis[k] = vold[k] - voldr * zaxis[k];
            voldp = voldp + orc_powi(xaxis[k], 2);
        }
        voldp = sqrt(voldp);
        if (voldp < eps) {
            xaxis[0] = eps * (1 - orc_powi(zaxis[0], 2));
            xaxis[1] = eps * (-zaxis[0] * zaxis[1]);
            xaxis[2] = eps * (-zaxis[0] * zaxis[2]);
            voldp = eps * sqrt(1 - orc_powi(zaxis[0], 2));
        }
        for (int k = 0; k < 3; k++) xaxis[k] = xaxis[k] / voldp;
        voldr = fmax(voldr, -2 * znear);
        yaxis[0] = zaxis[1] * xaxis[2] - zaxis[2] * xaxis[1];
        yaxis[1] = zaxis[2] * xaxis[0] - zaxis[0] * xaxis[2];
        yaxis[2] = zaxis[0] * xaxis[1] - zaxis[1] * xaxis[0];
        voldp = voldp * fbias;
        double zeta, co;
        m6_zeta_co(zcusp, voldr, rmino, &zeta, &co);
        /* maximum of the radial density on [rbot, rtop] for rejection sampling */
        double root = sqrt(orc_powi(1.5 * co - zeta, 2) + 2 * zeta * co);
        double rmax1 = (1.5 * co - zeta + root) / (2 * zeta * co);
        if (rmax1 < rbot) rmax1 = rbot;
        if (rmax1 > rtop) rmax1 = rtop;
        double fmx = sqrt(rmax1) * fabs(1 + co * rmax1) * exp(-zeta * rmax1);
        if (co < 0) {
            double rmax2 = (1.5 * co - zeta - root) / (2 * zeta * co);
            if (rmax2 < rbot) rmax2 = rbot;
            if (rmax2 > rtop) rmax2 = rtop;
            const double fmax2 = sqrt(rmax2) * fabs(1 + co * rmax2) * exp(-zeta * rmax2);
            fmx = fmx > fmax2 ? fmx : fmax2;
        }
        const double bot = sqrt(rmino) * fabs(1 + co * rmino) * exp(-zeta * rmino);
        double rtry = 0, top = 0;
        int ok = 0;
        for (int trial = 0; trial < 1000000 && !ok; trial++) {
            rtry = ((deltar - deltri) * orc_random_dp(w->rng) + deltri) * rmino;
            top = sqrt(rtry) * fabs(1 + co * rtry) * exp(-zeta * rtry);
            ok = top / fmx > orc_random_dp(w->rng);
        }
        fxop = fxop * top / bot;
        double areao = orc_m6_area(zeta, co, rbot, rtop, bot);
        /* cos(theta), then phi by rejection */
        const double raver = 0.5 * (rmino + rtry);
        const double deltt = m6_thetamx(raver, znear, deltat);
        double costht = 1 - deltt * orc_random_dp(w->rng);
        const double zprime = rtry * costht;
        double sintht = sqrt(1 - costht * costht);
        areao = areao * deltt;
        double term = fmin(voldp * raver * sintht, 1.0);
        fmx = 1 + term;
        double phitry = 0, cosphi = 0;
        int okp = 0;
        for (int trial = 0; trial < 1000000 && !okp; trial++) {
            phitry = pi * orc_random_dp(w->rng);
            cosphi = cos(phitry);
            top = 1 + term * cosphi;
            okp = top > orc_random_dp(w->rng) * fmx;
        }
        fxop = fxop * top;
        const double xprime = rtry * sintht * cos(phitry);
        double yprime = sqrt(fmax(0.0, rtry * rtry - orc_powi(xprime, 2) - orc_powi(zprime, 2)));
        if (orc_random_dp(w->rng) < 0.5) yprime = -yprime;
        for (int k = 0; k < 3; k++) xnew[k] = xaxis[k] * xprime + yaxis[k] * yprime + zaxis[k] * zprime + s->cent[3 * nearo + k];
        const double rminno = rtry;
        /* geometric rejection: the reverse move must be a possible forward move */
        const int nearn = m6_nearest(s, xnew, &d2);
        const double rminn = sqrt(d2);
        double rminon = 0, dot = 0;
        for (int k = 0; k < 3; k++) {
            rvminn[k] = xnew[k] - s->cent[3 * nearn + k];
            rvminon[k] = xold[k] - s->cent[3 * nearn + k];
            rminon = rminon + orc_powi(rvminon[k], 2);
            dot = dot + rvminn[k] * rvminon[k];
        }
        rminon = sqrt(rminon);
        dot = dot / (rminn * rminon);
        costht = dot;
        sintht = sqrt(1 - costht * costht);
        const double ravern = 0.5 * (rminn + rminon);
        const double znearn = s->znuc[s->iwctype[nearn]];
        const double delttn = m6_thetamx(ravern, znearn, deltat);
        double p = 0, psidn = 0, psijn = 0, psi2n = 0;
        const int geo = rminon > rminn * deltar || dot < 1 - delttn;
        if (!(ok && okp)) {
            p = 0;  /* a rejection loop hit its cap: the walker stays */
        } else if (!geo) {
            const double rratio = rminno / rminon;
            orc_psie(s, w, i, xnew, &psidn, &psijn, 0);
            orc_determinante_grad(s, w, i, 0, vnew);
            psi2n = 2 * (log(fabs(psidn)) + psijn);
            /* reverse transition */
            double fxnp = 1;
            rbot = rminn * deltri;
            rtop = rminn * deltar;
            double vnewr = 0;
            for (int k = 0; k < 3; k++) vnewr = vnewr + vnew[k] * rvminn[k];
            vnewr = vnewr / rminn;
            double vnewp = 0;
            for (int k = 0; k < 3; k++) {
                xaxis[k] = vnew[k] - vnewr * rvminn[k] / rminn;
                vnewp = vnewp + orc_powi(xaxis[k], 2);
            }
            vnewp = sqrt(vnewp);
            if (vnewp < eps) {
                xaxis[0] = eps * (1 - orc_powi(rvminn[0], 2));
                xaxis[1] = eps * (-rvminn[0] * rvminn[1]);
                xaxis[2] = eps * (-rvminn[0] * rvminn[2]);
                vnewp = eps * sqrt(1 - orc_powi(rvminn[0], 2) * (2 - orc_powi(rminn, 2)));
            }
            for (int k = 0; k < 3; k++) xaxis[k] = xaxis[k] / vnewp;
            zcusp = s->nloc == 0 ? znearn : 0;
            vnewr = fmax(vnewr, -2 * znearn);
            vnewp = vnewp * fbias;
            m6_zeta_co(zcusp, vnewr, rminn, &zeta, &co);
            const double botn = sqrt(rminn) * fabs(1 + co * rminn) * exp(-zeta * rminn);
            const double topn = sqrt(rminon) * fabs(1 + co * rminon) * exp(-zeta * rminon);
            fxnp = fxnp * topn / botn;
            double arean = orc_m6_area(zeta, co, rbot, rtop, botn);
            arean = arean * delttn;
            term = fmin(vnewp * ravern * sintht, 1.0);
            cosphi = 0;
            double rnorm = 0;
            for (int k = 0; k < 3; k++) {
                const double term2 = rvminon[k] / rminon - costht * rvminn[k] / rminn;
                rnorm = rnorm + term2 * term2;
                cosphi = cosphi + term2 * xaxis[k];
            }
            cosphi = cosphi / sqrt(rnorm);
            fxnp = fxnp * (1 + term * cosphi);
            p = orc_powi(rratio, 2) * exp(psi2n - w->psi2o) * fabs((fxnp * areao) / (fxop * arean));
            if (psidn == 0.0) p = 0.0;
            /* min(1, p); a NaN p rejects, as the reference's dmin1 (minsd) keeps it */
            p = p < 1.0 ? p : (p >= 1.0 ? 1.0 : 0.0);
        }
        if (orc_random_dp(w->rng) < p) {
            w->psi2o = psi2n;
            w->psido = psidn; w->psijo = psijn;
            orc_accept(s, w, i);
            acc++;
        } else if (!geo && ok && okp) {
            orc_reject(s, w, i);
        }
    }
    double psid, psij, ekin, energy;
    orc_hpsi(s, w, &psid, &psij, &ekin, &energy);
    return acc;
}

/* ------------------------------------------------------------------------ */
/* DMC (dmc.f90 with dmc_ps, dmc_ps_mov1.f90:4-838, or dmc_ae,           */
/* dmc_ae_mov1.f90, when nloc = 0) at the default controls                   */
/* idmc = 2, icasula = 0, ipq = 1, itau_eff = 1, iacc_rej = 1, icross = 1,  */
/* icut_br = 0, icut_e = 0, nforce = 1, node_cutoff = 0.  One stream per    */
/* walker slot (moves and quadrature rotation) and one population stream    */
/* for splitj; a copied walker keeps its destination slot's stream.         */

/* deep copy of one walker's state except its stream */
static void walker_copy_state(oracle_walker_t *d, const oracle_walker_t *s, const champ_system_t *sy)
{
    const int ne = s->nelec, nc = s->ncent, nm = s->nmax, nb = s->nbasis, no = s->norb, np = s->npair;
#define CP(f, n) memcpy(d->f, s->f, sizeof(double) * (size_t)(n))
    CP(x, 3 * ne); CP(xnew_e, 3); CP(r_en, ne * nc); CP(rvec_en, 3 * ne * nc); CP(r_ee, np); CP(rvec_ee, 3 * np);
    CP(r_en_sav, nc); CP(rvec_en_sav, 3 * nc); CP(r_ee_sav, ne); CP(rvec_ee_sav, 3 * ne);
    CP(fjo, 3 * ne); CP(fjn, 3 * ne); CP(fso, ne * ne); CP(fsn, ne * ne); CP(fijo, 3 * ne * ne); CP(fijn, 3 * ne * ne);
    CP(d2ijo, ne * ne); CP(d2ijn, ne * ne); CP(vj, 3 * ne); CP(vjn, 3 * ne); CP(orb, ne * no); CP(dorb, 3 * ne * no);
    CP(ddorb, ne * no); CP(orbn, no); CP(dorbn, 3 * no); CP(ddorbn, no); CP(slmi, 2 * nm * nm); CP(slmin, nm * nm);
    CP(fp, 3 * nm * nm); CP(fpp, nm * nm); CP(ddx, 3 * ne); CP(d2dx2, ne); CP(phin, nb); CP(dphin, 3 * nb);
    CP(d2phin, nb); CP(vps, ne * nc * CHAMP_MPS_L);
    if (s->ndet > 1) {
        CP(md_inv, s->ndet * 2 * nm * nm); CP(md_det, s->ndet * 2); CP(md_invn, s->ndet * nm * nm);
        CP(md_detn, s->ndet); CP(vmd, ne * no);
    }
#undef CP
    (void)sy;
    memcpy(d->xq, s->xq, sizeof(d->xq)); memcpy(d->yq, s->yq, sizeof(d->yq)); memcpy(d->zq, s->zq, sizeof(d->zq));
    d->fsumo = s->fsumo; d->d2o = s->d2o; d->fsumn = s->fsumn; d->d2n = s->d2n;
    d->detiab[0] = s->detiab[0]; d->detiab[1] = s->detiab[1]; d->detn = s->detn;
    d->psid = s->psid; d->psij = s->psij; d->ekin = s->ekin; d->energy = s->energy; d->pe_local = s->pe_local;
    d->d2j = s->d2j; d->vpsp_det[0] = s->vpsp_det[0]; d->vpsp_det[1] = s->vpsp_det[1]; d->nxquad = s->nxquad;
    d->psido = s->psido; d->psijo = s->psijo; d->psi2o = s->psi2o; d->eold = s->eold; d->ratio_kref = s->ratio_kref;
    d->md_psidn = s->md_psidn;
}

struct orc_dmc {
    int cap, nwalk, nconf, nfprod, ipass, first;
    int icasula;  /* 0: locality approximation; -1: one T-move per walker after the step */
    double tau, etrial, eigv, eest, fprod, wdsumo;
    double *ff, *wtgen, *wt, *fratio;
    int *age;
    uint64_t pop[4];
    oracle_walker_t **w;
    double wgsum, egsum, tausum, wgcum, egcum, taucum, wgcm2, egcm2;
    long long nbrnch;
};

/* dmc_eloc_cutoff for nloc > 0 (branching.f90:226-238), tratio = 1 */
static double dmc_fratio(const oracle_walker_t *w, double tau)
{
    const double adrift = 0.5;
    double vav2sum = 0, v2sum = 0;
    for (int i = 0; i < w->nelec; i++) {
        double v[3];
        for (int k = 0; k < 3; k++) v[k] = w->ddx[3 * i + k] + w->vj[3 * i + k];
        const double v2 = orc_powi(v[0], 2) + orc_powi(v[1], 2) + orc_powi(v[2], 2);
        const double vavvt = (sqrt(1.0 + 2.0 * adrift * v2 * tau) - 1.0) / (adrift * v2);
        const double vavvn = vavvt / tau;
        vav2sum = vav2sum + orc_powi(vavvn, 2) * v2;
        v2sum = v2sum + v2;
    }
    return sqrt(vav2sum / v2sum);
}

/* ---- all-electron mover (dmc_ae_mov1.f90, dmc.f90:98-102 for nloc = 0) ---- */

typedef struct {
    int ic;
    double rmin, vr, va, vp, v2, zeta, zeff, xaxis[3], zaxis[3];
} ae_frame_t;

/* frame at x with drift v: nearest nucleus, radial and perpendicular drift
   (dmc_ae_mov1.f90:226-270, :388-440) */
static void ae_frame(const champ_system_t *s, double tau, const double x[3], const double v[3], ae_frame_t *f)
{
    const double eps = 1.e-10;
    double ren2mn = 1.e100;
    f->ic = 0;
    for (int ic = 0; ic < s->ncent; ic++) {
        const double ren2 = orc_powi(x[0] - s->cent[3 * ic], 2) + orc_powi(x[1] - s->cent[3 * ic + 1], 2) +
                            orc_powi(x[2] - s->cent[3 * ic + 2], 2);
        if (ren2 < ren2mn) { ren2mn = ren2; f->ic = ic; }
    }
    double rv[3], rmin = 0, vr = 0, v2 = 0;
    for (int k = 0; k < 3; k++) {
        rv[k] = x[k] - s->cent[3 * f->ic + k];
        rmin = rmin + rv[k] * rv[k];
        vr = vr + v[k] * rv[k];
        v2 = v2 + v[k] * v[k];
    }
    rmin = sqrt(rmin);
    f->rmin = rmin; f->vr = vr / rmin; f->v2 = v2; f->va = sqrt(v2);
    f->zeff = s->znuc[s->iwctype[f->ic]];
    f->zeta = sqrt(1.0 / tau + f->zeff * f->zeff);
    double vp = 0;
    for (int k = 0; k < 3; k++) {
        f->zaxis[k] = rv[k] / rmin;
        f->xaxis[k] = v[k] - f->vr * f->zaxis[k];
        vp = vp + f->xaxis[k] * f->xaxis[k];
    }
    vp = sqrt(vp);
    if (vp < eps) {
        f->xaxis[0] = eps * (1 - f->zaxis[0] * f->zaxis[0]);
        f->xaxis[1] = eps * (-f->zaxis[0] * f->zaxis[1]);
        f->xaxis[2] = eps * (-f->zaxis[0] * f->zaxis[2]);
        vp = eps * sqrt(1 + eps - f->zaxis[0] * f->zaxis[0]);
    }
    for (int k = 0; k < 3; k++) f->xaxis[k] = f->xaxis[k] / vp;
    f->vp = vp;
}

/* adrift near the nucleus and vavvt (dmc_ae_mov1.f90:273-277, adrift0 = 0.1) */
static double ae_vavvt(const ae_frame_t *f, double tau)
{
    const double eps = 1.e-10, adrift0 = 0.1;
    const double hafzr2 = orc_powi(0.5 * f->zeff * f->rmin, 2);
    const double adrift = (0.5 * (1 + eps + f->vr / f->va)) + adrift0 * hafzr2 / (1 + hafzr2);
    return (sqrt(1 + 2 * adrift * f->v2 * tau) - 1) / (adrift * f->v2);
}

static double ae_term(double tau)
{
    const double pi = 4 * atan(1.0);
    return orc_powi(sqrt(2 * pi * tau), 3) / pi;
}

/* forward move (dmc_ae_mov1.f90:272-353): Gaussian around the drifted point
   with probability pgaus, else exponential around the nucleus */
static void ae_propose(const champ_system_t *s, double tau, const double xo[3], const double vo[3], uint64_t rng[4],
                       double xn[3], double *dfus2o, double *fnormo)
{
    const double pi = 4 * atan(1.0), rttau = sqrt(tau);
    ae_frame_t f;
    ae_frame(s, tau, xo, vo, &f);
    const double *c = s->cent + 3 * f.ic;
    const double vavvt = ae_vavvt(&f, tau);
    double rtry = f.rmin + vavvt * f.vr, qgaus;
    if (rtry > 0) {
        qgaus = 0.5 * erfc(rtry / sqrt(2 * tau));
        const double xprime = vavvt * f.vp * rtry / (0.5 * (f.rmin + rtry)), zprime = rtry;
        for (int k = 0; k < 3; k++) xn[k] = c[k] + f.xaxis[k] * xprime + f.zaxis[k] * zprime;
    } else {
        qgaus = 0.5 * (1 + erf(-rtry / sqrt(2 * tau)));
        for (int k = 0; k < 3; k++) xn[k] = c[k];
    }
    const double pgaus = 1 - qgaus;
    double dfus2a = 0, dfusb;
    if (orc_random_dp(rng) < pgaus) {
        double dfus2b = 0;
        for (int k = 0; k < 3; k++) {
            const double dfus = orc_gauss(rng) * rttau;
            dfus2a = dfus2a + dfus * dfus;
            xn[k] = xn[k] + dfus;
            dfus2b = dfus2b + orc_powi(xn[k] - c[k], 2);
        }
        dfusb = sqrt(dfus2b);
    } else {
        const double u1 = orc_random_dp(rng), u2 = orc_random_dp(rng), u3 = orc_random_dp(rng);
        dfusb = (-0.5 / f.zeta) * log(u1 * u2 * u3);
        const double costht = 2 * (orc_random_dp(rng) - 0.5), sintht = sqrt(1 - costht * costht);
        const double phi = 2 * pi * orc_random_dp(rng);
        const double dfus3[3] = {dfusb * sintht * cos(phi), dfusb * sintht * sin(phi), dfusb * costht};
        for (int k = 0; k < 3; k++) {
            dfus2a = dfus2a + orc_powi(c[k] + dfus3[k] - xn[k], 2);
            xn[k] = c[k] + dfus3[k];
        }
    }
    *dfus2o = dfus2a;
    *fnormo = pgaus + qgaus * ae_term(tau) * orc_powi(f.zeta, 3) * exp(-2 * f.zeta * dfusb + 0.5 * dfus2a / tau);
}

/* reverse Green function from xn back to xo (dmc_ae_mov1.f90:388-480) */
static void ae_reverse(const champ_system_t *s, double tau, const double xo[3], const double xn[3], const double vn[3],
                       double *dfus2n, double *fnormn)
{
    ae_frame_t f;
    ae_frame(s, tau, xn, vn, &f);
    const double *c = s->cent + 3 * f.ic;
    const double vavvt = ae_vavvt(&f, tau);
    double rtry = f.rmin + vavvt * f.vr, qgaus, dfus2a = 0, dfus2b = 0;
    if (rtry > 0) {
        qgaus = 0.5 * erfc(rtry / sqrt(2 * tau));
        const double xprime = vavvt * f.vp * rtry / (0.5 * (f.rmin + rtry)), zprime = rtry;
        for (int k = 0; k < 3; k++) {
            const double xbac = c[k] + f.xaxis[k] * xprime + f.zaxis[k] * zprime;
            dfus2b = dfus2b + orc_powi(c[k] - xo[k], 2);
            dfus2a = dfus2a + orc_powi(xbac - xo[k], 2);
        }
    } else {
        qgaus = 0.5 * (1 + erf(-rtry / sqrt(2 * tau)));
        for (int k = 0; k < 3; k++) dfus2b = dfus2b + orc_powi(c[k] - xo[k], 2);
        dfus2a = dfus2b;
    }
    *dfus2n = dfus2a;
    const double pgaus = 1 - qgaus, dfusb = sqrt(dfus2b);
    *fnormn = pgaus + qgaus * ae_term(tau) * orc_powi(f.zeta, 3) * exp(-2 * f.zeta * dfusb + 0.5 * dfus2a / tau);
}

/* dmc_eloc_cutoff for nloc = 0 (branching.f90:240-278), adrift = adrift0 = 0.1, tratio = 1 */
static double dmc_fratio_ae(const champ_system_t *s, const oracle_walker_t *w, double tau)
{
    double vav2sum = 0, v2sum = 0;
    for (int i = 0; i < w->nelec; i++) {
        double v[3], x[3];
        for (int k = 0; k < 3; k++) { v[k] = w->ddx[3 * i + k] + w->vj[3 * i + k]; x[k] = w->x[3 * i + k]; }
        ae_frame_t f;
        ae_frame(s, tau, x, v, &f);
        const double vavvn = ae_vavvt(&f, tau) / tau;
        vav2sum = vav2sum + vavvn * vavvn * f.v2;
        v2sum = v2sum + f.v2;
    }
    return sqrt(vav2sum / v2sum);
}

orc_dmc_t *orc_dmc_new(const champ_system_t *s, int cap, const double *x0, const uint64_t *rng0, int nconf, double tau,
                       double etrial, int nfprod, const uint64_t pop[4])
{
    orc_dmc_t *d = (orc_dmc_t *)xcalloc(1, sizeof(*d));
    d->cap = cap; d->nwalk = nconf; d->nconf = nconf; d->nfprod = nfprod; d->tau = tau; d->etrial = etrial;
    d->eigv = 1; d->eest = etrial; d->fprod = 1; d->wdsumo = nconf; d->first = 1;
    d->ff = xcalloc(nfprod + 1, 8); d->wtgen = xcalloc(nfprod + 1, 8);
    for (int i = 0; i <= nfprod; i++) { d->ff[i] = 1; d->wtgen[i] = nconf; }
    d->wt = xcalloc(cap, 8); d->fratio = xcalloc(cap, 8); d->age = (int *)xcalloc(cap, sizeof(int));
    memcpy(d->pop, pop, 32);
    d->w = (oracle_walker_t **)xcalloc(cap, sizeof(*d->w));
    for (int iw = 0; iw < cap; iw++) {
        d->w[iw] = orc_walker_new(s);
        orc_walker_set_x(d->w[iw], x0 + (size_t)iw * 3 * s->nelec);
        orc_walker_set_rng(d->w[iw], rng0 + 4 * (size_t)iw);
    }
    for (int iw = 0; iw < nconf; iw++) {
        double a, b, c2, e;
        d->wt[iw] = 1;
        orc_hpsi(s, d->w[iw], &a, &b, &c2, &e);
    }
    return d;
}

void orc_dmc_free(orc_dmc_t *d)
{
    if (!d) return;
    for (int iw = 0; iw < d->cap; iw++) orc_walker_free(d->w[iw]);
    free(d->w); free(d->ff); free(d->wtgen); free(d->wt); free(d->fratio); free(d->age);
    free(d);
}

/* One T-move of the walker after its step (icasula = -1): nonloc_grid with
 * all electrons (nonloc_grid.f90:60-137), heat-bath choice among the
 * quadrature points with negative t_vpsp in the order (electron, point,
 * centre), tauprim = tau, then psiedmc + jassav + detsav of the chosen
 * electron at its point (dmc_ps_mov1.f90:701-730). */
static void dmc_tmove(const champ_system_t *s, oracle_walker_t *w, double tau)
{
    const int ne = w->nelec, nc = w->ncent, nq = s->nquad;
    double t_norm = 0.0;
    for (int i = 0; i < ne; i++)
        for (int iq = 0; iq < nq; iq++)
            for (int ic = 0; ic < nc; ic++) {
                double t = w->tvpsp[((size_t)i * nc + ic) * nq + iq];
                if (t > 0.0) t = 0.0;
                t_norm = t_norm - t;
            }
    t_norm = 1.0 + t_norm * tau;
    if (t_norm == 1.0) return;
    const double t_normi = 1.0 / t_norm;
    const double p = orc_random_dp(w->rng);
    double t_cum = 0.0;
    for (int i = 0; i < ne; i++)
        for (int iq = 0; iq < nq; iq++)
            for (int ic = 0; ic < nc; ic++) {
                double t = w->tvpsp[((size_t)i * nc + ic) * nq + iq];
                if (t > 0.0) t = 0.0;
                t_cum = t_cum - tau * t * t_normi;
                if (t_cum > p) {
                    const double ren = w->r_en[(size_t)i * nc + ic];
                    const double q[3] = {w->xq[iq], w->yq[iq], w->zq[iq]};
                    double x[3], psidn, psijn;
                    for (int k = 0; k < 3; k++) x[k] = ren * q[k] + s->cent[3 * ic + k];
                    orc_psie(s, w, i, x, &psidn, &psijn, 0);
                    w->psido = psidn; w->psijo = psijn;
                    w->psi2o = 2 * (log(fabs(psidn)) + psijn);
                    orc_accept(s, w, i);
                    return;
                }
            }
}

void orc_dmc_set_icasula(orc_dmc_t *d, int icasula) { d->icasula = icasula; }

/* splitj (splitj.f90:3-233) with deep copies of the walker states */
static int dmc_splitj(const champ_system_t *s, orc_dmc_t *d)
{
    int *iwundr = (int *)xcalloc(d->cap + 1, sizeof(int));
    int iunder = 0, ipair = 0, iw2 = 0;
    for (int iw = 0; iw < d->nwalk; iw++) {
        if (d->wt[iw] < 0.5) {
            if (d->wt[iw] == 0) {
                d->nbrnch++;
                iwundr[iunder++] = iw;
            } else if (ipair == 0) {
                ipair = 1;
                iw2 = iw;
            } else {
                d->nbrnch++;
                ipair = 0;
                const double wttot = d->wt[iw] + d->wt[iw2];
                if (orc_random_dp(d->pop) > d->wt[iw] / wttot) {
                    d->wt[iw2] = wttot;
                    iwundr[iunder++] = iw;
                } else {
                    d->wt[iw] = wttot;
                    iwundr[iunder++] = iw2;
                }
            }
        }
    }
    int nwalk2 = d->nwalk;
    for (int iw = 0; iw < d->nwalk; iw++) {
        if (d->wt[iw] >= 2) {
            d->nbrnch++;
            int dst;
            if (iunder > 0) dst = iwundr[--iunder];
            else {
                dst = nwalk2++;
                if (nwalk2 > d->cap) { free(iwundr); return -1; }
            }
            d->wt[iw] = d->wt[iw] * 0.5;
            d->wt[dst] = d->wt[iw];
            d->age[dst] = d->age[iw];
            d->fratio[dst] = d->fratio[iw];
            walker_copy_state(d->w[dst], d->w[iw], s);
        }
    }
    for (int j = iunder - 1; j >= 0; j--) {
        const int dst = iwundr[j], src = nwalk2 - 1;
        nwalk2--;
        d->wt[dst] = d->wt[src];
        d->age[dst] = d->age[src];
        d->fratio[dst] = d->fratio[src];
        walker_copy_state(d->w[dst], d->w[src], s);
    }
    free(iwundr);
    d->nwalk = nwalk2;
    return nwalk2;
}

/* one DMC step of the population; out: wsum1, esum1, wgsum1, egsum1, eest,
 * eigv, fprod, nodecr.  Returns the new walker count (-1: capacity).      */
int orc_dmc_step(const champ_system_t *s, orc_dmc_t *d, double out[8])
{
    const double adrift = 0.5, tau = d->tau, rttau = sqrt(tau);
    const int ne = s->nelec;
    d->ipass++;
    const int ipmod = d->ipass % d->nfprod;
    const double ginv = tau < 1.0 ? tau : 1.0;
    const double ffn = d->eigv * pow(d->wdsumo / d->nconf, ginv), ffi = 1 / ffn;
    d->fprod = d->fprod * ffn / d->ff[ipmod];
    d->ff[ipmod] = ffn;
    if (d->first) {
        for (int iw = 0; iw < d->nwalk; iw++)
            d->fratio[iw] = s->nloc > 0 ? dmc_fratio(d->w[iw], tau) : dmc_fratio_ae(s, d->w[iw], tau);
        d->first = 0;
    }
    double wsum1 = 0, esum1 = 0, nod = 0;
    for (int iw = 0; iw < d->nwalk; iw++) {
        oracle_walker_t *w = d->w[iw];
        if (s->nloc > 0) orc_rotqua(s, w->rng, w->xq, w->yq, w->zq);
        double dfus2ac = 0, dfus2un = 0;
        int iaccept = 0;
        for (int i = 0; i < ne; i++) {
            double vold[3], vnew[3], xold[3], xnew[3];
            for (int k = 0; k < 3; k++) xold[k] = w->x[3 * i + k];
            orc_determinante_grad(s, w, i, 1, vold);
            double dfus2o = 0, fnormo = 1, fnormn = 1;
            if (s->nloc == 0) {
                ae_propose(s, tau, xold, vold, w->rng, xnew, &dfus2o, &fnormo);
            } else {
                const double v2old = orc_powi(vold[0], 2) + orc_powi(vold[1], 2) + orc_powi(vold[2], 2);
                const double vavvt = (sqrt(1.0 + 2.0 * adrift * v2old * tau) - 1.0) / (adrift * v2old);
                for (int k = 0; k < 3; k++) {
                    const double drift = vavvt * vold[k];
                    const double dfus = orc_gauss(w->rng) * rttau;
                    const double dx = drift + dfus;
                    dfus2o = dfus2o + orc_powi(dfus, 2);
                    xnew[k] = xold[k] + dx;
                }
            }
            double psidn, psijn;
            orc_psie(s, w, i, xnew, &psidn, &psijn, 0);
            orc_determinante_grad(s, w, i, 0, vnew);
            if (psidn * w->psido <= 0) nod++;
            double dfus2n = 0;
            if (s->nloc == 0) {
                ae_reverse(s, tau, xold, xnew, vnew, &dfus2n, &fnormn);
            } else {
                const double v2new = orc_powi(vnew[0], 2) + orc_powi(vnew[1], 2) + orc_powi(vnew[2], 2);
                const double vavvt = (sqrt(1.0 + 2.0 * adrift * v2new * tau) - 1.0) / (adrift * v2new);
                for (int k = 0; k < 3; k++) {
                    const double drift = vavvt * vnew[k];
                    const double dfus = xnew[k] + drift - xold[k];
                    dfus2n = dfus2n + orc_powi(dfus, 2);
                }
            }
            double p = orc_powi(psidn / w->psido, 2) * exp(2 * (psijn - w->psijo)) * exp((dfus2o - dfus2n) / (2 * tau)) *
                       (fnormn / fnormo);
            if (d->age[iw] > 50) p = p * pow(1.1, d->age[iw] - 50);
            p = p < 1.0 ? p : (p >= 1.0 ? 1.0 : 0.0);
            dfus2ac = dfus2ac + p * dfus2o;
            dfus2un = dfus2un + dfus2o;
            if (orc_random_dp(w->rng) < p) {
                iaccept = 1;
                d->age[iw] = 0;
                w->psi2o = 2 * (log(fabs(psidn)) + psijn);
                w->psido = psidn; w->psijo = psijn;
                orc_accept(s, w, i);
            } else {
                orc_reject(s, w, i);
            }
        }
        const double taunow = tau * dfus2ac / dfus2un;
        const double eprev = w->eold;
        double psid, psij, ekin, enew;
        orc_hpsi(s, w, &psid, &psij, &ekin, &enew);
        const double fration = s->nloc > 0 ? dmc_fratio(w, tau) : dmc_fratio_ae(s, w, tau);
        const double deo = d->eest - eprev, den = d->eest - enew;
        const double ewto = d->eest - deo * d->fratio[iw], ewtn = d->eest - den * fration;
        const double expon = (d->etrial - 0.5 * (ewto + ewtn)) * taunow;
        const double dwt = exp(expon) * ffi;
        d->wt[iw] = d->wt[iw] * dwt;
        const double wtnow = d->wt[iw], wtg = wtnow * d->fprod;
        d->tausum = d->tausum + wtg * taunow;
        d->fratio[iw] = fration;
        if (!iaccept) d->age[iw]++;
        if (d->icasula < 0 && s->nloc > 0) dmc_tmove(s, w, tau);
        wsum1 = wsum1 + wtnow;
        esum1 = esum1 + wtnow * enew;
    }
    const double wgsum1 = wsum1 * d->fprod, egsum1 = esum1 * d->fprod;
    if (dmc_splitj(s, d) < 0) return -1;
    d->wgsum += wgsum1; d->egsum += egsum1;
    d->eest = (d->egcum + d->egsum) / (d->wgcum + d->wgsum);
    d->eigv = exp((d->etrial - d->eest) * (d->taucum + d->tausum) / (d->wgcum + d->wgsum));
    d->wdsumo = wsum1;
    d->wtgen[ipmod] = wsum1;
    out[0] = wsum1; out[1] = esum1; out[2] = wgsum1; out[3] = egsum1; out[4] = d->eest; out[5] = d->eigv;
    out[6] = d->fprod; out[7] = nod;
    return d->nwalk;
}

/* end of block (acuest): block energy egsum/wgsum, cumulative sums */
double orc_dmc_block(orc_dmc_t *d)
{
    const double en = d->egsum / d->wgsum;
    d->wgcum += d->wgsum; d->egcum += d->egsum; d->taucum += d->tausum;
    d->wgcm2 += d->wgsum * d->wgsum; d->egcm2 += d->egsum * en;
    d->wgsum = d->egsum = d->tausum = 0;
    return en;
}

int orc_dmc_get(const orc_dmc_t *d, double *wt, double *x, double *eold)
{
    for (int iw = 0; iw < d->nwalk; iw++) {
        if (wt) wt[iw] = d->wt[iw];
        if (x) orc_walker_get_x(d->w[iw], x + (size_t)iw * 3 * d->w[iw]->nelec);
        if (eold) eold[iw] = d->w[iw]->eold;
    }
    return d->nwalk;
}

void orc_walker_scalars(const oracle_walker_t *w, double out[4])
{
    out[0] = w->psido; out[1] = w->psijo; out[2] = w->eold; out[3] = w->psi2o;
}
void orc_get_vj(const oracle_walker_t *w, double *vj) { memcpy(vj, w->vj, sizeof(double) * 3 * w->nelec); }
void orc_get_vjn(const oracle_walker_t *w, double *vjn) { memcpy(vjn, w->vjn, sizeof(double) * 3 * w->nelec); }
void orc_get_ddx(const oracle_walker_t *w, double *ddx, double *d2dx2)
{
    memcpy(ddx, w->ddx, sizeof(double) * 3 * w->nelec);
    memcpy(d2dx2, w->d2dx2, sizeof(double) * w->nelec);
}
void orc_get_slmi(const oracle_walker_t *w, double *slmi) { memcpy(slmi, w->slmi, sizeof(double) * 2 * w->nmax * w->nmax); }
void orc_get_detiab(const oracle_walker_t *w, double det[2]) { det[0] = w->detiab[0]; det[1] = w->detiab[1]; }
void orc_get_orbn(const oracle_walker_t *w, double *orbn)
{
    memcpy(orbn, w->orbn, sizeof(double) * w->norb);
    memcpy(orbn + w->norb, w->dorbn, sizeof(double) * 3 * w->norb);
}
void orc_get_parts(const oracle_walker_t *w, double out[6])
{
    out[0] = w->pe_local; out[1] = w->d2j; out[2] = w->vpsp_det[0]; out[3] = w->vpsp_det[1];
    out[4] = (double)w->nxquad; out[5] = w->d2n;
}

/* CPU baseline: nwalk independent chains, each thread advancing whole walkers;
 * the initial hpsi of every walker is done before the timed region.        */
double orc_vmc_run(const champ_system_t *s, int nwalk, const double *x0, const uint64_t *rng0, int nsweep,
                   double tau, int nthreads, double *esum, double *accepted)
{
    double es = 0, acc = 0;
    struct timespec t0, t1;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
    oracle_walker_t **ws = (oracle_walker_t **)calloc(nwalk, sizeof(*ws));
#pragma omp parallel for schedule(dynamic, 1)
    for (int iw = 0; iw < nwalk; iw++) {
        oracle_walker_t *w = orc_walker_new(s);
        orc_walker_set_x(w, x0 + (size_t)iw * 3 * s->nelec);
        orc_walker_set_rng(w, rng0 + 4 * (size_t)iw);
        double psid, psij, ek, en;
        orc_hpsi(s, w, &psid, &psij, &ek, &en);
        ws[iw] = w;
    }
    clock_gettime(CLOCK_MONOTONIC, &t0);
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : es, acc)
    for (int iw = 0; iw < nwalk; iw++) {
        oracle_walker_t *w = ws[iw];
        for (int isw = 0; isw < nsweep; isw++) {
            acc += orc_metrop1_sweep(s, w, tau);
            es += w->eold;
        }
    }
    clock_gettime(CLOCK_MONOTONIC, &t1);
    for (int iw = 0; iw < nwalk; iw++) orc_walker_free(ws[iw]);
    free(ws);
    if (esum) *esum = es;
    if (accepted) *accepted = acc;
    return (t1.tv_sec - t0.tv_sec) + 1e-9 * (t1.tv_nsec - t0.tv_nsec);
}
