/*
 * champ_oracle.h -- CPU restatement (FP64, one walker at a time) of CHAMP's
 * per-move wavefunction path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the parity checker for the GPU engine and
 * the "reference CPU path" timed beside it.  Only tests/, the smoke check and
 * the CPU-baseline leg of bench.py may load it; the engine never does.
 *
 * Parity pinning (which check fixes each function) is listed in the header
 * comment of champ_oracle.c.
 */
#ifndef CHAMP_ORACLE_H
#define CHAMP_ORACLE_H

#include <stdint.h>
#include "../include/champ_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oracle_walker oracle_walker_t;

/* --- primitives --------------------------------------------------------- */
double orc_powi(double x, int n);
void   orc_splfit(const champ_system_t *s, int it, int irb, double r, double wfv[4], int ider);
void   orc_slm(int l, const double rvec[3], double r2, double *y, double dy[3], double *ddy_lap, int ider);
void   orc_basis_fns_point(const champ_system_t *s, const double *rvec /*[ncent][3]*/,
                           const double *r /*[ncent]*/, int ider, double *phi /*[nbasis]*/,
                           double *dphi /*[3][nbasis]*/, double *d2phi /*[nbasis]*/);
int    orc_matinv(double *a, int n, double *det);   /* column-major, in place */
int    orc_nterms4(int nord);
void   orc_asymp(const champ_system_t *s, double *asymp_jasa /*[nctype]*/, double asymp_jasb[2]);
double orc_pot_nn(const champ_system_t *s);
void   orc_gauss_pot(const champ_system_t *s, double r, int l, int ict, double *vpot, double *dvpot);
double orc_yl0(int l, double costh);
void   orc_gesqua(int nq, double *xq, double *yq, double *zq, double *wq);

/* xoshiro256** (Blackman & Vigna 2018) */
uint64_t orc_xoshiro_next(uint64_t s[4]);
void     orc_xoshiro_jump(uint64_t s[4]);
double   orc_random_dp(uint64_t s[4]);
double   orc_gauss(uint64_t s[4]);
void     orc_setrn(uint64_t s[4], const int32_t seed[8]);
void     orc_rotqua(const champ_system_t *s, uint64_t rng[4], double *xq, double *yq, double *zq);

/* --- walker state ------------------------------------------------------- */
oracle_walker_t *orc_walker_new(const champ_system_t *s);
void   orc_walker_free(oracle_walker_t *w);
void   orc_walker_set_x(oracle_walker_t *w, const double *x);
void   orc_walker_get_x(const oracle_walker_t *w, double *x);
void   orc_walker_set_quad(oracle_walker_t *w, const double *xq, const double *yq, const double *zq);
void   orc_walker_set_rng(oracle_walker_t *w, const uint64_t st[4]);
void   orc_walker_get_rng(const oracle_walker_t *w, uint64_t st[4]);

/* hpsi at the walker's xold.  Fills psid, psij, ekin, energy. */
void   orc_hpsi(const champ_system_t *s, oracle_walker_t *w, double *psid, double *psij,
                double *ekin, double *energy);
/* psie for electron iel (0-based) moved to xnew[3]. */
void   orc_psie(const champ_system_t *s, oracle_walker_t *w, int iel, const double xnew[3],
                double *psidn, double *psijn, int iflag);
void   orc_determinante_grad(const champ_system_t *s, oracle_walker_t *w, int iel,
                             int iflag_move, double vd[3]);
void   orc_accept(const champ_system_t *s, oracle_walker_t *w, int iel);
void   orc_reject(const champ_system_t *s, oracle_walker_t *w, int iel);
/* one metrop1 sweep (rotqua first when nloc>0); returns accepted moves */
int    orc_metrop1_sweep(const champ_system_t *s, oracle_walker_t *w, double tau);
/* one metrop6 sweep (accelerated Metropolis, metrop_mov1_slat.f90), then hpsi */
int    orc_metrop6_sweep(const champ_system_t *s, oracle_walker_t *w, double deltar, double deltat, double fbias);
double orc_gammai(double a, double x, double xae, int *iflag);
double orc_m6_area(double zeta, double co, double rbot, double rtop, double bot);
/* DMC population (dmc.f90 / dmc_ps / splitj / acues1 at default controls) */
typedef struct orc_dmc orc_dmc_t;
orc_dmc_t *orc_dmc_new(const champ_system_t *s, int cap, const double *x0, const uint64_t *rng0, int nconf, double tau,
                       double etrial, int nfprod, const uint64_t pop[4]);
void   orc_dmc_free(orc_dmc_t *d);
void   orc_dmc_set_icasula(orc_dmc_t *d, int icasula);
int    orc_dmc_step(const champ_system_t *s, orc_dmc_t *d, double out[8]);
double orc_dmc_block(orc_dmc_t *d);
int    orc_dmc_get(const orc_dmc_t *d, double *wt, double *x, double *eold);
void   orc_walker_scalars(const oracle_walker_t *w, double out[4]); /* psido psijo eold psi2o */

/* state getters for checking */
void   orc_get_vj(const oracle_walker_t *w, double *vj);          /* [nelec][3] */
void   orc_get_vjn(const oracle_walker_t *w, double *vjn);        /* [nelec][3] */
void   orc_get_ddx(const oracle_walker_t *w, double *ddx, double *d2dx2);
void   orc_get_slmi(const oracle_walker_t *w, double *slmi);      /* [2][nmax][nmax] */
void   orc_get_detiab(const oracle_walker_t *w, double det[2]);
void   orc_get_orbn(const oracle_walker_t *w, double *orbn);      /* [4][norb] */
void   orc_get_parts(const oracle_walker_t *w, double out[6]);    /* pe_local d2j vpsp(2) nxquad d2n */
void   orc_jastrow_full(const champ_system_t *s, oracle_walker_t *w, double *fsum,
                        double *d2, double *fj /*[nelec][3]*/);

/* cpu baseline: nwalk independent chains, nsweep sweeps, OpenMP threads */
double orc_vmc_run(const champ_system_t *s, int nwalk, const double *x0, const uint64_t *rng0,
                   int nsweep, double tau, int nthreads, double *esum, double *accepted);

#ifdef __cplusplus
}
#endif
#endif
