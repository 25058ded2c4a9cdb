/*
 * champ_gpu.h -- C-ABI of the MI355X engine for CHAMP's per-move wavefunction
 * evaluation (walker-population batched).
 *
 * Every entry point acts on a whole walker population held on one GPU by an
 * engine context.  The routine names, argument order and meaning follow the
 * CHAMP Fortran subroutines they replace (cited per declaration, paths relative
 * to the CHAMP source tree); the one-walker module-global state of the
 * reference becomes per-walker device state inside the context, and every
 * per-walker output gains a leading walker index.
 *
 * Host arrays are C row-major; sizes are given per declaration.  Indices of
 * electrons, centres and orbitals are 0-based here (the Fortran side adds 1).
 * All calls are ordered on the context's single HIP stream and return when
 * their host outputs are written.  Errors are reported by the return code;
 * a walker whose state went bad (singular Slater matrix, zero ratio,
 * non-finite value) is flagged in its status word, never aborts the batch.
 */
#ifndef CHAMP_GPU_H
#define CHAMP_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- limits (match CHAMP's compile-time parameters) --------------------- */
#define CHAMP_NCOEF     5   /* small-r polynomial terms, m_vmc.f90:96          */
#define CHAMP_MPS_L     5   /* max pseudopotential components, m_pseudo.f90:3 */
#define CHAMP_MPS_QUAD  86  /* max quadrature points, m_pseudo.f90:4          */
#define CHAMP_MGAUSS    8   /* max Gaussian terms per ECP component (CHAMP: 100) */
#define CHAMP_MAX_SLM   20  /* Cartesian monomials up to f, slm.f90:22-28      */
#define CHAMP_MAX_NORD  8   /* max Jastrow polynomial order                    */

/* ---- return codes -------------------------------------------------------- */
#define CHAMP_OK               0
#define CHAMP_ERR_ARG         -1
#define CHAMP_ERR_HIP         -2
#define CHAMP_ERR_UNSUPPORTED -3
#define CHAMP_ERR_NOMEM       -4
#define CHAMP_ERR_STATE       -5
#define CHAMP_ERR_NUMERIC     -6

/* ---- per-walker status bits ---------------------------------------------- */
#define CHAMP_WALKER_SINGULAR   1  /* LU pivot was exactly zero (matinv.f90:59)   */
#define CHAMP_WALKER_ZERO_RATIO 2  /* determinant ratio exactly 0 (determinante.f90:46) */
#define CHAMP_WALKER_NONFINITE  4  /* NaN/Inf in psi, drift or local energy       */

/* ---- precision modes ----------------------------------------------------- */
#define CHAMP_PREC_FP64  0     /* every stage in IEEE double                      */
/* FP32 arithmetic for AO values, the Jastrow (moves and hpsi) and the ECP
 * quadrature terms; FP64 for the AO->MO contraction (matrix cores), the
 * Sherman-Morrison updates, a full recompute of the Slater inverse every 80
 * moves, every sum across lanes and the local energy. */
#define CHAMP_PREC_MIXED 1

/*
 * System description (read-only tables shared by all walkers).  Mirrors the
 * module globals CHAMP fills at input time: `system` (cent, znuc, iwctype),
 * `numbas`/`numexp`/`numbas1` (grid-3 splines, read_data_trexio.f90:991-1179),
 * `coefs`/`slater` (MO coefficients, determinants), `jastrow` (ijas=4, isc=2;
 * m_jastrow.f90:145-147), `gauss_ecp`/`pseudo` (nloc=4) and `qua`.
 */
typedef struct champ_system {
    int32_t nelec, nup;             /* ndn = nelec - nup                          */
    int32_t ncent, nctype;
    const double  *cent;            /* [ncent][3]                                 */
    const int32_t *iwctype;         /* [ncent] centre -> type (0-based)           */
    const double  *znuc;            /* [nctype] (effective) nuclear charge        */

    /* numerical radial basis on grid type 3: r_i = r0*arg^(i-1) - r0            */
    int32_t nr;                     /* grid points (2000)                         */
    double  grid_r0, grid_arg;      /* r0 already divided by (arg^(nr-1)-1)       */
    int32_t nrbas_max;              /* leading dimension for radial shells        */
    const int32_t *nrbas;           /* [nctype] radial shells per type            */
    const double  *rwf;             /* [nctype][nrbas_max][nr]  R(r_i)            */
    const double  *d2rwf;           /* [nctype][nrbas_max][nr]  spline R''        */
    const double  *ce;              /* [nctype][nrbas_max][NCOEF] small-r poly    */
    const double  *ae;              /* [nctype][nrbas_max][2]  a*exp(-b r) tail   */
    const double  *rmaxwf;          /* [nctype][nrbas_max] cut radius             */
    int32_t nbastyp_max;            /* leading dimension for AOs per type         */
    const int32_t *nbastyp;         /* [nctype] AOs on a centre of this type      */
    const int32_t *iwrwf;           /* [nctype][nbastyp_max] radial shell (0-based) */
    const int32_t *iwlbas;          /* [nctype][nbastyp_max] slm index (1-based, 1..20) */
    int32_t nbasis;                 /* sum over centres of nbastyp(iwctype)       */

    /* orbitals and determinants                                                 */
    int32_t norb;                   /* MOs referenced by any determinant          */
    const double  *coef;            /* [norb][nbasis]  coef(m,iorb) of CHAMP      */
    int32_t ndet;                   /* determinants (1 = single Slater det)       */
    const int32_t *iworbd;          /* [ndet][nelec] orbital of electron slot (0-based);
                                       determinant 0 is the reference (kref = 1); every
                                       other one may differ from it by at most 4
                                       orbitals per spin, in any slot order          */
    const double  *cdet;            /* [ndet] CI coefficients                     */

    /* Jastrow ijas=4, isc=2                                                     */
    int32_t norda, nordb, nordc;
    int32_t nspin2b, nocuspb;       /* CHAMP defaults 1 and 0                     */
    double  scalek;                 /* kappa                                      */
    const double *a4;               /* [nctype][norda+1]                          */
    const double *b;                /* [nspin2b][nordb+1]                         */
    const double *c;                /* [nctype][nterms4(nordc)]                   */

    /* Gaussian ECP (nloc = 4) or all-electron (nloc = 0)                        */
    int32_t nloc;
    const int32_t *lpot;            /* [nctype] local component index (1-based)   */
    const int32_t *necp_term;       /* [nctype][MPS_L] terms of component l (l=1..) */
    const int32_t *necp_power;      /* [nctype][MPS_L][MGAUSS]                    */
    const double  *ecp_coef;        /* [nctype][MPS_L][MGAUSS]                    */
    const double  *ecp_exponent;    /* [nctype][MPS_L][MGAUSS]                    */
    int32_t nquad;                  /* quadrature points                          */
    const double  *xq0, *yq0, *zq0, *wq;  /* [nquad] unrotated rule (gesqua)     */
} champ_system_t;

/* Opaque engine context: tables + walker population resident in HBM.        */
typedef struct champ_ctx champ_ctx;

/* Sweep statistics returned by the population driver.                       */
typedef struct champ_sweep_stats {
    double accepted;        /* accepted single-electron moves (all walkers)      */
    double proposed;        /* proposed moves                                    */
    double esum;            /* sum over walkers and sweeps of E_L (FP64)         */
    double esum2;           /* sum of E_L^2                                      */
    int64_t nflagged;       /* walkers with a non-zero status word               */
} champ_sweep_stats;

/* --------------------------------------------------------------------------
 * Context life cycle.  Replaces the input-time uploads of CHAMP's module
 * globals (read_data_trexio.f90:991-1179 tables, readps_gauss, set_scale_dist
 * scale_dist.f90:3-60) by one upload of `sys` for a population of `nwalk`.
 * ------------------------------------------------------------------------ */
int champ_gpu_init(champ_ctx **ctx, const champ_system_t *sys, int32_t nwalk,
                   int32_t precision);
int champ_gpu_free(champ_ctx *ctx);
const char *champ_gpu_last_error(void);

/* Walker positions xold, [nwalk][nelec][3].  mc_configs.f90:33-150 reads the
 * same quantity from `mc_configs_start`.                                      */
int champ_gpu_set_walkers(champ_ctx *ctx, const double *x);
int champ_gpu_get_walkers(champ_ctx *ctx, double *x);

/* Per-walker xoshiro256** state, [nwalk][4] uint64.  seed_rng reproduces
 * parser.f90:651-661: the 32-character seed read as 8 x i4 (setrn), then
 * walker w takes the stream jump()ed w times (one walker per MPI rank).
 * init seeds the streams with CHAMP's default seed string
 * "18123437465549275475255231234865" (parser.f90:315).                      */
int champ_gpu_set_rng(champ_ctx *ctx, const uint64_t *state);
int champ_gpu_get_rng(champ_ctx *ctx, uint64_t *state);
int champ_gpu_seed_rng(champ_ctx *ctx, const int32_t seed[8]);

/* Rotated quadrature per walker, [nwalk][nquad] each: the state rotqua
 * (rotqua.f90:357-430) leaves in xq,yq,zq.  rotqua draws the rotation from
 * each walker's own stream exactly as the reference does once per step.      */
int champ_gpu_set_quadrature(champ_ctx *ctx, const double *xq, const double *yq,
                             const double *zq);
int champ_gpu_rotqua(champ_ctx *ctx);

/* hpsi (hpsi.f90:3-273): full recompute of distances, pot_local, jastrow_factor4,
 * determinant (orbitals + LU inverse), nonloc_pot and E_L at the current xold,
 * refreshing every cached per-walker quantity.  Outputs per walker [nwalk]
 * (any may be NULL): psid (Slater part), psij (log Jastrow), ekin, energy.     */
int champ_gpu_hpsi(champ_ctx *ctx, double *psid, double *psij, double *ekin,
                   double *energy);

/* determinant (determinant.f90:9-140) on its own: orbitals of all electrons
 * at the current positions, both Slater matrices, their LU inverse (matinv,
 * matinv.f90:22-92), the spin determinants, ddx/d2dx2 and the orbital
 * gradients of every electron (plus the multideterminant state when ndet > 1),
 * all in FP64; then psid and psi2o refreshed with the unchanged Jastrow.
 * The mixed path runs this same recompute every lu_every moves inside
 * metrop1.  psid (may be NULL) [nwalk].                                      */
int champ_gpu_determinant(champ_ctx *ctx, double *psid);

/* Moves between the FP64 recomputes of the Slater inverse inside a metrop1
 * sweep (0: only the one in hpsi).  Defaults: 80 on the mixed path with one
 * determinant, 0 on the FP64 path and for multideterminant trial functions
 * (whose recompute rebuilds the whole excitation state).  No counterpart in
 * the reference, which rebuilds the inverse once per sweep in hpsi.          */
int champ_gpu_set_lu_every(champ_ctx *ctx, int32_t nmoves);

/* With ndet > 1 the multideterminant parts run inside these same calls:
 * init builds the excitation tables (multideterminants_define,
 * set_input_data.f90:46-414), hpsi adds multideterminant_hpsi
 * (multideterminant.f90:3-255) and psie / compute_determinante_grad / accept
 * add multideterminante(_grad) and its saves (multideterminante.f90:3-222,
 * detsav.f90).  psid and psidn are then sum_k cdet_k D_up,k D_dn,k.        */

/* compute_determinante_grad (determinante.f90:69-283):
 * iflag_move=1 -> drift of electron iel at its current position,
 * iflag_move=0 -> drift at the position last proposed by psie.
 * vd: [nwalk][3].                                                             */
int champ_gpu_compute_determinante_grad(champ_ctx *ctx, int32_t iel,
                                        int32_t iflag_move, double *vd);

/* psie (hpsie.f90:3-79): distances(iel), jastrowe/jastrow4e, determinante
 * (orbitalse + Sherman-Morrison ratio) for electron iel moved to xnew
 * ([nwalk][3]) in every walker.  psidn: new Slater value, psijn: new log
 * Jastrow ([nwalk] each).  iflag=0 as in metrop1 (no Laplacians).            */
int champ_gpu_psie(champ_ctx *ctx, int32_t iel, const double *xnew,
                   double *psidn, double *psijn, int32_t iflag);

/* jassav + detsav (jassav.f90:3-78, detsav.f90:3-94) for walkers with
 * accept[w] != 0, distancese_restore (distances.f90:169-211) for the others.  */
int champ_gpu_accept(champ_ctx *ctx, int32_t iel, const uint8_t *accept);

/* metrop1 (metrop_mov1_driftdif.f90:4-363) over the whole population for
 * nsweep sweeps: rotqua (if nloc>0), Ne drift-diffusion single-electron moves
 * with accept/reject drawn from each walker's stream, then hpsi.  All state
 * stays on the GPU; stats accumulate the local energies of every sweep.      */
int champ_gpu_metrop1(champ_ctx *ctx, int32_t nsweep, double tau,
                      champ_sweep_stats *stats);

/* metrop6, CHAMP's default mover (imetro=6, metrop_mov1_slat.f90:3-738; the
 * accelerated Metropolis method of Umrigar, PRL 71, 408 (1993)) over the
 * whole population for nsweep sweeps: rotqua (if nloc>0), Ne single-electron
 * moves sampled in local spherical coordinates about each electron's nearest
 * nucleus with the drift-biased radial and azimuthal densities, geometric
 * rejections and the reverse-move acceptance, then hpsi.  iguiding = 0,
 * node_cutoff = 0.  Parameters from champ_gpu_set_metrop6 (CHAMP's defaults
 * deltar 5, deltat 1, fbias 1 until then).                                   */
int champ_gpu_metrop6(champ_ctx *ctx, int32_t nsweep, champ_sweep_stats *stats);

/* deltar, deltat, fbias of metrop6 (the metropolis module of parser.f90:
 * 402-404), with the reference's input checks (:774-787): deltar < 1 and
 * deltat outside [0, 2] are reset to 2, fbias is truncated to [0, 2].       */
int champ_gpu_set_metrop6(champ_ctx *ctx, double deltar, double deltat, double fbias);

/* Result of a blocked VMC run (champ_gpu_vmc).                            */
typedef struct champ_vmc_result {
    double efin;            /* total E: sum of E_L over the accumulated blocks / samples   */
    double eerr;            /* its error from the block averages (finwrt.f90:182-187):
                               sqrt(|sum_b w_b e_b^2 / W - efin^2| / nblk)                  */
    double sigma;           /* sqrt(<E_L^2> - <E_L>^2) over every sample                   */
    double tcorr;           /* (eerr / (sigma / sqrt(samples)))^2: the correlation time in
                               sweeps, 1 for uncorrelated samples (CHAMP's Tcor column)      */
    double accept;          /* accepted / proposed moves of the accumulated blocks          */
    int64_t nsamples;       /* walker-sweeps accumulated: nwalk * nstep * nblk             */
    int64_t nflagged;       /* walkers with a non-zero status word at the end              */
} champ_vmc_result;

/* vmc (vmc.f90:3-159) over the population: nblkeq equilibration blocks, then
 * nblk blocks of nstep sweeps whose energies are accumulated per block as
 * acuest does (acuest.f90; the block weight is nwalk * nstep, as the sum over
 * CHAMP's one-walker ranks in acuest_reduce).  imetro = 1: metrop1 sweeps
 * (drift-diffusion, time step tau); imetro = 6: metrop6 sweeps (tau unused,
 * as in the reference).  enow (may be NULL) [nblk]: the block averages.
 * res (may be NULL).                                                        */
int champ_gpu_vmc(champ_ctx *ctx, int32_t nblk, int32_t nstep, int32_t nblkeq, double tau, int32_t imetro,
                  double *enow, champ_vmc_result *res);

/* --------------------------------------------------------------------------
 * Diffusion Monte Carlo (dmc.f90 with the dmc_ps mover, dmc_ps_mov1.f90:4-838,
 * or for all-electron systems (nloc = 0) the dmc_ae mover, dmc_ae_mov1.f90,
 * as dmc.f90:98-102 chooses) at CHAMP's default controls: idmc = 2,
 * icasula = 0 (the non-local
 * pseudopotential in the locality approximation; -1, T-moves, through
 * champ_gpu_dmc_set_icasula), ipq = 1, itau_eff = 1,
 * iacc_rej = 1, icross = 1, icut_br = 0, icut_e = 0, nforce = 1,
 * node_cutoff = 0.  The context's walker
 * count at the first champ_gpu_dmc_init is the population's capacity (MWALK);
 * from then on the context's walkers are the live population, slots
 * 0 .. nwalk-1.  Each walker slot draws its moves from its own stream; the
 * branching (splitj) uniforms come from the population stream given here
 * (CHAMP draws both from one stream per MPI rank).
 * ------------------------------------------------------------------------ */
typedef struct champ_dmc_step_stats {
    int32_t nwalk;          /* live walkers after the step's branching           */
    int32_t pad;
    double wsum, esum;      /* sum of weights and of weight * E_L (wsum1, esum1)  */
    double wgsum, egsum;    /* the same times fprod (growth estimator)           */
    double eest, eigv, fprod;  /* energy estimate, population factor, fprod     */
    double accept;          /* mean acceptance probability of the step's moves   */
    int64_t nodecr;         /* node crossings of the step                        */
    int64_t nbranch;        /* branching events since the last zerest            */
} champ_dmc_step_stats;

typedef struct champ_dmc_result {
    double egave, egerr;    /* growth estimator of the energy and its error
                               (finwrt.f90:228-229: egcum/wgcum, errg)            */
    double eest;            /* energy estimate at the end                        */
    double accept;          /* acceptance of the accumulated moves               */
    double tau_eff;         /* taucum / wgcum                                    */
    int32_t nwalk, nwalk_max;
    int64_t nbranch, nodecr;
} champ_dmc_result;

/* init (dmc/init.f90:3-191): the first nconf walkers (set by set_walkers)
 * start with weight 1; hpsi of each; eest = etrial, eigv = fprod = 1,
 * nfprod population-control steps (parser.f90:440), time step tau.
 * pop_state: the xoshiro256** state of the branching stream.                 */
int champ_gpu_dmc_init(champ_ctx *ctx, int32_t nconf, double tau, double etrial, int32_t nfprod,
                       const uint64_t pop_state[4]);

/* One DMC step of the population: rotqua, dmc_ps (every electron of every
 * walker), the reweighting, splitj (splitj.f90:3-233) and acues1
 * (acues1.f90:3-95).  stats may be NULL.                                      */
int champ_gpu_dmc_step(champ_ctx *ctx, champ_dmc_step_stats *stats);

/* dmc (dmc.f90:3-125): 2*nblkeq equilibration blocks (zerest after each
 * half), then nblk blocks of nstep steps; egnow (may be NULL) [nblk]: the
 * block energies (acuest.f90); res (may be NULL).                           */
int champ_gpu_dmc(champ_ctx *ctx, int32_t nblk, int32_t nstep, int32_t nblkeq, double *egnow,
                  champ_dmc_result *res);

/* icasula of the dmc module (parser.f90): 0, the locality approximation (the
 * default), or -1, one Casula T-move per walker after each step's moves and
 * hpsi (dmc_ps_mov1.f90:701-730, nonloc_grid.f90:3-137 over all electrons:
 * heat-bath choice among the quadrature points with negative t_vpsp,
 * nonloc.f90:297, in the order electron, point, centre; the walker's stream
 * draws the choice).  The local energy and the weights are those of the
 * configuration before the T-move, as in the reference.                    */
int champ_gpu_dmc_set_icasula(champ_ctx *ctx, int32_t icasula);

/* The live walker count and their weights [nwalk] (wt may be NULL).          */
int champ_gpu_dmc_weights(champ_ctx *ctx, double *wt, int32_t *nwalk);

/* --------------------------------------------------------------------------
 * Parameter derivatives d ln psi / d p at the walkers' current configurations
 * (deriv_jastrow4.f90:3-667 gvalue for the Jastrow coefficients; optci.f90:
 * 55-57 for the determinant coefficients) and the population sums of
 * stochastic reconfiguration (optwf_sr.f90:504-620).  They use the state the
 * last hpsi left (metrop1/metrop6 sweeps end with one).  Raw parameters, in
 * this order and in the layouts of champ_system_t's arrays, selected by the
 * bits of `kinds`:
 *   CHAMP_DERIV_JAS_EN   a4[it][p], p = 0..norda, per centre type (p = 0, 1:
 *                        the Pade numerator and denominator, CHAMP's a(1),
 *                        a(2); p >= 2 the polynomial, CHAMP's a(p+1))
 *   CHAMP_DERIV_JAS_EE   b[isb][p], p = 0..nordb, per spin class (nspin2b)
 *   CHAMP_DERIV_JAS_EEN  c[it][t] per centre type, every term of order nordc
 *                        in jastrow4.f90's order (no cusp elimination)
 *   CHAMP_DERIV_CI       cdet[k] of every determinant (ndet > 1 only)
 *   CHAMP_DERIV_ORB      the orbital terms of champ_gpu_set_orb_terms
 *   CHAMP_DERIV_MAPPED   the parameters of the map set by
 *                        champ_gpu_set_deriv_map (other bits ignored)
 * ------------------------------------------------------------------------ */
#define CHAMP_DERIV_JAS_EN  1
#define CHAMP_DERIV_JAS_EE  2
#define CHAMP_DERIV_JAS_EEN 4
#define CHAMP_DERIV_CI      8
#define CHAMP_DERIV_MAPPED  16
#define CHAMP_DERIV_ORB     32
int champ_gpu_deriv_count(champ_ctx *ctx, int32_t kinds, int32_t *nparm);

/* A linear map from the raw parameters of `kinds` (nraw of them) to a
 * caller's nout parameters: O' = O map, map [nraw][nout] row-major.  This is
 * how CHAMP's own parameter list is obtained: the iwjasa / iwjasb / iwjasc
 * selections of the jastrow_der file (read_data_parser.f90:1521-1700), the
 * e-e-n coefficients fixed by the cusp conditions folded into the free ones
 * (deriv_jastrow4.f90:394-460 with cdep of checkdepend4), and the CSF
 * combinations of optci.f90:70-80.  map NULL or nout 0 removes the map.      */
int champ_gpu_set_deriv_map(champ_ctx *ctx, int32_t kinds, int32_t nout, const double *map);

/* The orbital terms of CHAMP_DERIV_ORB: ideriv [nterm][2] = (io, jo),
 * 0-based orbital indices, each the excitation phi_io += kappa phi_jo that
 * compute_lcao applies (optwf_handle_wf.f90:1128-1134, orbitals_ortho off;
 * CHAMP's ideriv list of optorb_define, optorb.f90:838-1060).  The derivative
 * is sum_e V[e][io] phi_jo(r_e), V = d ln psi_D / d phi(r_e) (optorb_deriv,
 * optorb.f90:5-97, orbprim).  nterm 0 removes the terms.                      */
int champ_gpu_set_orb_terms(champ_ctx *ctx, int32_t nterm, const int32_t *ideriv);

/* O [nwalk][nparm]: d ln psi / d p of every walker.                         */
int champ_gpu_deriv_psi(champ_ctx *ctx, int32_t kinds, double *o);

/* Sums over the walkers of O_k O_l (s [nparm][nparm], on the matrix cores),
 * O_k (osum [nparm]), E_L O_k (eosum [nparm]) and E_L (esum); any may be NULL. */
int champ_gpu_sr_sums(champ_ctx *ctx, int32_t kinds, double *s, double *osum, double *eosum, double *esum);

/* New wave-function parameters (compute_jastrow / compute_ci,
 * optwf_handle_wf.f90:1050-1104, 1214-1284, after the caller has applied the
 * step and, for the e-e-n coefficients, cuspexact4): arrays in the layouts of
 * champ_system_t; a NULL array keeps its values.  The Pade denominators
 * a4[it][1] and b[isb][1] must exceed -scalek (check_parms_jas,
 * optwf_handle_wf.f90:1286-1336): CHAMP_ERR_ARG otherwise.  Every walker's
 * state is recomputed (hpsi) before the call returns.                       */
int champ_gpu_set_params(champ_ctx *ctx, const double *a4, const double *b, const double *c,
                         const double *cdet);

/* New orbital coefficients coef [norb][nbasis] (the orbital part of
 * compute_lcao, optwf_handle_wf.f90:1106-1150); every walker's state is
 * recomputed (hpsi) before the call returns.                                */
int champ_gpu_set_orbitals(champ_ctx *ctx, const double *coef);

typedef struct champ_sr_info {
    double adiag;           /* the diagonal shift finally used                   */
    double dparm_norm;      /* rms of the step                                   */
    int32_t cg_iters;       /* conjugate-gradient iterations of the last solve   */
    int32_t retries;        /* times the shift was multiplied by ten             */
} champ_sr_info;

/* The SR step from the sums (host only, no GPU work): sr_hs
 * (optwf_sr.f90:504-620) with S_kl = <O_k O_l> - <O_k><O_l> plus sr_adiag
 * times its diagonal, h = -2 (<E O_k> - <E><O_k>), the Jacobi-preconditioned
 * conjugate gradients of pcg (optwf_sr_more.f90:31-97) to relative residual
 * sr_eps, dparm = -sr_tau x (optwf_sr.f90:213-214), and the retry of
 * optwf_sr.f90:215-231 with ten times the shift while the rms of dparm
 * exceeds dparm_norm_min (CHAMP's default 1).  The sums are those of
 * champ_gpu_sr_sums over walkers of total weight wsum.  New parameters are
 * p - dparm (compute_jastrow subtracts).  info may be NULL.                  */
int champ_sr_solve(int32_t nparm, double wsum, const double *s, const double *osum, const double *eosum,
                   double esum, double sr_tau, double sr_eps, double sr_adiag, double dparm_norm_min,
                   double *dparm, champ_sr_info *info);

/* SR sums kept on the GPU, for parameter sets whose S (nparm^2 doubles) is
 * too large to bring to the host at every sample: adds the sums of
 * champ_gpu_sr_sums over the current walkers (weight 1 each) to accumulators
 * on the device; reset = 1 clears them first (needed when kinds or the
 * parameter count change).  wsum and esum (may be NULL) receive the running
 * total weight and sum of E_L.  The reference accumulates the same sums over
 * its samples (sr_hs's inputs, optwf_sr.f90:504-620).                        */
int champ_gpu_sr_accumulate(champ_ctx *ctx, int32_t kinds, int32_t reset, double *wsum, double *esum);

/* The SR step of champ_sr_solve from the sums champ_gpu_sr_accumulate holds
 * (same equations, shift retries and info), with the products S z of the
 * conjugate gradients (atimes_n, optwf_sr_more.f90:31-160) computed on the
 * GPU over the resident S.  dparm [nparm] on the host.                       */
int champ_gpu_sr_solve(champ_ctx *ctx, double sr_tau, double sr_eps, double sr_adiag, double dparm_norm_min,
                       double *dparm, champ_sr_info *info);

/* metrop1 over n populations at once, typically different molecules (each
 * context its own system, walker count and precision): every context's
 * kernels run on its own stream in the order of its own champ_gpu_metrop1,
 * so each population evolves exactly as under champ_gpu_metrop1 with the
 * same tau, while the GPU overlaps the populations.  stats (may be NULL)
 * [n].  The contexts must be distinct; sweeps are launched, not replayed
 * from a graph.  No counterpart in the reference, which runs one molecule
 * per process.                                                                */
int champ_gpu_metrop1_batch(champ_ctx *const *ctxs, int32_t n, int32_t nsweep,
                            double tau, champ_sweep_stats *stats);

/* Autotuner: times each candidate kernel form of the tunable launches on the
 * context's current walkers and keeps the fastest for the context's later
 * calls.  choices (may be NULL) [4]: the kernel kind of the AO->MO GEMM of the
 * moves, of the hpsi orbitals, of the ECP coefficient rows (0: 16 x 16 tiles
 * per wavefront, 1: 16 x 80 strips with LDS-shared B chunks, 2: VALU, 3: FP32
 * operands and accumulation, a candidate for the moves of the mixed path,
 * 4: the tile form split over two halves of K with their partials added in a
 * fixed order, a candidate for the moves),
 * and the ECP quadrature form (0: matrix cores, 1: VALU); -1 where nothing was
 * tuned.  Leaves the walker state as
 * it was (hpsi is recomputed).  No counterpart in the reference.              */
int champ_gpu_autotune(champ_ctx *ctx, int32_t *choices);

/* Applies kernel choices without timing them, in the layout autotune returns
 * (choices [4]; -1 restores the shape rule of that site).  GEMM kind 3 (FP32)
 * is accepted for the moves of the mixed path only, kind 4 (split K) for the
 * moves only.  No counterpart in the reference.                              */
int champ_gpu_set_tuning(champ_ctx *ctx, const int32_t *choices);

/* on = 1: metrop1 replays each sweep as a HIP graph captured on first use
 * (per tau); on = 0 (default): every kernel is launched from the host.  No
 * counterpart in the reference.  Kernel-class timing (champ_gpu_kernel_times)
 * always runs without the graph.                                              */
int champ_gpu_set_graphs(champ_ctx *ctx, int32_t on);

/* Per-walker state readback for checking: psido, psijo, eold, psi2o [nwalk]
 * and the status words.                                                       */
int champ_gpu_get_walker_scalars(champ_ctx *ctx, double *psido, double *psijo,
                                 double *eold, double *psi2o);
int champ_gpu_get_status(champ_ctx *ctx, int32_t *status);

/* Array readback for checking (host buffer sized by champ_gpu_array_size).   */
#define CHAMP_ARR_SLMI   1   /* [nwalk][2][nmax][nmax] transpose inverse rows  */
#define CHAMP_ARR_VJ     2   /* [nwalk][nelec][3]  Jastrow gradient (vj)       */
#define CHAMP_ARR_DDX    3   /* [nwalk][nelec][3]  grad_i psi_D / psi_D (all determinants) */
#define CHAMP_ARR_D2DX2  4   /* [nwalk][nelec]     lap_i psi_D / psi_D (all determinants)  */
#define CHAMP_ARR_DETIAB 5   /* [nwalk][2]         reference spin determinants */
#define CHAMP_ARR_VPSP   6   /* [nwalk][2]         nonlocal energy per spin (psi_D with all determinants) */
#define CHAMP_ARR_PELOC  7   /* [nwalk]            local potential             */
#define CHAMP_ARR_D2J    8   /* [nwalk]            Jastrow Laplacian sum       */
#define CHAMP_ARR_ORBN   9   /* [nwalk][4][norb]   orbn,dorbn of last psie     */
#define CHAMP_ARR_VJN   10   /* [nwalk][nelec][3]  vjn of last psie            */
#define CHAMP_ARR_NXQUAD 11  /* [nwalk]            quadrature points used      */
#define CHAMP_ARR_FSO    12  /* [nwalk][nelec][nelec] Jastrow pair values fso (diagonal: e-n) */
int64_t champ_gpu_array_size(champ_ctx *ctx, int32_t which);
int champ_gpu_get_array(champ_ctx *ctx, int32_t which, double *out);

/* basis_fns (basis_fns.f90:8-214) at arbitrary points: npt positions
 * [npt][3] -> phin [npt][5][nbasis] (value, d/dx, d/dy, d/dz, Laplacian),
 * evaluated with ider=2.                                                      */
int champ_gpu_basis_fns(champ_ctx *ctx, int32_t npt, const double *pos,
                        double *phin);

/* champ_gpu_basis_fns without any device allocation: no counterpart in the
 * reference; for callers that cannot allocate device memory (tight loops,
 * code that forbids hipMalloc).  The points and values are staged in a static
 * 16 MiB device array per GPU (held until the process ends), so npt is
 * limited to npt * (3 + 5 nbasis) doubles within it (about 690 points at 600
 * AOs); larger requests return CHAMP_ERR_ARG.  Every call on one GPU shares
 * that array: calls that overlap on one GPU (from several host threads or
 * contexts) race and must be serialised by the caller.  Same outputs as
 * champ_gpu_basis_fns.                                                       */
int champ_gpu_basis_fns_noalloc(champ_ctx *ctx, int32_t npt, const double *pos,
                                double *phin);

/* Per-kernel-class timing with HIP events on the context stream.  With names
 * and ms both NULL the call switches timing on (*n > 0) or off (*n = 0) and
 * clears the sums, returning the number of classes in *n.  Otherwise it
 * returns the milliseconds and launch counts summed since then: names
 * (n x 32 chars), ms, launches (any may be NULL) sized by *n on input.  No
 * counterpart in the reference (m_timer times on the host).                   */
int champ_gpu_kernel_times(champ_ctx *ctx, int32_t *n, char *names /* n*32 */,
                           double *ms, int64_t *launches);

#ifdef __cplusplus
}
#endif
#endif /* CHAMP_GPU_H */
