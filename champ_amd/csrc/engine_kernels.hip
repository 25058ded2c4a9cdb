// engine_kernels.hip -- FP64 walker-population kernels of the per-move path.
//
// Layout (walker-major, lanes run over the innermost index):
//   xT[w][3][NE], r_en/rr_en[w][NC][NE], fso[w][NE][NE] (symmetric),
//   g1/g2[w][NE][3][NE] (pair gradients on the first/second electron),
//   vjT[w][3][NE], slmi[w][2][NM][NM] (row i = electron, CHAMP slmi(j+(i-1)n)),
//   dorb[w][NE][3][NO].
// One 64-lane wavefront serves one walker in the per-move kernels; the
// AO->MO contractions are population-wide GEMMs.
#include "een_poly.hpp"
#include "engine_internal.hpp"
#include "dmc_ae.hpp"
#include "metrop6.hpp"

#include <algorithm>
#include <cmath>

namespace champ {

struct MoveArgs {
    int W, NE, NUP, NM, NC, NB, NO, iel, mode, pad;
    double tau, cdet0;
    const double *cent;
    const int *iwctype;
    SplineTab tab;
    const double *rmax_ao;
    const int *ao_cent, *ao_type, *ao_rad, *ao_slm, *iworbd;
    int NRI;
    const int *ri_cent, *ri_type, *ri_rad, *ao_ri;
    const double *ri_rmax;
    const JasParams *jas;
    double *xT, *r_en, *rr_en, *fso, *g1, *g2, *vjT, *fsumo, *dorb, *slmi, *detiab, *psido, *psijo, *psi2o;
    uint64_t *rng;
    double *xnew, *dfus2o, *rnew, *rrnew, *fsn_row, *g1n, *g2n, *vjnT, *fsumn, *aomove, *orbn;
    double *ratio, *psidn, *psijn, *vnew, *vold, *esum_dev;
    unsigned char *acc;
    int *status;
    int md;                            // multideterminant trial function
    const double *md_vcol, *md_vcoln;  // [W][NO] d ln psi_D / d phi_o(r_iel), current / proposed
    double *md_S, *md_Sn;              // [W] psi_D / (D_up,ref D_dn,ref), current / proposed
    int *md_cur;                       // [W][2] current buffer of the multideterminant state
    int ao32;                          // AO block stored as FP32 for the FP32 contraction
    int imetro;                        // mode-0 mover: 1 metrop1, 6 metrop6, 2 DMC (dmc_ps)
    M6Params m6p;
    double *dmc_acc;                   // [W][DMC_ACC] DMC move accumulators of the current step
    int *dmc_age;                      // [W] steps since the walker last moved
    double *m6rec;                     // [W][M6_REC]
    const EcpParams *ecp;              // znuc per type and nloc (metrop6's cusp and theta range)
    float2 *enp;                       // mixed path: [W][NC][NE] partner pair table (enp_pack), else null
};

static MoveArgs move_args(champ_ctx *c, int iel, int mode, double tau)
{
    MoveArgs a;
    a.W = c->W; a.NE = c->NE; a.NUP = c->NUP; a.NM = c->NM; a.NC = c->NC; a.NB = c->NB; a.NO = c->NO;
    a.iel = iel; a.mode = mode; a.pad = 0; a.tau = tau; a.cdet0 = c->cdet0;
    a.cent = c->cent; a.iwctype = c->iwctype; a.tab = c->tab; a.rmax_ao = c->rmax_ao;
    a.ao_cent = c->ao_cent; a.ao_type = c->ao_type; a.ao_rad = c->ao_rad; a.ao_slm = c->ao_slm;
    a.iworbd = c->iworbd; a.jas = c->jas;
    a.NRI = c->NRI; a.ri_cent = c->ri_cent; a.ri_type = c->ri_type; a.ri_rad = c->ri_rad; a.ao_ri = c->ao_ri;
    a.ri_rmax = c->ri_rmax;
    a.enp = c->enp;
    a.xT = c->xT; a.r_en = c->r_en; a.rr_en = c->rr_en; a.fso = c->fso; a.g1 = c->g1; a.g2 = c->g2;
    a.vjT = c->vjT; a.fsumo = c->fsumo; a.dorb = c->dorb; a.slmi = c->slmi; a.detiab = c->detiab;
    a.psido = c->psido; a.psijo = c->psijo; a.psi2o = c->psi2o; a.rng = c->rng;
    a.xnew = c->xnew; a.dfus2o = c->dfus2o; a.rnew = c->rnew; a.rrnew = c->rrnew; a.fsn_row = c->fsn_row;
    a.g1n = c->g1n; a.g2n = c->g2n; a.vjnT = c->vjnT; a.fsumn = c->fsumn; a.aomove = c->aomove; a.orbn = c->orbn;
    a.ratio = c->ratio; a.psidn = c->psidn; a.psijn = c->psijn; a.vnew = c->vnew; a.vold = c->vold;
    a.esum_dev = c->esum_dev; a.acc = c->acc; a.status = c->status;
    a.md = c->md; a.md_vcol = c->md_vcol; a.md_vcoln = c->md_vcoln; a.md_S = c->md_S; a.md_Sn = c->md_Sn;
    a.md_cur = c->md_cur;
    a.ao32 = c->prec == CHAMP_PREC_MIXED && c->tune_gemm[0] == GEMM_F32;
    a.imetro = c->imetro;
    a.m6p = M6Params{c->m6_deltar, c->m6_deltat, c->m6_fbias};
    a.m6rec = c->m6rec;
    a.ecp = c->ecp;
    a.dmc_acc = c->dmc_acc;
    a.dmc_age = c->dmc_age;
    return a;
}

// ------------------------------------------------------------------------
// Drift at the current position of iel: compute_determinante_grad with
// iflag_move=1 (determinante.f90:127-157, determinante_ref_grad :285-328).
__device__ __forceinline__ void drift_old(const MoveArgs &a, int w, int lane, double v[3])
{
    const int iel = a.iel, iab = iel < a.NUP ? 0 : 1, nel = iab ? a.NE - a.NUP : a.NUP, ish = iab ? a.NUP : 0;
    const double *srow = a.slmi + ((size_t)w * 2 + iab) * a.NM * a.NM + (size_t)(iel - ish) * a.NM;
    const double *dor = a.dorb + ((size_t)w * a.NE + iel) * 3 * a.NO;
    double v0 = 0, v1 = 0, v2 = 0;
    if (a.md) {  // sum over every orbital of d ln psi_D / d phi_o grad phi_o
        const double *vc = a.md_vcol + (size_t)w * a.NO;
        for (int o = lane; o < a.NO; o += WAVE) {
            const double s = vc[o];
            v0 += s * dor[o]; v1 += s * dor[a.NO + o]; v2 += s * dor[2 * a.NO + o];
        }
    } else {
        for (int j = lane; j < nel; j += WAVE) {
            const double s = srow[j];
            const int jo = a.iworbd[ish + j];
            v0 += s * dor[jo];
            v1 += s * dor[a.NO + jo];
            v2 += s * dor[2 * a.NO + jo];
        }
    }
    v0 = wave_sum(v0); v1 = wave_sum(v1); v2 = wave_sum(v2);
    const double *vj = a.vjT + (size_t)w * 3 * a.NE;
    v[0] = 0.0 + v0 + vj[iel];
    v[1] = 0.0 + v1 + vj[a.NE + iel];
    v[2] = 0.0 + v2 + vj[2 * a.NE + iel];
}

__global__ void __launch_bounds__(256) k_drift_old(MoveArgs a)
{
    const int lane = threadIdx.x % WAVE;
    const int w = blockIdx.x * (blockDim.x / WAVE) + threadIdx.x / WAVE;
    if (w >= a.W) return;
    double v[3];
    drift_old(a, w, lane, v);
    if (lane == 0) for (int k = 0; k < 3; k++) a.vold[3 * w + k] = v[k];
}

// ------------------------------------------------------------------------
// Move kernel A, one wavefront per walker:
//   mode 0: proposal from the drift at the old position: metrop1 (3 gauss
//           draws; metrop_mov1_driftdif.f90:105-117) or metrop6 (local
//           spherical move about the nearest nucleus; metrop6.hpp);
//   mode 1: xnew given.
//   then distances(iel) (distances.f90:3-167), jastrow4e with iflag=0
//   (jastrow4e.f90:3-358) and basis_fns(iel, ider=1) (basis_fns.f90:8-214)
//   writing the AO block that the MO GEMM contracts (orbitalse).
// Tuning of k_move_a: partner-electron slots per lane in the pair loop and
// the minimum waves per SIMD asked of the register allocator.
#ifndef CHAMP_MJB
#define CHAMP_MJB 1
#endif
#ifndef CHAMP_MOVE_WPE
#define CHAMP_MOVE_WPE 1
#endif
constexpr int MJB = CHAMP_MJB;
// centres whose partner distances are loaded one group ahead: 4 on the FP64
// path (0.30 against 0.335 ms per move at 160 e x 4096 walkers), 1 on the FP32
// path (its register budget; 4 ran at 0.32 against 0.24 ms)
#ifndef CHAMP_MOVE_PF
#define CHAMP_MOVE_PF 4
#endif
template <typename T> constexpr int move_pf() { return sizeof(T) == sizeof(double) ? CHAMP_MOVE_PF : 1; }

// LDS of one wave of k_move_a, in doubles: r, rr, -kap (1 - kap rr) / r,
// rvec (3) per centre (double), then in the arithmetic type (tsize bytes) the
// centre's e-e-n tables (value and d/dx, na entries each) and R, R' of each
// radial instance at the moved electron
__host__ __device__ __forceinline__ int move_a_lds_doubles(int NC, int na, int NRI, int tsize)
{
    return 6 * NC + ((2 * na * NC + 2 * NRI) * tsize + 7) / 8;
}

// T: arithmetic type of the pair loop, the e-e-n tables and the AO values
// (double on the FP64 path, float on the mixed path); distances of the moved
// electron, the state and every sum across lanes stay double.
template <int NORDC, typename T>
__global__ void __launch_bounds__(256) CHAMP_WAVES_PER_EU(CHAMP_MOVE_WPE) k_move_a(MoveArgs a)
{
    HIP_DYNAMIC_SHARED(double, lds)
    constexpr int NA = een_naz<NORDC>();
    const int wv = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
    const int w = blockIdx.x * (blockDim.x / WAVE) + wv;
    const bool active = w < a.W;
    const int NE = a.NE, NC = a.NC, iel = a.iel;
    const JasParams &jp = *a.jas;
    const double kap = jp.kap, asymp_r = jp.asymp_r;
    double *Lr = lds + (size_t)wv * move_a_lds_doubles(NC, NORDC > 1 ? NA : 0, a.NRI, (int)sizeof(T));
    double *Lrr = Lr + NC, *Lfm = Lr + 2 * NC, *Lrv = Lr + 3 * NC;
    T *Ltab = (T *)(Lr + 6 * NC);
    T *Lrad = Ltab + (NORDC > 1 ? 2 * NA * NC : 0);

    double xn0 = 0, xn1 = 0, xn2 = 0;
    if (active) {
        const double *xw = a.xT + (size_t)w * 3 * NE;
        if (a.mode == 0) {
            double vo[3];
            drift_old(a, w, lane, vo);
            if (lane == 0 && a.imetro == 6) {
                uint64_t s[4];
                for (int k = 0; k < 4; k++) s[k] = a.rng[4 * (size_t)w + k];
                const double xo[3] = {xw[iel], xw[NE + iel], xw[2 * NE + iel]};
                double xn[3];
                m6_propose(a.cent, a.iwctype, a.ecp->znuc, a.ecp->nloc, NC, a.m6p, xo, vo, s, xn,
                           a.m6rec + (size_t)w * M6_REC);
                for (int k = 0; k < 4; k++) a.rng[4 * (size_t)w + k] = s[k];
                for (int k = 0; k < 3; k++) { a.xnew[3 * w + k] = xn[k]; a.vold[3 * w + k] = vo[k]; }
                xn0 = xn[0]; xn1 = xn[1]; xn2 = xn[2];
            } else if (lane == 0 && a.imetro == 2 && a.ecp->nloc == 0) {
                // dmc_ae proposal (dmc_ae_mov1.f90:226-353)
                uint64_t s[4];
                for (int k = 0; k < 4; k++) s[k] = a.rng[4 * (size_t)w + k];
                const double xo[3] = {xw[iel], xw[NE + iel], xw[2 * NE + iel]};
                double xn[3], dfus2o, fnormo;
                ae_propose(a.cent, a.iwctype, a.ecp->znuc, NC, a.tau, xo, vo, s, xn, dfus2o, fnormo);
                for (int k = 0; k < 4; k++) a.rng[4 * (size_t)w + k] = s[k];
                for (int k = 0; k < 3; k++) { a.xnew[3 * w + k] = xn[k]; a.vold[3 * w + k] = vo[k]; }
                a.dfus2o[w] = dfus2o;
                a.dmc_acc[(size_t)w * DMC_ACC + DMC_FNORMO] = fnormo;
                xn0 = xn[0]; xn1 = xn[1]; xn2 = xn[2];
            } else if (lane == 0 && a.imetro == 2) {
                // dmc_ps drift-diffusion proposal (dmc_ps_mov1.f90:291-303): the
                // drift is vavvt v, without metrop1's extra factor tau
                const double tau = a.tau, rttau = sqrt(tau), adrift = 0.5;
                uint64_t s[4];
                for (int k = 0; k < 4; k++) s[k] = a.rng[4 * (size_t)w + k];
                const double v2old = vo[0] * vo[0] + vo[1] * vo[1] + vo[2] * vo[2];
                const double vavvt = (sqrt(1.0 + 2.0 * adrift * v2old * tau) - 1.0) / (adrift * v2old);
                double dfus2o = 0.0, xn[3];
                for (int k = 0; k < 3; k++) {
                    const double drift = vavvt * vo[k];
                    const double dfus = xo_gauss(s) * rttau;
                    const double dx = drift + dfus;
                    dfus2o = dfus2o + dfus * dfus;
                    xn[k] = xw[k * NE + iel] + dx;
                }
                for (int k = 0; k < 4; k++) a.rng[4 * (size_t)w + k] = s[k];
                for (int k = 0; k < 3; k++) { a.xnew[3 * w + k] = xn[k]; a.vold[3 * w + k] = vo[k]; }
                a.dfus2o[w] = dfus2o;
                xn0 = xn[0]; xn1 = xn[1]; xn2 = xn[2];
            } else if (lane == 0) {
                const double tau = a.tau, rttau = sqrt(tau), adrift = 0.5;
                uint64_t s[4];
                for (int k = 0; k < 4; k++) s[k] = a.rng[4 * (size_t)w + k];
                double v2old = vo[0] * vo[0] + vo[1] * vo[1] + vo[2] * vo[2];
                double vavvt = (sqrt(1.0 + 2.0 * adrift * v2old * tau) - 1.0) / (adrift * v2old);
                double dfus2o = 0.0, xn[3];
                for (int k = 0; k < 3; k++) {
                    double drift = vavvt * vo[k] * tau;
                    double dfus = xo_gauss(s) * rttau;
                    double dx = drift + dfus;
                    dfus2o = dfus2o + dfus * dfus;
                    xn[k] = xw[k * NE + iel] + dx;
                }
                for (int k = 0; k < 4; k++) a.rng[4 * (size_t)w + k] = s[k];
                for (int k = 0; k < 3; k++) { a.xnew[3 * w + k] = xn[k]; a.vold[3 * w + k] = vo[k]; }
                a.dfus2o[w] = dfus2o;
                xn0 = xn[0]; xn1 = xn[1]; xn2 = xn[2];
            }
            xn0 = __shfl(xn0, 0, WAVE); xn1 = __shfl(xn1, 0, WAVE); xn2 = __shfl(xn2, 0, WAVE);
        } else {
            xn0 = a.xnew[3 * w]; xn1 = a.xnew[3 * w + 1]; xn2 = a.xnew[3 * w + 2];
        }
    }

    // ---- e-n distances of the moved electron + its e-n Jastrow term; per
    // centre the e-e-n coefficient tables at the moved electron (een_poly.hpp)
    double fs_en = 0, ge0 = 0, ge1 = 0, ge2 = 0;
    if (active) {
        for (int c = lane; c < NC; c += WAVE) {
            const double rv0 = xn0 - a.cent[3 * c], rv1 = xn1 - a.cent[3 * c + 1], rv2 = xn2 - a.cent[3 * c + 2];
            double r2 = 0;
            r2 = r2 + rv0 * rv0; r2 = r2 + rv1 * rv1; r2 = r2 + rv2 * rv2;
            const double r = sqrt(r2);
            const double rr = scale_r(kap, r);
            const double dd7 = 1.0 - kap * rr;
            Lr[c] = r; Lrr[c] = rr; Lfm[c] = -kap * dd7 / r;
            Lrv[3 * c] = rv0; Lrv[3 * c + 1] = rv1; Lrv[3 * c + 2] = rv2;
            a.rnew[(size_t)w * NC + c] = r;
            a.rrnew[(size_t)w * NC + c] = rr;
            if constexpr (NORDC > 1) een_tables<NORDC, 1, T>(jp.c[a.iwctype[c]], (T)(1.0 - kap * rr), Ltab + (size_t)c * 2 * NA);
            if (rr != asymp_r) {
                const int it = a.iwctype[c];
                const double *a4 = jp.a4[it];
                double top = a4[0] * rr, topi = a4[0];
                double bot = 1 + a4[1] * rr, boti = a4[1], bot2 = bot * bot;
                double fen = top / bot - jp.ja[it];
                double feni = topi / bot - boti * top / bot2;
                double rprev = rr;  // rri(iord-1)
                for (int iord = 2; iord <= jp.norda; iord++) {
                    double rp = powi(rr, iord);
                    fen = fen + a4[iord] * rp;
                    feni = feni + a4[iord] * iord * rprev;
                    rprev = rp;
                }
                feni = feni * dd7 / r;
                fs_en += fen;
                ge0 += feni * rv0; ge1 += feni * rv1; ge2 += feni * rv2;
            }
        }
        fs_en = wave_sum(fs_en); ge0 = wave_sum(ge0); ge1 = wave_sum(ge1); ge2 = wave_sum(ge2);
    }
    __syncthreads();

    if (active) {
        const double *xw = a.xT + (size_t)w * 3 * NE;
        const double *fso_row = a.fso + ((size_t)w * NE + iel) * NE;
        const double *g1_row = a.g1 + ((size_t)w * NE + iel) * 3 * NE;
        const double *g2_row = a.g2 + ((size_t)w * NE + iel) * 3 * NE;
        const double *vj = a.vjT + (size_t)w * 3 * NE;
        double *vjn = a.vjnT + (size_t)w * 3 * NE;
        double *fsn = a.fsn_row + (size_t)w * NE;
        double *g1n = a.g1n + (size_t)w * 3 * NE, *g2n = a.g2n + (size_t)w * 3 * NE;
        const double *rw = a.r_en + (size_t)w * NC * NE, *rrw = a.rr_en + (size_t)w * NC * NE;
        double dfs = 0, dg0 = 0, dg1 = 0, dg2 = 0;
        if (lane == 0) {  // the moved electron's own entries: its e-n terms
            fsn[iel] = fs_en;
            g1n[iel] = ge0; g1n[NE + iel] = ge1; g1n[2 * NE + iel] = ge2;
            g2n[iel] = ge0; g2n[NE + iel] = ge1; g2n[2 * NE + iel] = ge2;
            dfs += fs_en - fso_row[iel];
            dg0 += ge0 - g1_row[iel]; dg1 += ge1 - g1_row[NE + iel]; dg2 += ge2 - g1_row[2 * NE + iel];
        }
        // Partners in slots of WAVE electrons, MJB slots per lane; centres in
        // the outer loop so each centre's table is read once per slot group.
        // Per pair: gm/go = gradient on the moved / the partner electron.
        const T tkap = (T)kap, tasymp = (T)asymp_r;
        for (int j0 = 0; j0 < NE; j0 += WAVE * MJB) {
            T fs[MJB], gm[MJB][3], go[MJB][3], xj[MJB][3], uu1[MJB], fuf[MJB], sfu[MJB];
            bool on[MJB];
#pragma unroll
            for (int s = 0; s < MJB; s++) {
                const int j = j0 + s * WAVE + lane;
                on[s] = j < NE && j != iel;
                const int jq = j < NE ? j : 0;
                fs[s] = 0; sfu[s] = 0;
                for (int k = 0; k < 3; k++) { gm[s][k] = 0; go[s][k] = 0; }
                const double xd0 = xw[jq], xd1 = xw[NE + jq], xd2 = xw[2 * NE + jq];
                xj[s][0] = (T)xd0; xj[s][1] = (T)xd1; xj[s][2] = (T)xd2;
                const T rmo[3] = {(T)(xn0 - xd0), (T)(xn1 - xd1), (T)(xn2 - xd2)};
                T r2 = 0;
                r2 = r2 + rmo[0] * rmo[0]; r2 = r2 + rmo[1] * rmo[1]; r2 = r2 + rmo[2] * rmo[2];
                const T rij = sqrt(r2);
                const T u = scale_r<T>(tkap, rij);
                const T dd1 = switched_from<T>(kap, rij, u);
                uu1[s] = dd1;
                fuf[s] = -tkap * dd1 / rij;
                // e-e term (jastrow4e.f90: psi_ee with its u-derivative)
                const int I = j > iel ? j : iel, J = j > iel ? iel : j;
                T sspinn = 1; int isb = 0, ipar = 0;
                if (I < a.NUP || J >= a.NUP) {
                    if (jp.nspin2b == 2) isb = 1;
                    else if (jp.nocuspb == 0) sspinn = T(0.5);
                    ipar = 1;
                }
                if (on[s] && (sizeof(T) == sizeof(float) || u != tasymp)) {
                    const double *b = jp.b[isb];
                    T top = sspinn * (T)b[0] * u, topu = sspinn * (T)b[0];
                    T bot = 1 + (T)b[1] * u, botu = (T)b[1], bot2 = bot * bot;
                    T fee = top / bot - (T)jp.jb[ipar];
                    T feeu = topu / bot - botu * top / bot2;
                    T up = u;
                    for (int iord = 2; iord <= jp.nordb; iord++) {
                        T un = u * up;
                        fee = fee + (T)b[iord] * un;
                        feeu = feeu + (T)b[iord] * (T)iord * up;
                        up = un;
                    }
                    feeu = feeu * dd1 / rij;
                    if constexpr (sizeof(T) == sizeof(float))  // value minus asymptote without cancellation
                        fee = pade_minus_asym<T>(b, jp.nordb, sspinn, u, asym_offset<T>(asymp_r, dd1), tasymp);
                    fs[s] += fee;
                    for (int k = 0; k < 3; k++) { gm[s][k] += feeu * rmo[k]; go[s][k] -= feeu * rmo[k]; }
                }
            }
            if constexpr (NORDC > 1) {
                // distances of the partners to the centres, MOVE_PF centres at a
                // time with the next group's loads in flight (slots past NE read
                // electron 0 and are not used)
                int jcl[MJB];
#pragma unroll
                for (int s = 0; s < MJB; s++) {
                    const int j = j0 + s * WAVE + lane;
                    jcl[s] = j < NE ? j : 0;
                }
                constexpr int MOVE_PF = move_pf<T>();
                // FP32 path: the partner pair table (x, gradient factor), 8
                // bytes per pair instead of the two FP64 distances
                constexpr bool F32 = sizeof(T) == sizeof(float);
                const float2 *ew = F32 ? a.enp + (size_t)w * NC * NE : nullptr;
                double rjc[MOVE_PF][MJB], rrjc[MOVE_PF][MJB], rjn[MOVE_PF][MJB], rrjn[MOVE_PF][MJB];
                float2 ejc[MOVE_PF][MJB], ejn[MOVE_PF][MJB];
#pragma unroll
                for (int p = 0; p < MOVE_PF; p++)
#pragma unroll
                    for (int s = 0; s < MJB; s++) {
                        const int c = p < NC ? p : 0;
                        if constexpr (F32) {
                            ejn[p][s] = ew[(size_t)c * NE + jcl[s]];
                        } else {
                            rjn[p][s] = rw[(size_t)c * NE + jcl[s]];
                            rrjn[p][s] = rrw[(size_t)c * NE + jcl[s]];
                        }
                    }
                for (int c0 = 0; c0 < NC; c0 += MOVE_PF) {
#pragma unroll
                    for (int p = 0; p < MOVE_PF; p++)
#pragma unroll
                        for (int s = 0; s < MJB; s++) {
                            const int cn = c0 + MOVE_PF + p;
                            if constexpr (F32) {
                                ejc[p][s] = ejn[p][s];
                                if (cn < NC) ejn[p][s] = ew[(size_t)cn * NE + jcl[s]];
                            } else {
                                rjc[p][s] = rjn[p][s]; rrjc[p][s] = rrjn[p][s];
                                if (cn < NC) {
                                    rjn[p][s] = rw[(size_t)cn * NE + jcl[s]];
                                    rrjn[p][s] = rrw[(size_t)cn * NE + jcl[s]];
                                }
                            }
                        }
#pragma unroll
                    for (int p = 0; p < MOVE_PF; p++) {
                        const int c = c0 + p;
                        if (c >= NC) break;
                        const T *tab = Ltab + (size_t)c * 2 * NA;
                        const double rrm = Lrr[c];
                        const T fmf = (T)Lfm[c];
                        const T rvm0 = (T)Lrv[3 * c], rvm1 = (T)Lrv[3 * c + 1], rvm2 = (T)Lrv[3 * c + 2];
                        const T cx = (T)a.cent[3 * c], cy = (T)a.cent[3 * c + 1], cz = (T)a.cent[3 * c + 2];
#pragma unroll
                        for (int s = 0; s < MJB; s++) {
                            if (j0 + s * WAVE >= NE) break;
                            if (!on[s]) continue;
                            T swj;
                            if constexpr (F32) {
                                swj = ejc[p][s].x;
                                if (rrm == asymp_r && swj == 0.0f) continue;
                            } else {
                                if (rrm == asymp_r && rrjc[p][s] == asymp_r) continue;
                                swj = (T)(1.0 - kap * rrjc[p][s]);
                            }
                            const EenOut<T> e = een_eval_tab<NORDC, 1, T>(tab, uu1[s], swj);
                            fs[s] += e.f;
                            sfu[s] += e.fu;
                            const T fmv = e.fx * fmf;
                            T fot;
                            if constexpr (F32) fot = e.fy * ejc[p][s].y;
                            else fot = e.fy * (-tkap * swj) / (T)rjc[p][s];
                            gm[s][0] += fmv * rvm0; gm[s][1] += fmv * rvm1; gm[s][2] += fmv * rvm2;
                            go[s][0] += fot * (xj[s][0] - cx); go[s][1] += fot * (xj[s][1] - cy);
                            go[s][2] += fot * (xj[s][2] - cz);
                        }
                    }
                }
            }
#pragma unroll
            for (int s = 0; s < MJB; s++) {
                const int j = j0 + s * WAVE + lane;
                if (!on[s]) continue;
                const T fu = sfu[s] * fuf[s];
                const T rmo[3] = {(T)(xn0 - (double)xj[s][0]), (T)(xn1 - (double)xj[s][1]), (T)(xn2 - (double)xj[s][2])};
                double gi[3], gj[3];
                for (int k = 0; k < 3; k++) { gi[k] = (double)(gm[s][k] + fu * rmo[k]); gj[k] = (double)(go[s][k] - fu * rmo[k]); }
                fsn[j] = (double)fs[s];
                dfs += (double)fs[s] - fso_row[j];
                for (int k = 0; k < 3; k++) {
                    g1n[k * NE + j] = gi[k];
                    g2n[k * NE + j] = gj[k];
                    vjn[k * NE + j] = vj[k * NE + j] + gj[k] - g2_row[k * NE + j];
                }
                dg0 += gi[0] - g1_row[j]; dg1 += gi[1] - g1_row[NE + j]; dg2 += gi[2] - g1_row[2 * NE + j];
            }
        }
        dfs = wave_sum(dfs); dg0 = wave_sum(dg0); dg1 = wave_sum(dg1); dg2 = wave_sum(dg2);
        if (lane == 0) {
            const double fsn_tot = a.fsumo[w] + dfs;
            a.fsumn[w] = fsn_tot;
            a.psijn[w] = fsn_tot;
            vjn[iel] = vj[iel] + dg0; vjn[NE + iel] = vj[NE + iel] + dg1; vjn[2 * NE + iel] = vj[2 * NE + iel] + dg2;
        }

        // ---- AOs of the moved electron (ider = 1) into the GEMM operand
        // radial functions (R, R') once per (centre, shell), then each AO
        for (int ri = lane; ri < a.NRI; ri += WAVE) {
            const double r = Lr[a.ri_cent[ri]];
            T wfv[3] = {T(0), T(0), T(0)};
            if (r <= a.ri_rmax[ri]) splfit_jx<T>(a.tab, a.ri_type[ri], a.ri_rad[ri], (T)r, grid_jx(a.tab, r), wfv, 1);
            Lrad[2 * ri] = wfv[0]; Lrad[2 * ri + 1] = wfv[1];
        }
        __syncwarp();
        // the AO block as FP64, or as FP32 for the FP32 contraction (a.ao32)
        double *ao = a.aomove + (size_t)w * 4 * a.NB;
        float *ao32 = (float *)a.aomove + (size_t)w * 4 * a.NB;
        for (int m = lane; m < a.NB; m += WAVE) {
            const int c = a.ao_cent[m], ri = a.ao_ri[m];
            const T wfv[3] = {Lrad[2 * ri], Lrad[2 * ri + 1], T(0)};
            T out[5] = {T(0), T(0), T(0), T(0), T(0)};
            if (wfv[0] != T(0) || wfv[1] != T(0))  // else beyond the cut radius: exactly zero
                ao_combine<T>(a.ao_slm[m], (T)Lrv[3 * c], (T)Lrv[3 * c + 1], (T)Lrv[3 * c + 2], (T)Lr[c], wfv, 1, out);
            if (a.ao32) {
                for (int k = 0; k < 4; k++) ao32[(size_t)k * a.NB + m] = (float)out[k];
            } else {
                for (int k = 0; k < 4; k++) ao[(size_t)k * a.NB + m] = (double)out[k];
            }
        }
    }
}

// Sherman-Morrison rows of one wave of a 4-wave block: rows wv + 4 r (r < MR,
// i != row), columns lane and lane + WAVE (nel <= 2 WAVE).  Same arithmetic
// as the row-by-row loop; all loads are issued before the reductions.
template <int MR>
__device__ __forceinline__ void sm_rows(double *S, int NM, int nel, int row, double rt, const double *s_row,
                                        const double *s_orb, int wv, int lane)
{
    const bool j0 = lane < nel, j1 = lane + WAVE < nel;
    double v0[MR], v1[MR];
#pragma unroll
    for (int r = 0; r < MR; r++) {
        const int i = wv + 4 * r;
        const bool ok = i < nel && i != row;
        v0[r] = (ok && j0) ? S[(size_t)i * NM + lane] : 0.0;
        v1[r] = (ok && j1) ? S[(size_t)i * NM + lane + WAVE] : 0.0;
    }
    const double o0 = j0 ? s_orb[lane] : 0.0, o1 = j1 ? s_orb[lane + WAVE] : 0.0;
    const double w0 = j0 ? s_row[lane] : 0.0, w1 = j1 ? s_row[lane + WAVE] : 0.0;
#pragma unroll
    for (int r = 0; r < MR; r++) {
        const int i = wv + 4 * r;
        if (i >= nel) break;  // uniform in the wave
        double sum = 0;
        sum += v0[r] * o0;
        sum += v1[r] * o1;
        sum = wave_sum(sum);
        sum = sum / rt;
        if (i == row) continue;
        if (j0) S[(size_t)i * NM + lane] = v0[r] - w0 * sum;
        if (j1) S[(size_t)i * NM + lane + WAVE] = v1[r] - w1 * sum;
    }
}

// ------------------------------------------------------------------------
// Move kernel B, one 256-thread block per walker: determinante ratio
// (determinante.f90:38-46), drift at the new position (iflag_move=0,
// :158-199), metropolis decision (metrop_mov1_driftdif.f90:166-205) and on
// acceptance jassav + detsav with the Sherman-Morrison update applied in
// place (determinante.f90:47-62 + detsav.f90:42-93).
//   mode 0: decide with the walker's stream; 1: compute only; 2: commit acc[w].
//   MR: rows of the inverse per wave held in registers during the update
//   (0: row-by-row loop for spin blocks above 2 WAVE electrons).
template <int MR>
__global__ void __launch_bounds__(256) k_move_b(MoveArgs a)
{
    HIP_DYNAMIC_SHARED(double, lds)
    const int w = blockIdx.x, tid = threadIdx.x, lane = tid % WAVE, wv = tid / WAVE, nwv = blockDim.x / WAVE;
    const int NE = a.NE, NO = a.NO, iel = a.iel;
    const int iab = iel < a.NUP ? 0 : 1, nel = iab ? NE - a.NUP : a.NUP, ish = iab ? a.NUP : 0, row = iel - ish;
    double *S = a.slmi + ((size_t)w * 2 + iab) * a.NM * a.NM;
    const double *ob = a.orbn + (size_t)w * 4 * NO;
    double *s_row = lds, *s_orb = lds + a.NM;
    __shared__ double s_ratio, s_detn;
    __shared__ int s_acc;

    if (wv == 0) {
        const double *srow = S + (size_t)row * a.NM;
        double rt = 0;
        for (int j = lane; j < nel; j += WAVE) rt += srow[j] * ob[a.iworbd[ish + j]];
        rt = wave_sum(rt);
        double v0 = 0, v1 = 0, v2 = 0;
        if (a.md) {
            const double *vc = a.md_vcoln + (size_t)w * NO;
            for (int o = lane; o < NO; o += WAVE) {
                const double sn = vc[o];
                v0 += sn * ob[NO + o]; v1 += sn * ob[2 * NO + o]; v2 += sn * ob[3 * NO + o];
            }
        } else {
            for (int j = lane; j < nel; j += WAVE) {
                const double sn = srow[j] / rt;
                const int jo = a.iworbd[ish + j];
                v0 += sn * ob[NO + jo]; v1 += sn * ob[2 * NO + jo]; v2 += sn * ob[3 * NO + jo];
            }
        }
        v0 = wave_sum(v0); v1 = wave_sum(v1); v2 = wave_sum(v2);
        if (lane == 0) {
            const double *vjn = a.vjnT + (size_t)w * 3 * NE;
            double vn[3] = {0.0 + v0 + vjn[iel], 0.0 + v1 + vjn[NE + iel], 0.0 + v2 + vjn[2 * NE + iel]};
            const double detn = a.detiab[2 * w + iab] * rt;
            const double psidn = (rt == 0.0) ? 0.0
                                             : 0.0 + detn * a.detiab[2 * w + 1 - iab] * (a.md ? a.md_Sn[w] : a.cdet0);
            const double psijn = a.psijn[w];
            int accept = 0;
            if (a.mode == 0 && a.imetro == 6) {
                const double psi2n = 2 * (log(fabs(psidn)) + psijn);
                const double *xw = a.xT + (size_t)w * 3 * NE;
                const double xo[3] = {xw[iel], xw[NE + iel], xw[2 * NE + iel]};
                const double xn[3] = {a.xnew[3 * w], a.xnew[3 * w + 1], a.xnew[3 * w + 2]};
                const double *rec = a.m6rec + (size_t)w * M6_REC;
                const double p = m6_accept_prob(a.cent, a.iwctype, a.ecp->znuc, a.ecp->nloc, a.m6p, xo, xn, vn, psi2n,
                                                a.psi2o[w], psidn == 0.0, rec);
                uint64_t s[4];
                for (int k = 0; k < 4; k++) s[k] = a.rng[4 * (size_t)w + k];
                const double uacc = xo_uniform(s);
                for (int k = 0; k < 4; k++) a.rng[4 * (size_t)w + k] = s[k];
                accept = uacc < p;
                if (accept) {
                    a.psi2o[w] = psi2n;
                    atomicAdd(&a.esum_dev[0], 1.0);
                }
                // psi is evaluated at proposals that pass the geometric test only
                if (rt == 0.0 && rec[M6_LIVE] != 0.0) atomicOr(&a.status[w], CHAMP_WALKER_ZERO_RATIO);
            } else if (a.mode == 0 && a.imetro == 2) {
                // dmc_ps acceptance (dmc_ps_mov1.f90:325-399), ipq = 1, icross = 1,
                // iacc_rej = 1, node_cutoff = 0: node crossings are counted, not rejected
                const double tau = a.tau, adrift = 0.5;
                const double *xw = a.xT + (size_t)w * 3 * NE;
                double *ac = a.dmc_acc + (size_t)w * DMC_ACC;
                const double psido = a.psido[w];
                if (psidn * psido <= 0.0) ac[DMC_NODECR] += 1.0;
                double dfus2n = 0.0, fnorm_ratio = 1.0;
                if (a.ecp->nloc == 0) {
                    // dmc_ae's reverse Green function (dmc_ae_mov1.f90:388-490)
                    const double xo[3] = {xw[iel], xw[NE + iel], xw[2 * NE + iel]};
                    const double xn[3] = {a.xnew[3 * w], a.xnew[3 * w + 1], a.xnew[3 * w + 2]};
                    double fnormn;
                    ae_reverse(a.cent, a.iwctype, a.ecp->znuc, a.NC, tau, xo, xn, vn, dfus2n, fnormn);
                    fnorm_ratio = fnormn / ac[DMC_FNORMO];
                } else {
                    const double v2new = vn[0] * vn[0] + vn[1] * vn[1] + vn[2] * vn[2];
                    const double vavvt = (sqrt(1.0 + 2.0 * adrift * v2new * tau) - 1.0) / (adrift * v2new);
                    for (int k = 0; k < 3; k++) {
                        const double drift = vavvt * vn[k];
                        const double xbac = a.xnew[3 * w + k] + drift;
                        const double dfus = xbac - xw[k * NE + iel];
                        dfus2n = dfus2n + dfus * dfus;
                    }
                }
                const double dfus2o = a.dfus2o[w];
                const double rat = psidn / psido;
                double p = rat * rat * exp(2 * (psijn - a.psijo[w])) * exp((dfus2o - dfus2n) / (2.0 * tau)) * fnorm_ratio;
                const int age = a.dmc_age[w];
                if (age > 50) p = p * pow(1.1, (double)(age - 50));
                ac[DMC_PP] = ac[DMC_PP] * p;
                p = p < 1.0 ? p : (p >= 1.0 ? 1.0 : 0.0);  // dmin1(one, p); a NaN p never passes
                ac[DMC_PSUM] += p;
                ac[DMC_DFUS2AC] += p * dfus2o;
                ac[DMC_DFUS2UN] += dfus2o;
                uint64_t s[4];
                for (int k = 0; k < 4; k++) s[k] = a.rng[4 * (size_t)w + k];
                const double uacc = xo_uniform(s);
                for (int k = 0; k < 4; k++) a.rng[4 * (size_t)w + k] = s[k];
                accept = uacc < p;
                if (accept) {
                    a.psi2o[w] = 2 * (log(fabs(psidn)) + psijn);
                    atomicAdd(&a.esum_dev[0], 1.0);
                    a.dmc_age[w] = 0;
                    ac[DMC_IACCEPT] = 1.0;
                }
                if (rt == 0.0) atomicOr(&a.status[w], CHAMP_WALKER_ZERO_RATIO);
            } else if (a.mode == 0) {
                const double tau = a.tau, adrift = 0.5;
                const double psi2n = 2 * (log(fabs(psidn)) + psijn);
                const double v2new = vn[0] * vn[0] + vn[1] * vn[1] + vn[2] * vn[2];
                const double vavvt = (sqrt(1.0 + 2.0 * adrift * v2new * tau) - 1.0) / (adrift * v2new);
                const double *xw = a.xT + (size_t)w * 3 * NE;
                double dfus2n = 0.0;
                for (int k = 0; k < 3; k++) {
                    double drift = vavvt * vn[k] * tau;
                    double xbac = a.xnew[3 * w + k] + drift;
                    double dfus = xbac - xw[k * NE + iel];
                    dfus2n = dfus2n + dfus * dfus;
                }
                double p = exp(psi2n - a.psi2o[w]) * exp(-0.5 * (dfus2n - a.dfus2o[w]) / tau);
                if (psidn == 0.0) p = 0.0;
                p = p < 1.0 ? p : 1.0;
                uint64_t s[4];
                for (int k = 0; k < 4; k++) s[k] = a.rng[4 * (size_t)w + k];
                const double uacc = xo_uniform(s);
                for (int k = 0; k < 4; k++) a.rng[4 * (size_t)w + k] = s[k];
                accept = uacc < p;
                if (accept) {
                    a.psi2o[w] = psi2n;
                    atomicAdd(&a.esum_dev[0], 1.0);
                }
                if (rt == 0.0) atomicOr(&a.status[w], CHAMP_WALKER_ZERO_RATIO);
            } else if (a.mode == 2) {
                accept = a.acc[w] != 0;
                if (accept) a.psi2o[w] = 2 * (log(fabs(psidn)) + psijn);
            }
            a.ratio[w] = rt; a.psidn[w] = psidn;
            for (int k = 0; k < 3; k++) a.vnew[3 * w + k] = vn[k];
            if (accept) {
                a.psido[w] = psidn; a.psijo[w] = psijn;
                if (a.md) { a.md_S[w] = a.md_Sn[w]; a.md_cur[2 * w + iab] ^= 1; }
            }
            s_ratio = rt; s_detn = detn; s_acc = accept;
        }
    }
    __syncthreads();
    if (!s_acc) return;
    const double rt = s_ratio;
    for (int j = tid; j < nel; j += blockDim.x) {
        s_row[j] = S[(size_t)row * a.NM + j];
        s_orb[j] = ob[a.iworbd[ish + j]];
    }
    __syncthreads();
    // Sherman-Morrison: rows i != iel (all rows of a wave loaded before any
    // is reduced, so their loads are in flight together)
    if constexpr (MR > 0)
        sm_rows<MR>(S, a.NM, nel, row, rt, s_row, s_orb, wv, lane);
    else
        for (int i = wv; i < nel; i += nwv) {
            if (i == row) continue;
            double *Si = S + (size_t)i * a.NM;
            double sum = 0;
            for (int j = lane; j < nel; j += WAVE) sum += Si[j] * s_orb[j];
            sum = wave_sum(sum);
            sum = sum / rt;
            for (int j = lane; j < nel; j += WAVE) Si[j] = Si[j] - s_row[j] * sum;
        }
    for (int j = tid; j < nel; j += blockDim.x) S[(size_t)row * a.NM + j] = s_row[j] / rt;
    // detsav: orbital gradients of iel, spin determinant
    double *dor = a.dorb + ((size_t)w * NE + iel) * 3 * NO;
    for (int t = tid; t < 3 * NO; t += blockDim.x) dor[t] = ob[NO + t];
    if (tid == 0) a.detiab[2 * w + iab] = s_detn;
    // position and distances
    if (tid < 3) a.xT[((size_t)w * 3 + tid) * NE + iel] = a.xnew[3 * w + tid];
    for (int c = tid; c < a.NC; c += blockDim.x) {
        const double rn = a.rnew[(size_t)w * a.NC + c], rrn = a.rrnew[(size_t)w * a.NC + c];
        a.r_en[((size_t)w * a.NC + c) * NE + iel] = rn;
        a.rr_en[((size_t)w * a.NC + c) * NE + iel] = rrn;
        if (a.enp) a.enp[((size_t)w * a.NC + c) * NE + iel] = enp_pack(a.jas->kap, a.jas->asymp_r, rn, rrn);
    }
    // jassav (iflag = 0)
    const double *fsn = a.fsn_row + (size_t)w * NE;
    const double *g1n = a.g1n + (size_t)w * 3 * NE, *g2n = a.g2n + (size_t)w * 3 * NE;
    const double *vjn = a.vjnT + (size_t)w * 3 * NE;
    double *vj = a.vjT + (size_t)w * 3 * NE;
    double *F = a.fso + (size_t)w * NE * NE;
    double *G1 = a.g1 + (size_t)w * NE * 3 * NE, *G2 = a.g2 + (size_t)w * NE * 3 * NE;
    for (int j = tid; j < NE; j += blockDim.x) {
        const double v = fsn[j];
        F[(size_t)iel * NE + j] = v;
        F[(size_t)j * NE + iel] = v;
        for (int k = 0; k < 3; k++) {
            G1[((size_t)iel * 3 + k) * NE + j] = g1n[k * NE + j];
            G2[((size_t)iel * 3 + k) * NE + j] = g2n[k * NE + j];
            G1[((size_t)j * 3 + k) * NE + iel] = g2n[k * NE + j];
            G2[((size_t)j * 3 + k) * NE + iel] = g1n[k * NE + j];
            vj[k * NE + j] = vjn[k * NE + j];
        }
    }
    if (tid == 0) a.fsumo[w] = a.fsumn[w];
}

// ------------------------------------------------------------------------
// rotqua (rotqua.f90:357-430) per walker with its own stream.
__global__ void k_rotqua(int W, int nq, uint64_t *rng, const double *xq0, const double *yq0, const double *zq0,
                         double *xq, double *yq, double *zq)
{
    const int w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= W) return;
    uint64_t s[4];
    for (int k = 0; k < 4; k++) s[k] = rng[4 * (size_t)w + k];
    // rejection sampling as in rotqua; the trip cap (probability of reaching
    // it ~ 0.215^1000) only guarantees that every thread finishes
    double x1 = 0, x2 = 0, x3, xsum = 0, cfi, sfi, sthet;
    for (int tries = 0; tries < 1000; tries++) {
        x1 = 1.0 - 2.0 * xo_uniform(s);
        x2 = 1.0 - 2.0 * xo_uniform(s);
        xsum = x1 * x1 + x2 * x2;
        if (xsum < 1.0) break;
    }
    double xsum2 = 2.0 * sqrt(fabs(1.0 - xsum));
    x1 = x1 * xsum2; x2 = x2 * xsum2; x3 = 1.0 - 2.0 * xsum;
    double theta = acos(x3);
    sthet = sin(theta);
    if (sthet < 1.0e-05) { cfi = 1.0; sfi = 0.0; x1 = 0.0; x2 = 0.0; x3 = 1.0; sthet = 0.0; }
    else { cfi = x1 / sthet; sfi = x2 / sthet; }
    double yy1 = x3 * cfi, yy2 = x3 * sfi, yy3 = -sthet, zz1 = -sfi, zz2 = cfi, zz3 = 0.0;
    double u1 = 1, u2 = 0, usum = 1;
    for (int tries = 0; tries < 1000; tries++) {
        u1 = xo_uniform(s) * 2.0 - 1.0;
        u2 = xo_uniform(s) * 2.0 - 1.0;
        usum = u1 * u1 + u2 * u2;
        if (usum < 1.0) break;
    }
    double uu = sqrt(usum);
    u1 = u1 / uu; u2 = u2 / uu;
    double y1 = yy1 * u1 + zz1 * u2, y2 = yy2 * u1 + zz2 * u2, y3 = yy3 * u1 + zz3 * u2;
    double z1 = yy1 * u2 - zz1 * u1, z2 = yy2 * u2 - zz2 * u1, z3 = yy3 * u2 - zz3 * u1;
    for (int iq = 0; iq < nq; iq++) {
        xq[(size_t)w * nq + iq] = xq0[iq] * x1 + yq0[iq] * y1 + zq0[iq] * z1;
        yq[(size_t)w * nq + iq] = xq0[iq] * x2 + yq0[iq] * y2 + zq0[iq] * z2;
        zq[(size_t)w * nq + iq] = xq0[iq] * x3 + yq0[iq] * y3 + zq0[iq] * z3;
    }
    for (int k = 0; k < 4; k++) rng[4 * (size_t)w + k] = s[k];
}

// ------------------------------------------------------------------------
// launchers
static inline int nblocks(int n, int per) { return (n + per - 1) / per; }

// xnew of every walker = the current position of electron iel (a proposal of
// zero length; the autotuner times move steps on it)
__global__ void k_stay(int W, int NE, int iel, const double *xT, double *xnew)
{
    const int w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= W) return;
    for (int k = 0; k < 3; k++) xnew[3 * w + k] = xT[((size_t)w * 3 + k) * NE + iel];
}

int launch_stay(champ_ctx *c, int iel)
{
    hipLaunchKernelGGL(k_stay, dim3(nblocks(c->W, 256)), dim3(256), 0, c->stream, c->W, c->NE, iel, c->xT, c->xnew);
    HIPCHECK(hipGetLastError());
    return CHAMP_OK;
}

int launch_drift_old(champ_ctx *c, int iel)
{
    MoveArgs a = move_args(c, iel, 1, c->tau);
    hipLaunchKernelGGL(k_drift_old, dim3(nblocks(c->W, 4)), dim3(256), 0, c->stream, a);
    HIPCHECK(hipGetLastError());
    return CHAMP_OK;
}

template <typename T>
static int launch_move_a_t(champ_ctx *c, const MoveArgs &a, int nordc)
{
    const int naz[] = {0, 0, een_naz<2>(), een_naz<3>(), een_naz<4>(), een_naz<5>()};
    const size_t lds = (size_t)4 * move_a_lds_doubles(c->NC, nordc > 1 ? naz[nordc] : 0, c->NRI, (int)sizeof(T)) *
                       sizeof(double);
    const void *fn = nordc == 5 ? (const void *)k_move_a<5, T> : nordc == 4 ? (const void *)k_move_a<4, T>
                   : nordc == 3 ? (const void *)k_move_a<3, T> : nordc == 2 ? (const void *)k_move_a<2, T>
                                                                        : (const void *)k_move_a<0, T>;
    if (lds > 160 * 1024) {
        set_error("move: the move kernel's LDS tables exceed 160 KB for this system");
        return CHAMP_ERR_UNSUPPORTED;
    }
    if (lds > 64 * 1024) HIPCHECK(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    const dim3 grid(nblocks(c->W, 4)), block(256);
    switch (nordc) {
    case 5: hipLaunchKernelGGL((k_move_a<5, T>), grid, block, lds, c->stream, a); break;
    case 4: hipLaunchKernelGGL((k_move_a<4, T>), grid, block, lds, c->stream, a); break;
    case 3: hipLaunchKernelGGL((k_move_a<3, T>), grid, block, lds, c->stream, a); break;
    case 2: hipLaunchKernelGGL((k_move_a<2, T>), grid, block, lds, c->stream, a); break;
    default: hipLaunchKernelGGL((k_move_a<0, T>), grid, block, lds, c->stream, a); break;
    }
    HIPCHECK(hipGetLastError());
    return CHAMP_OK;
}

int launch_move_a(champ_ctx *c, int iel, int mode)
{
    MoveArgs a = move_args(c, iel, mode, c->tau);
    const int nordc = c->hjas.nordc;
    if (nordc > 5) {
        set_error("move: e-e-n order above 5 is not supported");
        return CHAMP_ERR_UNSUPPORTED;
    }
    int slot;
    tbegin(c, K_MOVE_A, &slot);
    const int rc = c->prec == CHAMP_PREC_MIXED ? launch_move_a_t<float>(c, a, nordc) : launch_move_a_t<double>(c, a, nordc);
    tend(c, slot);
    return rc;
}

int launch_move_gemm(champ_ctx *c)
{
    // orbn[w][comp][o] = sum_m aomove[w][comp][m] * coef[o][m]: FP64 matrix cores,
    // or (mixed path, when the autotuner finds it faster) FP32 operands and accumulation
    int slot;
    tbegin(c, K_MOVE_GEMM, &slot);
    int rc = (c->prec == CHAMP_PREC_MIXED && c->tune_gemm[0] == GEMM_F32)
                 ? launch_gemm_f32(c, 4 * c->W, c->NO, c->NB, (const float *)c->aomove, c->NB, c->coef32, c->NB,
                                   c->orbn, c->NO)
                 : launch_gemm(c, true, 4 * c->W, c->NO, c->NB, c->aomove, c->NB, 0, c->coef, c->NB, 0, c->orbn, c->NO,
                               0, 1, c->tune_gemm[0]);
    tend(c, slot);
    return rc;
}

int launch_move_b(champ_ctx *c, int iel, int mode)
{
    MoveArgs a = move_args(c, iel, mode, c->tau);
    size_t lds = (size_t)2 * c->NM * sizeof(double);
    int slot;
    tbegin(c, K_MOVE_B, &slot);
    // rows per wave for the larger spin block (4 waves per block)
    const int nmx = std::max(c->NUP, c->NDN), rpw = (nmx + 3) / 4;
    const dim3 grid(c->W), block(256);
    if (nmx > 2 * WAVE) hipLaunchKernelGGL(k_move_b<0>, grid, block, lds, c->stream, a);
    else if (rpw <= 4) hipLaunchKernelGGL(k_move_b<4>, grid, block, lds, c->stream, a);
    else if (rpw <= 8) hipLaunchKernelGGL(k_move_b<8>, grid, block, lds, c->stream, a);
    else if (rpw <= 12) hipLaunchKernelGGL(k_move_b<12>, grid, block, lds, c->stream, a);
    else if (rpw <= 16) hipLaunchKernelGGL(k_move_b<16>, grid, block, lds, c->stream, a);
    else if (rpw <= 20) hipLaunchKernelGGL(k_move_b<20>, grid, block, lds, c->stream, a);
    else if (rpw <= 24) hipLaunchKernelGGL(k_move_b<24>, grid, block, lds, c->stream, a);
    else hipLaunchKernelGGL(k_move_b<32>, grid, block, lds, c->stream, a);
    tend(c, slot);
    HIPCHECK(hipGetLastError());
    return CHAMP_OK;
}

int launch_rotqua(champ_ctx *c)
{
    hipLaunchKernelGGL(k_rotqua, dim3(nblocks(c->W, 256)), dim3(256), 0, c->stream, c->W, c->NQ, c->rng, c->xq0,
                       c->yq0, c->zq0, c->xq, c->yq, c->zq);
    HIPCHECK(hipGetLastError());
    return CHAMP_OK;
}


}  // namespace champ
