// hpsi_args.hpp -- argument block shared by the once-per-sweep (hpsi)
// kernels and their launchers.
#pragma once
#include "engine_internal.hpp"

namespace champ {

#define HB 0.5

struct HArgs {
    int W, NE, NUP, NM, NC, NB, NO, NQ, nloc, w0, nw, cap, accumulate, pad;
    double pecent, cdet0;
    const double *cent;
    const int *iwctype;
    SplineTab tab;
    const double *rmax_ao;
    const int *ao_cent, *ao_type, *ao_rad, *ao_slm, *iworbd;
    const JasParams *jas;
    const EcpParams *ecp;
    double *xT, *r_en, *rr_en, *fso, *g1, *g2, *vjT, *fsumo, *d2o, *dorb, *slmi, *detiab, *ddx, *d2dx2;
    double *psido, *psijo, *psi2o, *eold, *ekin, *pe_local, *vpsp;
    double *aoall, *orball, *wmat;
    int *pair_count, *pair_off, *pair_total, *pair_e, *pair_c, *nxquad;
    int *grp_off;  // [W+1] groups of NL_GROUP pairs per walker (prefix sums); pair_total[1] = groups
    double *pair_vps, *pair_val;
    double *pair_psi;  // [W*cap][NQ] psi ratios from k_nonloc_psi (psi_ext), else unused
    float2 *enp;       // mixed path: the FP32 partner pair table, else null
    int tstore;        // keep each point's t_vpsp (its term of the ECP sum) in pair_psi
    int psi_ext;
    double *rrT;  // [W][NE][NC] scaled e-n distances, centre index innermost
    int NRI;      // radial instances (centre, radial function)
    const int *ri_cent, *ri_type, *ri_rad, *ao_ri;
    const double *ri_rmax;
    const int *ri_perm, *ri_aoff, *ri_aol;  // instances by AO count; each instance's AOs (CSR)
    double *xq, *yq, *zq;
    double *esum_dev;
    int *status;
    int md;                 // multideterminant: psi_D = D_up,ref D_dn,ref md_S
    const double *md_S;     // [W]
    double *lu_ws;          // [2 hpsi_chunk][NM (NM+1)] Slater work matrices when they exceed LDS, else null
    int ocomp;              // components per electron in aoall / orball: 5 (value, gradient, Laplacian) or 1
};

inline HArgs hargs(champ_ctx *c)
{
    HArgs a;
    a.W = c->W; a.NE = c->NE; a.NUP = c->NUP; a.NM = c->NM; a.NC = c->NC; a.NB = c->NB; a.NO = c->NO; a.NQ = c->NQ;
    a.nloc = c->nloc; a.w0 = 0; a.nw = c->W; a.cap = c->pair_cap; a.accumulate = 0; a.pad = 0;
    a.pecent = c->pecent; a.cdet0 = c->cdet0;
    a.cent = c->cent; a.iwctype = c->iwctype; a.tab = c->tab; a.rmax_ao = c->rmax_ao;
    a.ao_cent = c->ao_cent; a.ao_type = c->ao_type; a.ao_rad = c->ao_rad; a.ao_slm = c->ao_slm; a.iworbd = c->iworbd;
    a.jas = c->jas; a.ecp = c->ecp;
    a.xT = c->xT; a.r_en = c->r_en; a.rr_en = c->rr_en; a.fso = c->fso; a.g1 = c->g1; a.g2 = c->g2; a.vjT = c->vjT;
    a.fsumo = c->fsumo; a.d2o = c->d2o; a.dorb = c->dorb; a.slmi = c->slmi; a.detiab = c->detiab; a.ddx = c->ddx;
    a.d2dx2 = c->d2dx2; a.psido = c->psido; a.psijo = c->psijo; a.psi2o = c->psi2o; a.eold = c->eold; a.ekin = c->ekin;
    a.pe_local = c->pe_local; a.vpsp = c->vpsp; a.aoall = c->aoall; a.orball = c->orball; a.wmat = c->wmat;
    a.pair_count = c->pair_count; a.pair_off = c->pair_off; a.pair_total = c->pair_total; a.pair_e = c->pair_e;
    a.pair_c = c->pair_c; a.nxquad = c->nxquad; a.grp_off = c->grp_off; a.pair_vps = c->pair_vps; a.pair_val = c->pair_val;
    a.pair_psi = c->pair_psi; a.psi_ext = 0;
    a.enp = c->enp;
    a.tstore = c->tstore;
    a.rrT = c->rrT;
    a.NRI = c->NRI; a.ri_cent = c->ri_cent; a.ri_type = c->ri_type; a.ri_rad = c->ri_rad; a.ao_ri = c->ao_ri;
    a.ri_rmax = c->ri_rmax;
    a.ri_perm = c->ri_perm; a.ri_aoff = c->ri_aoff; a.ri_aol = c->ri_aol;
    a.xq = c->xq; a.yq = c->yq; a.zq = c->zq; a.esum_dev = c->esum_dev; a.status = c->status;
    a.md = c->md; a.md_S = c->md_S;
    a.lu_ws = c->lu_ws;
    a.ocomp = 5;
    return a;
}

template <typename T>
__device__ T block_sum(T v, T *scratch)
{
    const int lane = threadIdx.x % WAVE, wv = threadIdx.x / WAVE, nwv = blockDim.x / WAVE;
    v = wave_sum(v);
    __syncthreads();
    if (lane == 0) scratch[wv] = v;
    __syncthreads();
    T s = 0;
    if (threadIdx.x == 0) for (int i = 0; i < nwv; i++) s += scratch[i];
    __syncthreads();
    return s;  // valid on thread 0
}

// yl0 (nonloc.f90:317-340)
__device__ __forceinline__ double yl0(int l, double c)
{
    if (l == 1) return 1.0;
    if (l == 2) return 3.0 * c;
    if (l == 3) return 2.5 * (3 * c * c - 1);
    if (l == 4) return 3.5 * (5 * c * c * c - 3 * c);
    return 0.0;
}

// static LDS of the Slater-inverse kernel k_hpsi_slater (perm[512], piv[512],
// s_p, s_bad, s_akk), counted with its dynamic work matrix when choosing
// between the LDS and the global-work-matrix forms
constexpr size_t SLATER_STATIC_LDS = 512 * sizeof(int) + 512 * sizeof(double) + 2 * sizeof(int) + sizeof(double) + 8;

// ECP quadrature over the active (electron, centre) pairs (engine_nonloc.hip)
int launch_nonloc(champ_ctx *c, const HArgs &a);
constexpr int NL_GROUP = 4;  // pairs of one walker per block of the matrix-core quadrature kernel

}  // namespace champ
