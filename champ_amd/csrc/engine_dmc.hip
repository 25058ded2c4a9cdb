// engine_dmc.hip -- diffusion Monte Carlo of the population: CHAMP's dmc_ps
// mover (dmc/dmc_ps_mov1.f90:4-838), or for all-electron systems (nloc = 0)
// its dmc_ae mover (dmc/dmc_ae_mov1.f90, dmc.f90:98-102; dmc_ae.hpp), with
// the default controls (idmc = 2,
// icasula = 0: the non-local pseudopotential in the locality approximation,
// ipq = 1, itau_eff = 1, iacc_rej = 1, icross = 1, icut_br = 0, icut_e = 0,
// limit_wt_dmc = 0, nforce = 1, node_cutoff = 0), the velocity-ratio
// reweighting of dmc/branching.f90, the population control of
// dmc/acues1.f90 (eigv, fprod over nfprod steps), the branching of
// dmc/splitj.f90 and the block estimators of dmc/acuest.f90 + finwrt.f90.
//
// The walkers stay on the GPU.  A step moves every electron of every walker
// (k_move_a / k_move_b with the DMC mover), runs hpsi, and k_dmc_weight forms
// each walker's new weight and its estimator terms.  The host sums those in
// walker order, runs splitj on the weights, and the GPU copies whole walker
// states from their sources to the slots splitj fills (k_dmc_copy); the live
// walkers occupy slots 0 .. nwalk-1 of the context's capacity.
//
// Streams: every walker slot keeps its own xoshiro stream (a copied walker
// continues with its new slot's stream); splitj's uniforms come from one
// population stream given at champ_gpu_dmc_init.  CHAMP draws all of them
// from one stream per MPI rank in walker order, which a population moved in
// parallel cannot reproduce; the oracle's DMC follows these stream rules.
#include "engine_internal.hpp"
#include "dmc_ae.hpp"
#include "hpsi_args.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

// one single-electron move of every walker with the context's mover (engine_api.hip)
int launch_move_step(champ_ctx *c, int iel);
int launch_forced_move(champ_ctx *c, int iel);

namespace champ {

int launch_hpsi_impl(champ_ctx *c, int accumulate);

// host side of the population (the module variables of CHAMP's DMC)
struct DmcState {
    int nconf = 0, nfprod = 100, ipass = 0, first = 1, iblk = 0;
    double tau = 0.05, etrial = 0.0;
    double eigv = 1.0, eest = 0.0, fprod = 1.0, wdsumo = 0.0;
    std::vector<double> ff, wtgen;
    uint64_t pop[4] = {0, 0, 0, 0};  // population stream (splitj)
    // per step (acues1)
    double wcum1 = 0, wfcum1 = 0, ecum1 = 0, efcum1 = 0, wcm21 = 0, wfcm21 = 0, ecm21 = 0, efcm21 = 0;
    double wgcum1 = 0, egcum1 = 0, wgcm21 = 0, egcm21 = 0;
    // per block (acuest)
    double wsum = 0, wfsum = 0, esum = 0, efsum = 0, wgsum = 0, egsum = 0, pesum = 0, tpbsum = 0, tausum = 0;
    // cumulative over blocks
    double wgcum = 0, egcum = 0, taucum = 0, wgcm2 = 0, egcm2 = 0, pecum = 0, tpbcum = 0;
    double acc = 0, trymove = 0, nodecr = 0;
    int64_t nbrnch = 0;
    int nwalk_max = 0;
    // the per-walker arrays a branching copy moves: device base and bytes per walker
    std::vector<std::pair<char *, size_t>> arrays;
    void *d_tab = nullptr;   // device copy of `arrays`
    int *d_pairs = nullptr;  // [2 W] (source, destination) slots of one step's copies
    // T-moves (icasula = -1): per walker the electron moved (-1: none) and its point
    int icasula = 0;
    int *d_tm_iel = nullptr;
    double *d_tm_x = nullptr;
};

void dmc_free(champ_ctx *c)
{
    if (c->dmc) {
        hipFree(c->dmc->d_tab);
        hipFree(c->dmc->d_pairs);
        if (c->dmc->d_tm_iel) hipFree(c->dmc->d_tm_iel);
        if (c->dmc->d_tm_x) hipFree(c->dmc->d_tm_x);
        delete c->dmc;
        c->dmc = nullptr;
    }
    void *ptrs[] = {c->dmc_acc, c->dmc_age, c->dmc_wt, c->dmc_fratio, c->dmc_eprev, c->dmc_out};
    for (void *p : ptrs) if (p) hipFree(p);
    c->dmc_acc = nullptr; c->dmc_age = nullptr; c->dmc_wt = nullptr; c->dmc_fratio = nullptr;
    c->dmc_eprev = nullptr; c->dmc_out = nullptr;
}

// ------------------------------------------------------------------------
// Per step, before the moves: the previous local energy and zeroed move
// accumulators (pp starts at 1).
__global__ void k_dmc_begin(int W, const double *eold, double *eprev, double *acc)
{
    const int w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= W) return;
    eprev[w] = eold[w];
    double *a = acc + (size_t)w * DMC_ACC;
    for (int k = 0; k < DMC_ACC; k++) a[k] = 0.0;
    a[DMC_PP] = 1.0;
}

struct DmcArgs {
    int W, NE, NC, nloc;
    double tau, etrial, eest, ffi, fprod;
    const double *ddx, *vjT, *eold, *eprev, *ekin, *acc;
    const double *xT, *cent;  // positions [W][3][NE], centres [NC][3] (the all-electron ratio)
    const int *iwctype;
    const EcpParams *ecp;
    double *wt, *fratio, *out;
    int *age;
};

// The velocity ratio of dmc_eloc_cutoff (branching.f90:226-278, tratio = 1):
// sqrt(sum vavvn^2 v^2 / sum v^2) over the electrons' drifts v = grad ln psi
// at the walker's configuration, with adrift 0.5 for nloc > 0 and, for
// nloc = 0, the nearest nucleus' adrift of the all-electron mover (adrift0
// 0.1).  One wavefront per walker.
// WEIGHT 0: only fratio (the first step, dmc_ps_mov1.f90:162-172);
// WEIGHT 1: the reweighting of the step (dmc_ps_mov1.f90:491-617:
// calculate_fratio, calculate_reweight with icut_e = 0, the population factor
// ffi) and the walker's estimator terms.
template <int WEIGHT>
__global__ void __launch_bounds__(256) k_dmc_weight(DmcArgs a)
{
    const int lane = threadIdx.x % WAVE;
    const int w = blockIdx.x * (blockDim.x / WAVE) + threadIdx.x / WAVE;
    if (w >= a.W) return;
    const double tau = a.tau, adrift = 0.5;
    const int NE = a.NE;
    double v2sum = 0.0, vav2sum = 0.0;
    for (int e = lane; e < NE; e += WAVE) {
        const double *dx = a.ddx + ((size_t)w * NE + e) * 3;
        const double *vj = a.vjT + (size_t)w * 3 * NE;
        const double v0 = dx[0] + vj[e], v1 = dx[1] + vj[NE + e], v2c = dx[2] + vj[2 * NE + e];
        if (a.nloc == 0) {
            const double *xw = a.xT + (size_t)w * 3 * NE;
            const double xe[3] = {xw[e], xw[NE + e], xw[2 * NE + e]}, ve[3] = {v0, v1, v2c};
            ae_fratio_terms(a.cent, a.iwctype, a.ecp->znuc, a.NC, tau, xe, ve, vav2sum, v2sum);
            continue;
        }
        const double v2 = v0 * v0 + v1 * v1 + v2c * v2c;
        const double vavvt = (sqrt(1.0 + 2.0 * adrift * v2 * tau) - 1.0) / (adrift * v2);
        const double vavvn = vavvt / tau;
        vav2sum += vavvn * vavvn * v2;
        v2sum += v2;
    }
    v2sum = wave_sum(v2sum);
    vav2sum = wave_sum(vav2sum);
    if (lane != 0) return;
    const double fration = sqrt(vav2sum / v2sum);
    if (WEIGHT) {
        const double *ac = a.acc + (size_t)w * DMC_ACC;
        const double taunow = tau * ac[DMC_DFUS2AC] / ac[DMC_DFUS2UN];  // tauprim, drifdifr = 1
        const double enew = a.eold[w], eo = a.eprev[w];
        const double deo = a.eest - eo, den = a.eest - enew;
        const double ewto = a.eest - deo * a.fratio[w];
        const double ewtn = a.eest - den * fration;
        const double expon = (a.etrial - 0.5 * (ewto + ewtn)) * taunow;
        const double dwt = exp(expon) * a.ffi;
        const double wtnow = a.wt[w] * dwt;
        a.wt[w] = wtnow;
        const double wtg = wtnow * a.fprod;
        double *o = a.out + (size_t)w * DMC_OUT;
        o[DMC_O_WT] = wtnow;
        o[DMC_O_WE] = wtnow * enew;
        o[DMC_O_PE] = wtg * (enew - a.ekin[w]);
        o[DMC_O_TPB] = wtg * a.ekin[w];
        o[DMC_O_TAU] = wtg * taunow;
        o[DMC_O_PSUM] = ac[DMC_PSUM];
        o[DMC_O_NODECR] = ac[DMC_NODECR];
        if (ac[DMC_IACCEPT] == 0.0) a.age[w] = a.age[w] + 1;
    }
    a.fratio[w] = fration;
}

// Copies of whole walker states: block (pair, array) copies array `y` of walker
// pairs[2x] into slot pairs[2x+1] in 4-byte words.
struct CopyTab {
    char *base;
    long long bytes;
};

__global__ void k_dmc_copy(const int *pairs, const CopyTab *tab)
{
    const int p = blockIdx.x, y = blockIdx.y;
    const int src = pairs[2 * p], dst = pairs[2 * p + 1];
    const long long nb = tab[y].bytes, nw = nb / 4;
    const uint32_t *s = (const uint32_t *)(tab[y].base + (size_t)src * nb);
    uint32_t *d = (uint32_t *)(tab[y].base + (size_t)dst * nb);
    for (long long i = threadIdx.x; i < nw; i += blockDim.x) d[i] = s[i];
}

static DmcArgs dmc_args(champ_ctx *c)
{
    DmcState &D = *c->dmc;
    DmcArgs a;
    a.W = c->W; a.NE = c->NE; a.tau = D.tau; a.etrial = D.etrial; a.eest = D.eest; a.ffi = 1.0; a.fprod = D.fprod;
    a.ddx = c->ddx; a.vjT = c->vjT; a.eold = c->eold; a.eprev = c->dmc_eprev; a.ekin = c->ekin; a.acc = c->dmc_acc;
    a.wt = c->dmc_wt; a.fratio = c->dmc_fratio; a.out = c->dmc_out; a.age = c->dmc_age;
    a.NC = c->NC; a.nloc = c->nloc; a.xT = c->xT; a.cent = c->cent; a.iwctype = c->iwctype; a.ecp = c->ecp;
    return a;
}

static double pop_uniform(DmcState &D)
{
    // xoshiro256** (Blackman and Vigna) and CHAMP's random_dp: the top 53 bits
    uint64_t *s = D.pop;
    const uint64_t r = ((s[1] * 5) << 7 | (s[1] * 5) >> 57) * 9;
    const uint64_t t = s[1] << 17;
    s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3]; s[2] ^= t;
    s[3] = (s[3] << 45) | (s[3] >> 19);
    return (double)(r >> 11) * 0x1.0p-53;
}

// splitj (splitj.f90:3-233) on the weights of the nwalk live walkers: pairs of
// light walkers (wt < 1/2) merge, heavy ones (wt >= 2) split into a dead or
// new slot, and the last walkers fill the remaining dead slots.  Returns the
// new walker count; `copies` receives each filled slot with the slot whose
// state (at the start of the step) it now holds.
static int splitj(DmcState &D, std::vector<double> &wt, std::vector<int> &age, int nwalk, int cap,
                  std::vector<std::pair<int, int>> &copies, std::string &err)
{
    std::vector<int> iwundr;
    int ipair = 0, iw2 = -1;
    for (int iw = 0; iw < nwalk; iw++) {
        if (wt[iw] < 0.5) {
            if (wt[iw] == 0.0) {
                D.nbrnch++;
                iwundr.push_back(iw);
            } else if (ipair == 0) {
                ipair = 1;
                iw2 = iw;
            } else {
                D.nbrnch++;
                ipair = 0;
                const double wttot = wt[iw] + wt[iw2];
                if (pop_uniform(D) > wt[iw] / wttot) {
                    wt[iw2] = wttot;
                    iwundr.push_back(iw);
                } else {
                    wt[iw] = wttot;
                    iwundr.push_back(iw2);
                }
            }
        }
    }
    // content[s]: the slot whose state slot s holds (states move only as copies)
    std::vector<int> content(cap);
    for (int s = 0; s < cap; s++) content[s] = s;
    int nwalk2 = nwalk;
    for (int iw = 0; iw < nwalk; iw++) {
        if (wt[iw] >= 2.0) {
            D.nbrnch++;
            int dst;
            if (!iwundr.empty()) {
                dst = iwundr.back();
                iwundr.pop_back();
            } else {
                dst = nwalk2++;
                if (nwalk2 > cap) { err = "dmc: splitj: walker capacity exceeded (MWALK)"; return -1; }
            }
            wt[iw] = wt[iw] * 0.5;
            wt[dst] = wt[iw];
            age[dst] = age[iw];
            content[dst] = content[iw];
        }
    }
    for (int j = (int)iwundr.size() - 1; j >= 0; j--) {
        const int dst = iwundr[j], src = nwalk2 - 1;
        nwalk2--;
        wt[dst] = wt[src];
        age[dst] = age[src];
        content[dst] = content[src];
    }
    copies.clear();
    for (int s = 0; s < nwalk2; s++)
        if (content[s] != s) copies.emplace_back(content[s], s);
    return nwalk2;
}

}  // namespace champ

using namespace champ;

extern "C" int champ_gpu_dmc_init(champ_ctx *c, int32_t nconf, double tau, double etrial, int32_t nfprod,
                                  const uint64_t pop_state[4])
{
    if (!c || nconf < 1 || nconf > (c->dmc_on ? c->dmc_W : c->W) || !(tau > 0) || !std::isfinite(etrial) ||
        nfprod < 1 || !pop_state) {
        set_error("dmc_init: 1 <= nconf <= walkers, tau > 0, finite etrial, nfprod >= 1 and a population stream required");
        return CHAMP_ERR_ARG;
    }
    if (!(pop_state[0] | pop_state[1] | pop_state[2] | pop_state[3])) {
        set_error("dmc_init: the population stream is all-zero");
        return CHAMP_ERR_ARG;
    }
    if (!c->dmc_on) {
        c->dmc_W = c->W;
        const size_t W = c->W;
        HIPCHECK(hipMalloc(&c->dmc_acc, W * DMC_ACC * sizeof(double)));
        HIPCHECK(hipMalloc(&c->dmc_age, W * sizeof(int)));
        HIPCHECK(hipMalloc(&c->dmc_wt, W * sizeof(double)));
        HIPCHECK(hipMalloc(&c->dmc_fratio, W * sizeof(double)));
        HIPCHECK(hipMalloc(&c->dmc_eprev, W * sizeof(double)));
        HIPCHECK(hipMalloc(&c->dmc_out, W * DMC_OUT * sizeof(double)));
        c->dmc = new DmcState();
        DmcState &D = *c->dmc;
        const size_t NE = c->NE, NC = c->NC, NO = c->NO, NM = c->NM;
        auto add = [&](void *p, size_t bytes) { if (p) D.arrays.emplace_back((char *)p, bytes); };
        add(c->xT, 3 * NE * 8); add(c->r_en, NC * NE * 8); add(c->rr_en, NC * NE * 8); add(c->fso, NE * NE * 8);
        add(c->g1, NE * 3 * NE * 8); add(c->g2, NE * 3 * NE * 8); add(c->vjT, 3 * NE * 8); add(c->fsumo, 8);
        add(c->d2o, 8); add(c->dorb, NE * 3 * NO * 8); add(c->slmi, 2 * NM * NM * 8); add(c->detiab, 16);
        add(c->ddx, NE * 3 * 8); add(c->d2dx2, NE * 8); add(c->psido, 8); add(c->psijo, 8); add(c->psi2o, 8);
        add(c->eold, 8); add(c->ekin, 8); add(c->pe_local, 8); add(c->vpsp, 16); add(c->status, 4);
        add(c->dmc_age, 4); add(c->dmc_fratio, 8);
        add(c->enp, NC * NE * 8);  // mixed path only (null otherwise)
        if (c->md) {
            add(c->md_aa, (size_t)2 * 2 * NM * c->md_NV * 8); add(c->md_D, (size_t)2 * 2 * c->md_NU * 8);
            add(c->md_S, 8); add(c->md_cur, 8); add(c->md_V, NE * NO * 8);
        }
        std::vector<CopyTab> tab;
        for (auto &p : D.arrays) {
            if (p.second % 4) { set_error("dmc_init: a walker array is not a whole number of words"); return CHAMP_ERR_STATE; }
            tab.push_back(CopyTab{p.first, (long long)p.second});
        }
        HIPCHECK(hipMalloc(&D.d_tab, tab.size() * sizeof(CopyTab)));
        HIPCHECK(hipMemcpy(D.d_tab, tab.data(), tab.size() * sizeof(CopyTab), hipMemcpyHostToDevice));
        HIPCHECK(hipMalloc(&D.d_pairs, 2 * W * sizeof(int)));
        HIPCHECK(hipMalloc(&D.d_tm_iel, W * sizeof(int)));
        HIPCHECK(hipMalloc(&D.d_tm_x, 3 * W * sizeof(double)));
        c->dmc_on = 1;
    }
    DmcState &D = *c->dmc;
    D.nconf = nconf; D.tau = tau; D.etrial = etrial; D.nfprod = nfprod;
    D.ipass = 0; D.first = 1; D.iblk = 0;
    D.eigv = 1.0; D.eest = etrial; D.fprod = 1.0; D.wdsumo = nconf;
    D.ff.assign(nfprod + 1, 1.0);
    D.wtgen.assign(nfprod + 1, (double)nconf);
    std::memcpy(D.pop, pop_state, sizeof(D.pop));
    D.nwalk_max = nconf;
    // init.f90: the first nconf walkers with weight 1, their local energies and drifts (hpsi)
    c->W = nconf;
    std::vector<double> one(nconf, 1.0);
    HIPCHECK(hipMemcpy(c->dmc_wt, one.data(), nconf * sizeof(double), hipMemcpyHostToDevice));
    HIPCHECK(hipMemset(c->dmc_age, 0, nconf * sizeof(int)));
    int rc = launch_hpsi_impl(c, 0);
    if (rc) return rc;
    // zerest
    D.wcum1 = D.wfcum1 = D.ecum1 = D.efcum1 = D.wcm21 = D.wfcm21 = D.ecm21 = D.efcm21 = 0;
    D.wgcum1 = D.egcum1 = D.wgcm21 = D.egcm21 = 0;
    D.wsum = D.wfsum = D.esum = D.efsum = D.wgsum = D.egsum = D.pesum = D.tpbsum = D.tausum = 0;
    D.wgcum = D.egcum = D.taucum = D.wgcm2 = D.egcm2 = D.pecum = D.tpbcum = 0;
    D.acc = D.trymove = D.nodecr = 0;
    D.nbrnch = 0;
    HIPCHECK(hipStreamSynchronize(c->stream));
    return CHAMP_OK;
}

// zerest (dmc/zerest.f90): cumulative and block estimators and the ages
static int dmc_zerest(champ_ctx *c)
{
    DmcState &D = *c->dmc;
    D.iblk = 0;
    D.wcum1 = D.wfcum1 = D.ecum1 = D.efcum1 = D.wcm21 = D.wfcm21 = D.ecm21 = D.efcm21 = 0;
    D.wgcum1 = D.egcum1 = D.wgcm21 = D.egcm21 = 0;
    D.wsum = D.wfsum = D.esum = D.efsum = D.wgsum = D.egsum = D.pesum = D.tpbsum = D.tausum = 0;
    D.wgcum = D.egcum = D.taucum = D.wgcm2 = D.egcm2 = D.pecum = D.tpbcum = 0;
    D.acc = D.trymove = D.nodecr = 0;
    D.nbrnch = 0;
    HIPCHECK(hipMemsetAsync(c->dmc_age, 0, (size_t)c->W * sizeof(int), c->stream));
    return CHAMP_OK;
}

// ------------------------------------------------------------------------
// T-moves with icasula = -1 (nonloc_grid.f90:60-137 over all electrons,
// dmc_ps_mov1.f90:701-730): a thread per walker.  t_vpsp of every active
// (electron, centre) pair and quadrature point is in tv (written by the
// step's hpsi); positive ones count as zero.  t_norm = 1 + tau sum(-t); the
// walker's stream draws p (only when t_norm != 1) and the first (electron,
// point, centre), in that order, whose cumulative -tau t / t_norm exceeds p
// is the move.  The pair list is electron-major with the centres ascending.
__global__ void k_tmove_select(int W, int NE, int NC, int NQ, int cap, double tau, const int *pair_count,
                               const int *pair_e, const int *pair_c, const double *tv, const double *r_en,
                               const double *cent, const double *xq, const double *yq, const double *zq,
                               uint64_t *rng, int *tm_iel, double *tm_x)
{
    const int w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= W) return;
    const int n = pair_count[w];
    const size_t base = (size_t)w * cap;
    double tn = 0.0;
    for (int p0 = 0; p0 < n;) {
        const int e = pair_e[base + p0];
        int p1 = p0;
        while (p1 < n && pair_e[base + p1] == e) p1++;
        for (int iq = 0; iq < NQ; iq++)
            for (int p = p0; p < p1; p++) {
                double t = tv[(base + p) * NQ + iq];
                if (t > 0.0) t = 0.0;
                tn = tn - t;
            }
        p0 = p1;
    }
    tn = 1.0 + tn * tau;
    tm_iel[w] = -1;
    if (tn == 1.0) return;
    const double tni = 1.0 / tn;
    uint64_t s[4];
    for (int k = 0; k < 4; k++) s[k] = rng[4 * (size_t)w + k];
    const double u = xo_uniform(s);
    for (int k = 0; k < 4; k++) rng[4 * (size_t)w + k] = s[k];
    double cum = 0.0;
    for (int p0 = 0; p0 < n;) {
        const int e = pair_e[base + p0];
        int p1 = p0;
        while (p1 < n && pair_e[base + p1] == e) p1++;
        for (int iq = 0; iq < NQ; iq++)
            for (int p = p0; p < p1; p++) {
                double t = tv[(base + p) * NQ + iq];
                if (t > 0.0) t = 0.0;
                cum = cum - tau * t * tni;
                if (cum > u) {
                    const int ic = pair_c[base + p];
                    const double ren = r_en[((size_t)w * NC + ic) * NE + e];
                    const double q[3] = {xq[(size_t)w * NQ + iq], yq[(size_t)w * NQ + iq], zq[(size_t)w * NQ + iq]};
                    for (int k = 0; k < 3; k++) tm_x[3 * (size_t)w + k] = ren * q[k] + cent[3 * ic + k];
                    tm_iel[w] = e;
                    return;
                }
            }
        p0 = p1;
    }
}

// the walkers whose T-move moves electron iel: its point as xnew and a
// commit; every other walker proposes its electron's own position, unused
__global__ void k_tmove_prep(int W, int NE, int iel, const int *tm_iel, const double *tm_x, const double *xT,
                             double *xnew, unsigned char *acc)
{
    const int w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= W) return;
    const bool mv = tm_iel[w] == iel;
    for (int k = 0; k < 3; k++) xnew[3 * w + k] = mv ? tm_x[3 * (size_t)w + k] : xT[((size_t)w * 3 + k) * NE + iel];
    acc[w] = mv ? 1 : 0;
}

extern "C" int champ_gpu_dmc_step(champ_ctx *c, champ_dmc_step_stats *st)
{
    if (!c || !c->dmc_on) { set_error("dmc_step: champ_gpu_dmc_init first"); return CHAMP_ERR_STATE; }
    DmcState &D = *c->dmc;
    int rc;
    // population control factors (dmc_ps_mov1.f90:149-156)
    D.ipass++;
    const int ipmod = D.ipass % D.nfprod;
    const double ginv = std::min(1.0, D.tau);
    const double ffn = D.eigv * std::pow(D.wdsumo / D.nconf, ginv), ffi = 1.0 / ffn;
    D.fprod = D.fprod * ffn / D.ff[ipmod];
    D.ff[ipmod] = ffn;
    const int W = c->W;
    DmcArgs da = dmc_args(c);
    if (D.first) {
        hipLaunchKernelGGL(k_dmc_weight<0>, dim3((W + 3) / 4), dim3(256), 0, c->stream, da);
        HIPCHECK(hipGetLastError());
        D.first = 0;
    }
    if (c->nloc > 0 && (rc = launch_rotqua(c))) return rc;
    hipLaunchKernelGGL(k_dmc_begin, dim3((W + 255) / 256), dim3(256), 0, c->stream, W, c->eold, c->dmc_eprev,
                       c->dmc_acc);
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipMemsetAsync(c->esum_dev, 0, sizeof(double) * 4, c->stream));
    c->tau = D.tau;
    c->imetro = 2;
    for (int iel = 0; iel < c->NE; iel++)
        if ((rc = launch_move_step(c, iel))) { c->imetro = 1; return rc; }
    c->imetro = 1;
    const bool tmoves = D.icasula < 0 && c->nloc > 0;
    c->tstore = tmoves ? 1 : 0;
    rc = launch_hpsi_impl(c, 0);
    c->tstore = 0;
    if (rc) return rc;
    da = dmc_args(c);
    da.ffi = ffi;
    hipLaunchKernelGGL(k_dmc_weight<1>, dim3((W + 3) / 4), dim3(256), 0, c->stream, da);
    HIPCHECK(hipGetLastError());
    std::vector<int> tm_iel;
    if (tmoves) {
        hipLaunchKernelGGL(k_tmove_select, dim3((W + 255) / 256), dim3(256), 0, c->stream, W, c->NE, c->NC, c->NQ,
                           c->pair_cap, D.tau, c->pair_count, c->pair_e, c->pair_c, c->pair_psi, c->r_en, c->cent,
                           c->xq, c->yq, c->zq, c->rng, D.d_tm_iel, D.d_tm_x);
        HIPCHECK(hipGetLastError());
        tm_iel.resize(W);
        HIPCHECK(hipMemcpyAsync(tm_iel.data(), D.d_tm_iel, (size_t)W * sizeof(int), hipMemcpyDeviceToHost, c->stream));
    }
    std::vector<double> out((size_t)W * DMC_OUT), wt(c->dmc_W);
    std::vector<int> age(c->dmc_W);
    HIPCHECK(hipMemcpyAsync(out.data(), c->dmc_out, out.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHECK(hipMemcpyAsync(wt.data(), c->dmc_wt, (size_t)W * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHECK(hipMemcpyAsync(age.data(), c->dmc_age, (size_t)W * sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHECK(hipStreamSynchronize(c->stream));
    // sums over the walkers in walker order (dmc_ps_mov1.f90:606-617)
    double wsum1 = 0, esum1 = 0, psum = 0, nod = 0;
    for (int w = 0; w < W; w++) {
        const double *o = &out[(size_t)w * DMC_OUT];
        wsum1 += o[DMC_O_WT];
        esum1 += o[DMC_O_WE];
        D.pesum += o[DMC_O_PE];
        D.tpbsum += o[DMC_O_TPB];
        D.tausum += o[DMC_O_TAU];
        psum += o[DMC_O_PSUM];
        nod += o[DMC_O_NODECR];
    }
    D.acc += psum;
    D.trymove += (double)W * c->NE;
    D.nodecr += nod;
    const double wfsum1 = wsum1 * ffn, efsum1 = esum1 * ffn;
    const double wgsum1 = wsum1 * D.fprod, egsum1 = esum1 * D.fprod;
    // the T-moves, electron by electron over the electrons some walker moves
    if (tmoves) {
        std::vector<char> used(c->NE, 0);
        for (int w = 0; w < W; w++)
            if (tm_iel[w] >= 0) used[tm_iel[w]] = 1;
        for (int iel = 0; iel < c->NE; iel++) {
            if (!used[iel]) continue;
            hipLaunchKernelGGL(k_tmove_prep, dim3((W + 255) / 256), dim3(256), 0, c->stream, W, c->NE, iel, D.d_tm_iel,
                               D.d_tm_x, c->xT, c->xnew, c->acc);
            HIPCHECK(hipGetLastError());
            if ((rc = launch_forced_move(c, iel))) return rc;
        }
    }
    // branching
    std::vector<std::pair<int, int>> copies;
    std::string err;
    const int nw2 = splitj(D, wt, age, W, c->dmc_W, copies, err);
    if (nw2 < 0) { set_error(err); return CHAMP_ERR_NOMEM; }
    if (nw2 == 0) { set_error("dmc_step: the population died out (every weight zero)"); return CHAMP_ERR_STATE; }
    if (!copies.empty()) {
        std::vector<int> pr;
        for (auto &p : copies) { pr.push_back(p.first); pr.push_back(p.second); }
        HIPCHECK(hipMemcpyAsync(D.d_pairs, pr.data(), pr.size() * sizeof(int), hipMemcpyHostToDevice, c->stream));
        hipLaunchKernelGGL(k_dmc_copy, dim3((unsigned)copies.size(), (unsigned)D.arrays.size()), dim3(256), 0,
                           c->stream, D.d_pairs, (const CopyTab *)D.d_tab);
        HIPCHECK(hipGetLastError());
    }
    HIPCHECK(hipMemcpyAsync(c->dmc_wt, wt.data(), (size_t)nw2 * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIPCHECK(hipStreamSynchronize(c->stream));  // the host vectors above are sources of the copies
    c->W = nw2;
    D.nwalk_max = std::max(D.nwalk_max, nw2);
    // acues1 (dmc/acues1.f90:34-95), idmc = 2
    D.wcum1 += wsum1; D.wfcum1 += wfsum1; D.ecum1 += esum1; D.efcum1 += efsum1;
    D.wcm21 += wsum1 * wsum1; D.wfcm21 += wfsum1 * wfsum1;
    D.ecm21 += esum1 * esum1 / wsum1; D.efcm21 += efsum1 * efsum1 / wfsum1;
    D.wgcum1 += wgsum1; D.egcum1 += egsum1; D.wgcm21 += wgsum1 * wgsum1; D.egcm21 += egsum1 * egsum1 / wgsum1;
    D.wsum += wsum1; D.wfsum += wfsum1; D.esum += esum1; D.efsum += efsum1;
    D.wgsum += wgsum1; D.egsum += egsum1;
    D.eest = (D.egcum + D.egsum) / (D.wgcum + D.wgsum);
    D.eigv = std::exp((D.etrial - D.eest) * (D.taucum + D.tausum) / (D.wgcum + D.wgsum));
    D.wdsumo = wsum1;
    D.wtgen[ipmod] = wsum1;
    if (st) {
        st->nwalk = nw2;
        st->wsum = wsum1;
        st->esum = esum1;
        st->wgsum = wgsum1;
        st->egsum = egsum1;
        st->eest = D.eest;
        st->eigv = D.eigv;
        st->fprod = D.fprod;
        st->accept = psum / ((double)W * c->NE);
        st->nodecr = (int64_t)nod;
        st->nbranch = D.nbrnch;
    }
    return CHAMP_OK;
}

// acuest (dmc/acuest.f90:53-352) at the end of a block: the block's growth
// estimator egnow, then the cumulative sums
static void dmc_acuest(champ_ctx *c, double *egnow)
{
    DmcState &D = *c->dmc;
    D.iblk++;
    const double en = D.egsum / D.wgsum;
    if (egnow) *egnow = en;
    D.wgcum += D.wgsum; D.egcum += D.egsum; D.taucum += D.tausum;
    D.wgcm2 += D.wgsum * D.wgsum; D.egcm2 += D.egsum * en;
    D.pecum += D.pesum; D.tpbcum += D.tpbsum;
    D.wsum = D.wfsum = D.esum = D.efsum = 0;
    D.wgsum = D.egsum = D.pesum = D.tpbsum = D.tausum = 0;
}

extern "C" int champ_gpu_dmc(champ_ctx *c, int32_t nblk, int32_t nstep, int32_t nblkeq, double *egnow,
                             champ_dmc_result *res)
{
    if (!c || !c->dmc_on) { set_error("dmc: champ_gpu_dmc_init first"); return CHAMP_ERR_STATE; }
    if (nblk < 1 || nstep < 1 || nblkeq < 0) { set_error("dmc: nblk, nstep >= 1 and nblkeq >= 0 required"); return CHAMP_ERR_ARG; }
    DmcState &D = *c->dmc;
    int rc;
    // dmc.f90:60-108: 2 nblkeq equilibration blocks, zerest after each half
    for (int i = 1; i <= nblk + 2 * nblkeq; i++) {
        if ((i == nblkeq + 1 || i == 2 * nblkeq + 1) && (rc = dmc_zerest(c))) return rc;
        for (int j = 0; j < nstep; j++)
            if ((rc = champ_gpu_dmc_step(c, nullptr))) return rc;
        double en = 0;
        dmc_acuest(c, &en);
        if (egnow && i > 2 * nblkeq) egnow[i - 2 * nblkeq - 1] = en;
    }
    if (res) {
        const double egave = D.egcum / D.wgcum;
        const double rn = D.wgcum * D.wgcum / D.wgcm2;
        res->egave = egave;
        res->egerr = std::sqrt(std::max((D.egcm2 / D.wgcum - egave * egave) / (rn - 1.0), 0.0));
        res->eest = D.eest;
        res->accept = D.trymove > 0 ? D.acc / D.trymove : 0.0;
        res->tau_eff = D.taucum / D.wgcum;
        res->nwalk = c->W;
        res->nwalk_max = D.nwalk_max;
        res->nbranch = D.nbrnch;
        res->nodecr = (int64_t)D.nodecr;
    }
    return CHAMP_OK;
}

extern "C" int champ_gpu_dmc_weights(champ_ctx *c, double *wt, int32_t *nwalk)
{
    if (!c || !c->dmc_on) { set_error("dmc_weights: champ_gpu_dmc_init first"); return CHAMP_ERR_STATE; }
    if (nwalk) *nwalk = c->W;
    if (wt) HIPCHECK(hipMemcpy(wt, c->dmc_wt, (size_t)c->W * sizeof(double), hipMemcpyDeviceToHost));
    return CHAMP_OK;
}

extern "C" int champ_gpu_dmc_set_icasula(champ_ctx *c, int32_t icasula)
{
    if (!c || !c->dmc_on) { set_error("dmc_set_icasula: champ_gpu_dmc_init first"); return CHAMP_ERR_STATE; }
    if (icasula != 0 && icasula != -1) { set_error("dmc_set_icasula: icasula 0 or -1 is provided"); return CHAMP_ERR_UNSUPPORTED; }
    c->dmc->icasula = icasula;
    return CHAMP_OK;
}
