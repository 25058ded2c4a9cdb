// engine_api.hip -- the extern "C" boundary declared in include/champ_gpu.h.
// Context set-up uploads the CHAMP tables once; every other entry point acts
// on the whole walker population on the context's stream.
#include "engine_internal.hpp"
#include "hpsi_args.hpp"
#include "metrop6.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

namespace champ {
static thread_local std::string g_err;
void set_error(const std::string &s) { g_err = s; }
int launch_hpsi_impl(champ_ctx *c, int accumulate);

void tbegin(champ_ctx *c, int cls, int *slot)
{
    *slot = -1;
    if (!c->timing) return;
    size_t idx = c->ev_used.size();
    if (idx >= c->ev_pool.size()) {
        Timer t;
        hipEventCreate(&t.a);
        hipEventCreate(&t.b);
        c->ev_pool.push_back(t);
    }
    c->ev_used.push_back({cls, (int)idx});
    hipEventRecord(c->ev_pool[idx].a, c->stream);
    *slot = (int)idx;
}
void tend(champ_ctx *c, int slot)
{
    if (slot < 0) return;
    hipEventRecord(c->ev_pool[slot].b, c->stream);
}
static void tflush(champ_ctx *c)
{
    if (!c->timing || c->ev_used.empty()) return;
    hipStreamSynchronize(c->stream);
    for (auto &u : c->ev_used) {
        float ms = 0;
        hipEventElapsedTime(&ms, c->ev_pool[u.second].a, c->ev_pool[u.second].b);
        c->class_ms[u.first] += ms;
        c->class_n[u.first] += 1;
    }
    c->ev_used.clear();
}

// host xoshiro256** for seeding (Blackman & Vigna 2018)
static inline uint64_t hrotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
static uint64_t hnext(uint64_t s[4])
{
    const uint64_t result = hrotl(s[1] * 5, 7) * 9;
    const uint64_t t = s[1] << 17;
    s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3];
    s[2] ^= t;
    s[3] = hrotl(s[3], 45);
    return result;
}
static void hjump(uint64_t s[4])
{
    static const uint64_t J[4] = {0x180ec6d33cfd0abaULL, 0xd5a61266f0c9392cULL, 0xa9582618e03fc9aaULL,
                                  0x39abdc4529b1661cULL};
    uint64_t t[4] = {0, 0, 0, 0};
    for (int i = 0; i < 4; i++)
        for (int b = 0; b < 64; b++) {
            if (J[i] & ((uint64_t)1 << b)) for (int k = 0; k < 4; k++) t[k] ^= s[k];
            hnext(s);
        }
    for (int k = 0; k < 4; k++) s[k] = t[k];
}
static double hpowi(double x, int m)
{
    unsigned n = m < 0 ? (unsigned)(-m) : (unsigned)m;
    double y = (n % 2) ? x : 1.0;
    while (n >>= 1) { x = x * x; if (n % 2) y = y * x; }
    return m < 0 ? 1.0 / y : y;
}

// node radii of every grid-3 interval (see grid_refs)
__global__ void k_grid_refs(int nr, double r0, double arg, double *gref)
{
    const int jx = blockIdx.x * blockDim.x + threadIdx.x;
    if (jx >= nr) return;
    double ref0, ref1;
    grid_refs(r0, arg, jx, ref0, ref1);
    gref[2 * jx] = ref0;
    gref[2 * jx + 1] = ref1;
}

template <typename T>
static int dalloc(T **p, size_t n)
{
    if (n == 0) n = 1;
    hipError_t e = hipMalloc((void **)p, n * sizeof(T));
    if (e != hipSuccess) { set_error(std::string("hipMalloc: ") + hipGetErrorString(e)); return CHAMP_ERR_NOMEM; }
    return CHAMP_OK;
}
template <typename T>
static int upload(T **p, const T *h, size_t n)
{
    int rc = dalloc(p, n);
    if (rc) return rc;
    if (n) HIPCHECK(hipMemcpy(*p, h, n * sizeof(T), hipMemcpyHostToDevice));
    return CHAMP_OK;
}
}  // namespace champ

using namespace champ;

// LDS bytes up to which the Slater inverse works in LDS (variant builds lower
// it to exercise the global work-matrix form on small systems)
#ifndef CHAMP_SLATER_LDS_MAX
#define CHAMP_SLATER_LDS_MAX (160 * 1024)
#endif
#define TRY(x) do { int rc_ = (x); if (rc_) { champ_gpu_free(c); *ctxp = nullptr; return rc_; } } while (0)

extern "C" const char *champ_gpu_last_error(void) { return g_err.c_str(); }

// ------------------------------------------------------------------------
// Static scratch of champ_gpu_basis_fns_noalloc: the staged positions and AO
// values of one call, 16 MiB per device, held until the process ends.  In
// host code `scratch` names a host-side stand-in, not device memory: its
// address on each device comes from hipGetSymbolAddress, taken once per
// device at champ_gpu_init with that device current.
constexpr size_t NOALLOC_DOUBLES = (size_t)2 << 20;  // 16 MiB
static __device__ __align__(256) double scratch[NOALLOC_DOUBLES];
constexpr int NOALLOC_MAXDEV = 64;
static std::mutex g_scratch_mu;
static double *g_scratch_dev[NOALLOC_MAXDEV];

static int noalloc_scratch(int dev, double **out)
{
    if (dev < 0 || dev >= NOALLOC_MAXDEV) { set_error("init: device index beyond the scratch table"); return CHAMP_ERR_UNSUPPORTED; }
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            set_error(std::string("init: hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
            return CHAMP_ERR_HIP;
        }
        g_scratch_dev[dev] = (double *)p;
    }
    *out = g_scratch_dev[dev];
    return CHAMP_OK;
}

// The Jastrow coefficients of JasParams from the system-array layouts
// (a4 [nctype][norda+1], b [nspin2b][nordb+1], c [nctype][nterms]), with the
// asymptotic values subtracted by jastrow4 (set_scale_dist, scale_dist.f90:3-60).
static void jas_coeffs(JasParams &jp, const double *a4, const double *b, const double *cc)
{
    const double ar = jp.asymp_r;
    const int nt = jp.nterms;
    for (int it = 0; it < jp.nctype; it++) {
        for (int p = 0; p <= jp.norda; p++) jp.a4[it][p] = a4[it * (jp.norda + 1) + p];
        for (int p = 0; p < nt; p++) jp.c[it][p] = cc[it * nt + p];
        double ja = jp.a4[it][0] * ar / (1 + jp.a4[it][1] * ar);
        for (int iord = 2; iord <= jp.norda; iord++) ja = ja + jp.a4[it][iord] * hpowi(ar, iord);
        jp.ja[it] = ja;
    }
    for (int isp = 0; isp < jp.nspin2b; isp++)
        for (int p = 0; p <= jp.nordb; p++) jp.b[isp][p] = b[isp * (jp.nordb + 1) + p];
    for (int i = 0; i < 2; i++) {
        double sspinn; int isp;
        if (i == 0) { sspinn = 1; isp = 0; }
        else { sspinn = (jp.nspin2b == 1 && jp.nocuspb == 0) ? 0.5 : 1.0; isp = jp.nspin2b - 1; }
        double jb = sspinn * jp.b[isp][0] * ar / (1 + jp.b[isp][1] * ar);
        for (int iord = 2; iord <= jp.nordb; iord++) jb = jb + jp.b[isp][iord] * hpowi(ar, iord);
        jp.jb[i] = jb;
    }
}

extern "C" int champ_gpu_init(champ_ctx **ctxp, const champ_system_t *s, int32_t nwalk, int32_t precision)
{
    if (!ctxp || !s || nwalk <= 0) { set_error("init: bad arguments"); return CHAMP_ERR_ARG; }
    *ctxp = nullptr;
    if (precision != CHAMP_PREC_FP64 && precision != CHAMP_PREC_MIXED) { set_error("init: unknown precision"); return CHAMP_ERR_ARG; }
    if (s->nctype > CHAMP_MAXCT || s->ncent > 128 || s->nelec > 1024 || s->nelec < 2 || s->ndet < 1 ||
        s->norda >= CHAMP_NORDP || s->nordb >= CHAMP_NORDP || s->nordc >= CHAMP_NORDP || s->nspin2b < 1 ||
        s->nspin2b > 2 || (s->nloc != 0 && s->nloc != 4) || s->nquad > CHAMP_MPS_QUAD) {
        set_error("init: system outside the supported envelope (nctype<=4, ncent<=128, 2<=nelec<=1024, ndet>=1, orders<8, nloc 0|4)");
        return CHAMP_ERR_UNSUPPORTED;
    }
    for (int e = 0; e < s->nelec; e++)
        if (s->iworbd[e] < 0 || s->iworbd[e] >= s->norb) { set_error("init: iworbd outside 0..norb-1"); return CHAMP_ERR_ARG; }
    for (int ic = 0; ic < s->ncent; ic++)
        if (s->iwctype[ic] < 0 || s->iwctype[ic] >= s->nctype) { set_error("init: iwctype outside 0..nctype-1"); return CHAMP_ERR_ARG; }
    for (int it = 0; it < s->nctype; it++)
        for (int i = 0; i < s->nbastyp[it]; i++) {
            const int ir = s->iwrwf[it * s->nbastyp_max + i], l = s->iwlbas[it * s->nbastyp_max + i];
            if (ir < 0 || ir >= s->nrbas[it] || l < 1 || l > CHAMP_MAX_SLM) { set_error("init: AO table entry out of range"); return CHAMP_ERR_ARG; }
        }
    if (s->nr < 2 || !s->cdet) { set_error("init: radial grid or cdet missing"); return CHAMP_ERR_ARG; }
    if (s->nup < 0 || s->nup > s->nelec) { set_error("init: nup outside 0..nelec"); return CHAMP_ERR_ARG; }
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) { set_error("init: no HIP device"); return CHAMP_ERR_HIP; }
    champ_ctx *c = new champ_ctx();
    *ctxp = c;
    c->dev = dev;
    if (int rc = noalloc_scratch(dev, &c->scratch)) { champ_gpu_free(c); *ctxp = nullptr; return rc; }
    c->W = nwalk; c->NE = s->nelec; c->NUP = s->nup; c->NDN = s->nelec - s->nup;
    c->NM = std::max(c->NUP, c->NDN); c->NC = s->ncent; c->NCT = s->nctype; c->NB = s->nbasis; c->NO = s->norb;
    c->NQ = s->nloc ? s->nquad : 0; c->nloc = s->nloc; c->prec = precision; c->cdet0 = s->cdet[0];
    // mixed path: the FP32-fed Sherman-Morrison inverse is recomputed in FP64
    // every 80 moves for a single determinant (orbital values and one LU); a
    // multideterminant recompute rebuilds the whole excitation state, as much
    // as a quarter of a sweep at 5000 determinants, so there it is left to hpsi
    c->lu_every = (precision == CHAMP_PREC_MIXED && s->ndet == 1) ? 80 : 0;
    if (c->NM > 400) { set_error("init: more than 400 electrons of one spin"); delete c; *ctxp = nullptr; return CHAMP_ERR_UNSUPPORTED; }
    if (hipStreamCreate(&c->stream) != hipSuccess) { set_error("init: stream"); delete c; *ctxp = nullptr; return CHAMP_ERR_HIP; }

    const int NE = c->NE, NC = c->NC, NB = c->NB, NO = c->NO, W = c->W, NM = c->NM, NCT = c->NCT;
    // ---- pecent (pot_nn, pot.f90:3-50)
    double pecent = 0;
    for (int i = 1; i < NC; i++)
        for (int j = 0; j < i; j++) {
            double r2 = 0;
            for (int k = 0; k < 3; k++) { double d = s->cent[3 * i + k] - s->cent[3 * j + k]; r2 = r2 + d * d; }
            pecent = pecent + s->znuc[s->iwctype[i]] * s->znuc[s->iwctype[j]] / std::sqrt(r2);
        }
    c->pecent = pecent;

    TRY(upload(&c->cent, s->cent, (size_t)3 * NC));
    TRY(upload(&c->iwctype, s->iwctype, (size_t)NC));
    // ---- spline tables
    const size_t ntab = (size_t)NCT * s->nrbas_max;
    double *d_rwf, *d_d2, *d_ce, *d_ae, *d_rmax;
    TRY(upload(&d_rwf, s->rwf, ntab * s->nr));
    TRY(upload(&d_d2, s->d2rwf, ntab * s->nr));
    TRY(upload(&d_ce, s->ce, ntab * CHAMP_NCOEF));
    TRY(upload(&d_ae, s->ae, ntab * 2));
    TRY(upload(&d_rmax, s->rmaxwf, ntab));
    c->tab.rwf = d_rwf; c->tab.d2rwf = d_d2; c->tab.ce = d_ce; c->tab.ae = d_ae; c->tab.rmaxwf = d_rmax;
    c->tab.nr = s->nr; c->tab.nrbas_max = s->nrbas_max; c->tab.r0 = s->grid_r0; c->tab.arg = s->grid_arg;
    c->tab.dlogag = std::log(s->grid_arg);
    TRY(dalloc(&c->gref, (size_t)2 * s->nr));
    hipLaunchKernelGGL(k_grid_refs, dim3((s->nr + 255) / 256), dim3(256), 0, c->stream, s->nr, s->grid_r0, s->grid_arg,
                       c->gref);
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipStreamSynchronize(c->stream));
    c->tab.gref = c->gref;
    // ---- AO table (basis_fns ordering: centres in order, AOs of the centre's type)
    std::vector<int> aoc, aot, aor, aol;
    std::vector<double> aorm;
    for (int ic = 0; ic < NC; ic++) {
        const int it = s->iwctype[ic];
        for (int i = 0; i < s->nbastyp[it]; i++) {
            const int ir = s->iwrwf[it * s->nbastyp_max + i];
            aoc.push_back(ic); aot.push_back(it); aor.push_back(ir);
            aol.push_back(s->iwlbas[it * s->nbastyp_max + i]);
            aorm.push_back(s->rmaxwf[it * s->nrbas_max + ir]);
        }
    }
    if ((int)aoc.size() != NB) { set_error("init: nbasis does not match the per-type AO counts"); champ_gpu_free(c); *ctxp = nullptr; return CHAMP_ERR_ARG; }
    TRY(upload(&c->ao_cent, aoc.data(), NB)); TRY(upload(&c->ao_type, aot.data(), NB));
    TRY(upload(&c->ao_rad, aor.data(), NB)); TRY(upload(&c->ao_slm, aol.data(), NB));
    TRY(upload(&c->rmax_ao, aorm.data(), NB));
    // radial instances (centre, radial function) and each AO's instance
    {
        std::vector<int> ric, rit, rir, rfirst(NC);
        std::vector<double> rirm;
        for (int ic = 0; ic < NC; ic++) {
            const int it = s->iwctype[ic];
            rfirst[ic] = (int)ric.size();
            for (int ir = 0; ir < s->nrbas[it]; ir++) {
                ric.push_back(ic); rit.push_back(it); rir.push_back(ir);
                rirm.push_back(s->rmaxwf[it * s->nrbas_max + ir]);
            }
        }
        std::vector<int> aori(NB);
        for (int m = 0; m < NB; m++) aori[m] = rfirst[aoc[m]] + aor[m];
        c->NRI = (int)ric.size();
        TRY(upload(&c->ri_cent, ric.data(), ric.size())); TRY(upload(&c->ri_type, rit.data(), rit.size()));
        TRY(upload(&c->ri_rad, rir.data(), rir.size())); TRY(upload(&c->ri_rmax, rirm.data(), rirm.size()));
        TRY(upload(&c->ao_ri, aori.data(), NB));
        // each instance's AOs, and the instances in decreasing order of their AO count
        std::vector<int> aoff(c->NRI + 1, 0), aol(NB), perm(c->NRI);
        for (int m = 0; m < NB; m++) aoff[aori[m] + 1]++;
        for (int r = 0; r < c->NRI; r++) aoff[r + 1] += aoff[r];
        std::vector<int> fill(aoff.begin(), aoff.end() - 1);
        for (int m = 0; m < NB; m++) aol[fill[aori[m]]++] = m;
        for (int r = 0; r < c->NRI; r++) perm[r] = r;
        std::stable_sort(perm.begin(), perm.end(),
                         [&](int x, int y) { return aoff[x + 1] - aoff[x] > aoff[y + 1] - aoff[y]; });
        TRY(upload(&c->ri_perm, perm.data(), c->NRI)); TRY(upload(&c->ri_aoff, aoff.data(), c->NRI + 1));
        TRY(upload(&c->ri_aol, aol.data(), NB));
    }
    TRY(upload(&c->coef, s->coef, (size_t)NO * NB));
    if (precision == CHAMP_PREC_MIXED) {
        std::vector<float> cf((size_t)NO * NB);
        for (size_t t = 0; t < cf.size(); t++) cf[t] = (float)s->coef[t];
        TRY(upload(&c->coef32, cf.data(), cf.size()));
    }
    TRY(upload(&c->iworbd, s->iworbd, (size_t)NE));
    for (int iab = 0; iab < 2; iab++) {
        const int nel = iab ? c->NDN : c->NUP, ish = iab ? c->NUP : 0;
        std::vector<double> h((size_t)std::max(nel, 1) * NB, 0.0);
        for (int j = 0; j < nel; j++) std::memcpy(&h[(size_t)j * NB], s->coef + (size_t)s->iworbd[ish + j] * NB, sizeof(double) * NB);
        TRY(upload(&c->cocc[iab], h.data(), h.size()));
    }
    // ---- Jastrow parameters + asymptotes (set_scale_dist, scale_dist.f90:3-60)
    JasParams &jp = c->hjas;
    std::memset(&jp, 0, sizeof(jp));
    jp.norda = s->norda; jp.nordb = s->nordb; jp.nordc = s->nordc; jp.nspin2b = s->nspin2b; jp.nocuspb = s->nocuspb;
    jp.nctype = NCT; jp.kap = s->scalek; jp.asymp_r = 1.0 / s->scalek;
    int nt = 0;
    for (int n = 2; n <= s->nordc; n++)
        for (int k = n - 1; k >= 0; k--) {
            const int l_hi = (k == 0) ? n - k - 2 : n - k;
            for (int l = l_hi; l >= 0; l--) {
                const int m = (n - k - l) / 2;
                if (2 * m == n - k - l) {
                    if (nt >= CHAMP_MAXT) { set_error("init: too many e-e-n terms"); champ_gpu_free(c); *ctxp = nullptr; return CHAMP_ERR_UNSUPPORTED; }
                    jp.tk[nt] = (signed char)k; jp.tl[nt] = (signed char)l; jp.tm[nt] = (signed char)m; nt++;
                }
            }
        }
    jp.nterms = nt;
    jas_coeffs(jp, s->a4, s->b, s->c);
    TRY(upload(&c->jas, &c->hjas, 1));
    // ---- ECP
    EcpParams &ep = c->hecp;
    std::memset(&ep, 0, sizeof(ep));
    ep.nloc = s->nloc; ep.nquad = c->NQ;
    int nonloc_cent = 0;
    for (int it = 0; it < NCT; it++) {
        ep.znuc[it] = s->znuc[it];
        ep.lpot[it] = s->nloc ? s->lpot[it] : 0;
        if (s->nloc)
            for (int l = 0; l < CHAMP_MPS_L; l++) {
                ep.nterm[it][l] = s->necp_term[it * CHAMP_MPS_L + l];
                for (int g = 0; g < CHAMP_MGAUSS; g++) {
                    const size_t q = ((size_t)it * CHAMP_MPS_L + l) * CHAMP_MGAUSS + g;
                    ep.power[it][l][g] = s->necp_power[q];
                    ep.coef[it][l][g] = s->ecp_coef[q];
                    ep.expo[it][l][g] = s->ecp_exponent[q];
                }
            }
    }
    for (int q = 0; q < c->NQ; q++) ep.wq[q] = s->wq[q];
    for (int ic = 0; ic < NC; ic++) if (s->nloc && s->lpot[s->iwctype[ic]] > 1) nonloc_cent++;
    TRY(upload(&c->ecp, &c->hecp, 1));
    const int nq = std::max(c->NQ, 1);
    c->h_xq0.assign(nq, 0.0); c->h_yq0.assign(nq, 0.0); c->h_zq0.assign(nq, 0.0);
    for (int q = 0; q < c->NQ; q++) { c->h_xq0[q] = s->xq0[q]; c->h_yq0[q] = s->yq0[q]; c->h_zq0[q] = s->zq0[q]; }
    TRY(upload(&c->xq0, c->h_xq0.data(), nq)); TRY(upload(&c->yq0, c->h_yq0.data(), nq)); TRY(upload(&c->zq0, c->h_zq0.data(), nq));

    // ---- walker state
    TRY(dalloc(&c->xT, (size_t)W * 3 * NE));
    TRY(dalloc(&c->r_en, (size_t)W * NC * NE)); TRY(dalloc(&c->rr_en, (size_t)W * NC * NE));
    if (precision == CHAMP_PREC_MIXED) TRY(dalloc(&c->enp, (size_t)W * NC * NE));
    TRY(dalloc(&c->fso, (size_t)W * NE * NE));
    TRY(dalloc(&c->g1, (size_t)W * NE * 3 * NE)); TRY(dalloc(&c->g2, (size_t)W * NE * 3 * NE));
    TRY(dalloc(&c->vjT, (size_t)W * 3 * NE));
    TRY(dalloc(&c->fsumo, W)); TRY(dalloc(&c->d2o, W));
    TRY(dalloc(&c->dorb, (size_t)W * NE * 3 * NO));
    TRY(dalloc(&c->slmi, (size_t)W * 2 * NM * NM));
    TRY(dalloc(&c->detiab, (size_t)W * 2));
    TRY(dalloc(&c->ddx, (size_t)W * NE * 3)); TRY(dalloc(&c->d2dx2, (size_t)W * NE));
    TRY(dalloc(&c->psido, W)); TRY(dalloc(&c->psijo, W)); TRY(dalloc(&c->psi2o, W)); TRY(dalloc(&c->eold, W));
    TRY(dalloc(&c->ekin, W)); TRY(dalloc(&c->pe_local, W)); TRY(dalloc(&c->vpsp, (size_t)2 * W));
    TRY(dalloc(&c->status, W)); TRY(dalloc(&c->rng, (size_t)4 * W));
    TRY(dalloc(&c->xq, (size_t)W * nq)); TRY(dalloc(&c->yq, (size_t)W * nq)); TRY(dalloc(&c->zq, (size_t)W * nq));
    // move scratch
    TRY(dalloc(&c->xnew, (size_t)3 * W)); TRY(dalloc(&c->dfus2o, W));
    TRY(dalloc(&c->rnew, (size_t)W * NC)); TRY(dalloc(&c->rrnew, (size_t)W * NC));
    TRY(dalloc(&c->fsn_row, (size_t)W * NE));
    TRY(dalloc(&c->g1n, (size_t)W * 3 * NE)); TRY(dalloc(&c->g2n, (size_t)W * 3 * NE)); TRY(dalloc(&c->vjnT, (size_t)W * 3 * NE));
    TRY(dalloc(&c->fsumn, W));
    TRY(dalloc(&c->aomove, (size_t)W * 4 * NB)); TRY(dalloc(&c->orbn, (size_t)W * 4 * NO));
    TRY(dalloc(&c->ratio, W)); TRY(dalloc(&c->psidn, W)); TRY(dalloc(&c->psijn, W));
    TRY(dalloc(&c->vnew, (size_t)3 * W)); TRY(dalloc(&c->vold, (size_t)3 * W));
    TRY(dalloc(&c->acc, W)); TRY(dalloc(&c->esum_dev, 4));
    TRY(dalloc(&c->m6rec, (size_t)W * M6_REC));
    // hpsi scratch
    const double per_w = (double)NE * 5 * NB * 8;
    c->hpsi_chunk = (int)std::max(1.0, std::min((double)W, 3.0e9 / per_w));
    TRY(dalloc(&c->aoall, (size_t)c->hpsi_chunk * NE * 5 * NB));
    TRY(dalloc(&c->orball, (size_t)c->hpsi_chunk * NE * 5 * NO));
    c->pair_cap = std::max(1, NE * nonloc_cent);
    if (c->nloc) {
        TRY(dalloc(&c->wmat, (size_t)W * NE * NB));
        TRY(dalloc(&c->pair_e, (size_t)W * c->pair_cap)); TRY(dalloc(&c->pair_c, (size_t)W * c->pair_cap));
        TRY(dalloc(&c->pair_vps, (size_t)W * c->pair_cap * CHAMP_MPS_L));
        TRY(dalloc(&c->pair_val, (size_t)W * c->pair_cap));
        TRY(dalloc(&c->pair_psi, (size_t)W * c->pair_cap * std::max(c->NQ, 1)));
        TRY(dalloc(&c->rrT, (size_t)W * NE * NC));
    }
    TRY(dalloc(&c->pair_count, W)); TRY(dalloc(&c->pair_off, (size_t)W + 1)); TRY(dalloc(&c->pair_total, 2));
    TRY(dalloc(&c->grp_off, (size_t)W + 1));
    TRY(dalloc(&c->nxquad, W));
    HIPCHECK(hipMemset(c->status, 0, sizeof(int) * W));
    HIPCHECK(hipMemset(c->nxquad, 0, sizeof(int) * W));
    HIPCHECK(hipMemset(c->pair_count, 0, sizeof(int) * W));
    HIPCHECK(hipMemset(c->esum_dev, 0, sizeof(double) * 4));
    // quadrature starts unrotated
    {
        std::vector<double> hx((size_t)W * nq), hy((size_t)W * nq), hz((size_t)W * nq);
        for (int w = 0; w < W; w++)
            for (int q = 0; q < nq; q++) { hx[(size_t)w * nq + q] = c->h_xq0[q]; hy[(size_t)w * nq + q] = c->h_yq0[q]; hz[(size_t)w * nq + q] = c->h_zq0[q]; }
        HIPCHECK(hipMemcpy(c->xq, hx.data(), hx.size() * 8, hipMemcpyHostToDevice));
        HIPCHECK(hipMemcpy(c->yq, hy.data(), hy.size() * 8, hipMemcpyHostToDevice));
        HIPCHECK(hipMemcpy(c->zq, hz.data(), hz.size() * 8, hipMemcpyHostToDevice));
    }
    // the Slater inverse works in LDS up to ~140 electrons of one spin, above
    // that in a global work matrix per (walker, spin) of an hpsi chunk
    // (the kernel's own pivot arrays take SLATER_STATIC_LDS bytes beside the
    // dynamic work matrix; both count against one block's LDS)
    {
        const size_t lds = (size_t)(NM * (NM + 1) + NM) * sizeof(double) + SLATER_STATIC_LDS;
        if (lds > CHAMP_SLATER_LDS_MAX) TRY(dalloc(&c->lu_ws, (size_t)2 * c->hpsi_chunk * NM * (NM + 1)));
    }
    // excitation tables and state of a multideterminant trial function
    TRY(md_init(c, s));
    // walker streams from CHAMP's default seed "18123437465549275475255231234865"
    // read as 8 x i4 (parser.f90:315,654), until seed_rng / set_rng replace them
    {
        const int32_t seed0[8] = {1812, 3437, 4655, 4927, 5475, 2552, 3123, 4865};
        TRY(champ_gpu_seed_rng(c, seed0));
    }
    return CHAMP_OK;
}

extern "C" int champ_gpu_free(champ_ctx *c)
{
    if (!c) return CHAMP_OK;
    if (c->stream) hipStreamSynchronize(c->stream);
    void *ptrs[] = {c->cent, c->iwctype, (void *)c->tab.rwf, (void *)c->tab.d2rwf, (void *)c->tab.ce, (void *)c->tab.ae,
                    (void *)c->tab.rmaxwf, c->rmax_ao, c->ao_cent, c->ao_type, c->ao_rad, c->ao_slm, c->coef,
                    c->cocc[0], c->cocc[1], c->coef32, c->iworbd, c->jas, c->ecp, c->xq0, c->yq0, c->zq0, c->xT, c->r_en,
                    c->rr_en, c->fso, c->g1, c->g2, c->vjT, c->fsumo, c->d2o, c->dorb, c->slmi, c->detiab, c->ddx,
                    c->d2dx2, c->psido, c->psijo, c->psi2o, c->eold, c->ekin, c->pe_local, c->vpsp, c->status,
                    c->rng, c->xq, c->yq, c->zq, c->xnew, c->dfus2o, c->rnew, c->rrnew, c->fsn_row, c->g1n, c->g2n,
                    c->vjnT, c->fsumn, c->aomove, c->orbn, c->ratio, c->psidn, c->psijn, c->vnew, c->vold, c->acc,
                    c->esum_dev, c->aoall, c->orball, c->wmat, c->pair_count, c->pair_off, c->pair_total, c->grp_off,
                    c->pair_e, c->pair_c, c->pair_vps, c->pair_val, c->pair_psi, c->rrT, c->nxquad, c->ri_cent, c->ri_type,
                    c->ri_rad, c->ri_rmax, c->ao_ri, c->gref, c->lu_ws, c->m6rec, c->ri_perm,
                    c->ri_aoff, c->ri_aol, c->enp, c->gemm_tmp};
    for (void *p : ptrs) if (p) hipFree(p);
    md_free(c);
    dmc_free(c);
    deriv_free(c);
    if (c->sweep_exec) hipGraphExecDestroy(c->sweep_exec);
    for (auto &t : c->ev_pool) { hipEventDestroy(t.a); hipEventDestroy(t.b); }
    if (c->stream) hipStreamDestroy(c->stream);
    delete c;
    return CHAMP_OK;
}

extern "C" int champ_gpu_set_walkers(champ_ctx *c, const double *x)
{
    if (!c || !x) return CHAMP_ERR_ARG;
    std::vector<double> h((size_t)c->W * 3 * c->NE);
    for (int w = 0; w < c->W; w++)
        for (int e = 0; e < c->NE; e++)
            for (int k = 0; k < 3; k++) h[((size_t)w * 3 + k) * c->NE + e] = x[((size_t)w * c->NE + e) * 3 + k];
    HIPCHECK(hipMemcpyAsync(c->xT, h.data(), h.size() * 8, hipMemcpyHostToDevice, c->stream));
    HIPCHECK(hipStreamSynchronize(c->stream));
    return CHAMP_OK;
}

extern "C" int champ_gpu_set_params(champ_ctx *c, const double *a4, const double *b, const double *cc,
                                    const double *cdet)
{
    if (!c) return CHAMP_ERR_ARG;
    JasParams jp = c->hjas;
    // check_parms_jas (optwf_handle_wf.f90:1286-1336): the Pade denominators
    // a4[1] and b[1] must stay above -scalek
    const double scalem = -jp.kap;
    if (a4)
        for (int it = 0; it < jp.nctype; it++)
            if (jp.norda >= 1 && a4[it * (jp.norda + 1) + 1] <= scalem) {
                set_error("set_params: a4[1] <= -scalek");
                return CHAMP_ERR_ARG;
            }
    if (b)
        for (int isp = 0; isp < jp.nspin2b; isp++)
            if (jp.nordb >= 1 && b[isp * (jp.nordb + 1) + 1] <= scalem) {
                set_error("set_params: b[1] <= -scalek");
                return CHAMP_ERR_ARG;
            }
    if (a4 || b || cc) {
        std::vector<double> ha((size_t)jp.nctype * (jp.norda + 1)), hb((size_t)jp.nspin2b * (jp.nordb + 1)),
            hc((size_t)jp.nctype * std::max(jp.nterms, 1));
        for (int it = 0; it < jp.nctype; it++) {
            for (int p = 0; p <= jp.norda; p++) ha[it * (jp.norda + 1) + p] = a4 ? a4[it * (jp.norda + 1) + p] : jp.a4[it][p];
            for (int t = 0; t < jp.nterms; t++) hc[it * jp.nterms + t] = cc ? cc[it * jp.nterms + t] : jp.c[it][t];
        }
        for (int isp = 0; isp < jp.nspin2b; isp++)
            for (int p = 0; p <= jp.nordb; p++) hb[isp * (jp.nordb + 1) + p] = b ? b[isp * (jp.nordb + 1) + p] : jp.b[isp][p];
        jas_coeffs(jp, ha.data(), hb.data(), hc.data());
        HIPCHECK(hipStreamSynchronize(c->stream));
        c->hjas = jp;
        HIPCHECK(hipMemcpy(c->jas, &c->hjas, sizeof(JasParams), hipMemcpyHostToDevice));
    }
    if (cdet) {
        if (c->md) {
            if (int rc = md_set_cdet(c, cdet)) return rc;
        } else {
            c->cdet0 = cdet[0];
        }
    }
    // kernel arguments captured in a sweep graph hold the old coefficients
    if (c->sweep_exec) { hipGraphExecDestroy(c->sweep_exec); c->sweep_exec = nullptr; }
    int rc = launch_hpsi_impl(c, 0);
    if (!rc) HIPCHECK(hipStreamSynchronize(c->stream));
    return rc;
}

extern "C" int champ_gpu_set_orbitals(champ_ctx *c, const double *coef)
{
    if (!c || !coef) return CHAMP_ERR_ARG;
    const int NO = c->NO, NB = c->NB, NE = c->NE;
    HIPCHECK(hipStreamSynchronize(c->stream));
    HIPCHECK(hipMemcpy(c->coef, coef, (size_t)NO * NB * sizeof(double), hipMemcpyHostToDevice));
    if (c->coef32) {
        std::vector<float> cf((size_t)NO * NB);
        for (size_t t = 0; t < cf.size(); t++) cf[t] = (float)coef[t];
        HIPCHECK(hipMemcpy(c->coef32, cf.data(), cf.size() * sizeof(float), hipMemcpyHostToDevice));
    }
    // the occupied rows of the reference determinant per spin (the ECP rows' GEMM)
    std::vector<int> iw(NE);
    HIPCHECK(hipMemcpy(iw.data(), c->iworbd, (size_t)NE * sizeof(int), hipMemcpyDeviceToHost));
    for (int iab = 0; iab < 2; iab++) {
        const int nel = iab ? c->NDN : c->NUP, ish = iab ? c->NUP : 0;
        std::vector<double> h((size_t)std::max(nel, 1) * NB, 0.0);
        for (int j = 0; j < nel; j++) std::memcpy(&h[(size_t)j * NB], coef + (size_t)iw[ish + j] * NB, sizeof(double) * NB);
        HIPCHECK(hipMemcpy(c->cocc[iab], h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice));
    }
    if (c->sweep_exec) { hipGraphExecDestroy(c->sweep_exec); c->sweep_exec = nullptr; }
    int rc = launch_hpsi_impl(c, 0);
    if (!rc) HIPCHECK(hipStreamSynchronize(c->stream));
    return rc;
}

extern "C" int champ_gpu_get_walkers(champ_ctx *c, double *x)
{
    if (!c || !x) return CHAMP_ERR_ARG;
    std::vector<double> h((size_t)c->W * 3 * c->NE);
    HIPCHECK(hipMemcpyAsync(h.data(), c->xT, h.size() * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHECK(hipStreamSynchronize(c->stream));
    for (int w = 0; w < c->W; w++)
        for (int e = 0; e < c->NE; e++)
            for (int k = 0; k < 3; k++) x[((size_t)w * c->NE + e) * 3 + k] = h[((size_t)w * 3 + k) * c->NE + e];
    return CHAMP_OK;
}

extern "C" int champ_gpu_set_rng(champ_ctx *c, const uint64_t *st)
{
    if (!c || !st) return CHAMP_ERR_ARG;
    // an all-zero xoshiro256** state stays zero forever (u = 0 on every draw)
    for (int w = 0; w < c->W; w++)
        if ((st[4 * (size_t)w] | st[4 * (size_t)w + 1] | st[4 * (size_t)w + 2] | st[4 * (size_t)w + 3]) == 0) {
            set_error("set_rng: walker " + std::to_string(w) + " has an all-zero xoshiro256** state");
            return CHAMP_ERR_ARG;
        }
    HIPCHECK(hipMemcpy(c->rng, st, sizeof(uint64_t) * 4 * c->W, hipMemcpyHostToDevice));
    return CHAMP_OK;
}
extern "C" int champ_gpu_get_rng(champ_ctx *c, uint64_t *st)
{
    if (!c || !st) return CHAMP_ERR_ARG;
    HIPCHECK(hipStreamSynchronize(c->stream));
    HIPCHECK(hipMemcpy(st, c->rng, sizeof(uint64_t) * 4 * c->W, hipMemcpyDeviceToHost));
    return CHAMP_OK;
}
extern "C" int champ_gpu_seed_rng(champ_ctx *c, const int32_t seed[8])
{
    if (!c || !seed) return CHAMP_ERR_ARG;
    uint64_t s[4];
    for (int i = 0; i < 4; i++) s[i] = (uint64_t)(uint32_t)seed[2 * i] | ((uint64_t)(uint32_t)seed[2 * i + 1] << 32);
    std::vector<uint64_t> h((size_t)4 * c->W);
    for (int w = 0; w < c->W; w++) {
        for (int k = 0; k < 4; k++) h[4 * (size_t)w + k] = s[k];
        hjump(s);
    }
    return champ_gpu_set_rng(c, h.data());
}

extern "C" int champ_gpu_set_quadrature(champ_ctx *c, const double *xq, const double *yq, const double *zq)
{
    if (!c || !xq || !yq || !zq) return CHAMP_ERR_ARG;
    if (c->NQ == 0) return CHAMP_OK;
    const size_t n = (size_t)c->W * c->NQ * 8;
    HIPCHECK(hipMemcpy(c->xq, xq, n, hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(c->yq, yq, n, hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(c->zq, zq, n, hipMemcpyHostToDevice));
    return CHAMP_OK;
}
extern "C" int champ_gpu_rotqua(champ_ctx *c)
{
    if (!c) return CHAMP_ERR_ARG;
    if (c->NQ == 0) return CHAMP_OK;
    int rc = launch_rotqua(c);
    if (rc) return rc;
    HIPCHECK(hipStreamSynchronize(c->stream));
    return CHAMP_OK;
}

static int copy_out(champ_ctx *c, double *h, const double *d, size_t n)
{
    if (!h) return CHAMP_OK;
    HIPCHECK(hipMemcpyAsync(h, d, n * 8, hipMemcpyDeviceToHost, c->stream));
    return CHAMP_OK;
}

extern "C" int champ_gpu_hpsi(champ_ctx *c, double *psid, double *psij, double *ekin, double *energy)
{
    if (!c) return CHAMP_ERR_ARG;
    int rc = launch_hpsi_impl(c, 0);
    if (rc) return rc;
    if ((rc = copy_out(c, psid, c->psido, c->W))) return rc;
    if ((rc = copy_out(c, psij, c->psijo, c->W))) return rc;
    if ((rc = copy_out(c, ekin, c->ekin, c->W))) return rc;
    if ((rc = copy_out(c, energy, c->eold, c->W))) return rc;
    HIPCHECK(hipStreamSynchronize(c->stream));
    return CHAMP_OK;
}

extern "C" int champ_gpu_determinant(champ_ctx *c, double *psid)
{
    if (!c) return CHAMP_ERR_ARG;
    int rc = launch_slater_refresh(c);
    if (rc) return rc;
    if ((rc = copy_out(c, psid, c->psido, c->W))) return rc;
    HIPCHECK(hipStreamSynchronize(c->stream));
    return CHAMP_OK;
}

extern "C" int champ_gpu_compute_determinante_grad(champ_ctx *c, int32_t iel, int32_t iflag_move, double *vd)
{
    if (!c || !vd || iel < 0 || iel >= c->NE) return CHAMP_ERR_ARG;
    if (iflag_move == 1) {
        int rc = launch_md_pre(c, iel);
        if (rc) return rc;
        rc = launch_drift_old(c, iel);
        if (rc) return rc;
        HIPCHECK(hipMemcpyAsync(vd, c->vold, (size_t)3 * c->W * 8, hipMemcpyDeviceToHost, c->stream));
    } else if (iflag_move == 0) {
        HIPCHECK(hipMemcpyAsync(vd, c->vnew, (size_t)3 * c->W * 8, hipMemcpyDeviceToHost, c->stream));
    } else {
        return CHAMP_ERR_UNSUPPORTED;
    }
    HIPCHECK(hipStreamSynchronize(c->stream));
    return CHAMP_OK;
}

// multideterminant steps of a move, timed as one kernel class
static int launch_md_pre_timed(champ_ctx *c, int iel)
{
    if (!c->md) return CHAMP_OK;
    int slot;
    tbegin(c, K_MD, &slot);
    const int rc = launch_md_pre(c, iel);
    tend(c, slot);
    return rc;
}
static int launch_md_post_timed(champ_ctx *c, int iel)
{
    if (!c->md) return CHAMP_OK;
    int slot;
    tbegin(c, K_MD, &slot);
    const int rc = launch_md_post(c, iel);
    tend(c, slot);
    return rc;
}

extern "C" int champ_gpu_psie(champ_ctx *c, int32_t iel, const double *xnew, double *psidn, double *psijn, int32_t iflag)
{
    if (!c || !xnew || iel < 0 || iel >= c->NE) return CHAMP_ERR_ARG;
    if (iflag != 0) { set_error("psie: only iflag=0 (metrop1) is provided"); return CHAMP_ERR_UNSUPPORTED; }
    HIPCHECK(hipMemcpyAsync(c->xnew, xnew, (size_t)3 * c->W * 8, hipMemcpyHostToDevice, c->stream));
    int rc;
    if ((rc = launch_move_a(c, iel, 1))) return rc;
    if ((rc = launch_move_gemm(c))) return rc;
    if ((rc = launch_md_post_timed(c, iel))) return rc;
    if ((rc = launch_move_b(c, iel, 1))) return rc;
    if ((rc = copy_out(c, psidn, c->psidn, c->W))) return rc;
    if ((rc = copy_out(c, psijn, c->psijn, c->W))) return rc;
    HIPCHECK(hipStreamSynchronize(c->stream));
    return CHAMP_OK;
}

extern "C" int champ_gpu_accept(champ_ctx *c, int32_t iel, const uint8_t *accept)
{
    if (!c || !accept || iel < 0 || iel >= c->NE) return CHAMP_ERR_ARG;
    HIPCHECK(hipMemcpyAsync(c->acc, accept, c->W, hipMemcpyHostToDevice, c->stream));
    int rc = launch_move_b(c, iel, 2);
    if (rc) return rc;
    HIPCHECK(hipStreamSynchronize(c->stream));
    return CHAMP_OK;
}

// One single-electron move of every walker (drift, proposal, ratio, decision,
// commit) and the periodic refresh of the Slater inverses.
int launch_move_step(champ_ctx *c, int iel)
{
    int rc;
    if ((rc = launch_md_pre_timed(c, iel))) return rc;
    if ((rc = launch_move_a(c, iel, 0))) return rc;
    if ((rc = launch_move_gemm(c))) return rc;
    if ((rc = launch_md_post_timed(c, iel))) return rc;
    if ((rc = launch_move_b(c, iel, 0))) return rc;
    if (c->lu_every > 0 && (iel + 1) % c->lu_every == 0 && iel + 1 < c->NE && (rc = launch_slater_refresh(c)))
        return rc;
    return CHAMP_OK;
}

// Electron iel of every walker moved to xnew and committed where acc[w] is
// set (psie + jassav/detsav; DMC's T-moves), the others left as they are.
int launch_forced_move(champ_ctx *c, int iel)
{
    int rc;
    if ((rc = launch_move_a(c, iel, 1))) return rc;
    if ((rc = launch_move_gemm(c))) return rc;
    if ((rc = launch_md_post_timed(c, iel))) return rc;
    return launch_move_b(c, iel, 2);
}

// One metrop1 sweep: rotqua, Ne single-electron moves, hpsi accumulating the
// local energy.  Every launch argument depends only on the context and tau.
static int launch_sweep(champ_ctx *c)
{
    int rc;
    if (c->nloc > 0 && (rc = launch_rotqua(c))) return rc;
    for (int iel = 0; iel < c->NE; iel++)
        if ((rc = launch_move_step(c, iel))) return rc;
    return launch_hpsi_impl(c, 1);
}

// Estimator sums of the sweeps since the last clear, and the flagged walkers.
static int sweep_stats(champ_ctx *c, int32_t nsweep, champ_sweep_stats *stats)
{
    double h[4];
    HIPCHECK(hipMemcpyAsync(h, c->esum_dev, sizeof(h), hipMemcpyDeviceToHost, c->stream));
    HIPCHECK(hipStreamSynchronize(c->stream));
    tflush(c);
    if (stats) {
        std::vector<int> st(c->W);
        HIPCHECK(hipMemcpy(st.data(), c->status, sizeof(int) * c->W, hipMemcpyDeviceToHost));
        int64_t nf = 0;
        for (int v : st) nf += v != 0;
        stats->accepted = h[0];
        stats->proposed = (double)nsweep * c->W * c->NE;
        stats->esum = h[1];
        stats->esum2 = h[2];
        stats->nflagged = nf;
    }
    return CHAMP_OK;
}

// The sweep captured once per (tau, mover, metrop6 parameters) as a HIP graph
// and replayed; it holds kernel nodes only (the estimator sums are cleared
// before the first replay).
static int sweep_graph(champ_ctx *c, double tau)
{
    const double key[5] = {tau, (double)c->imetro, c->m6_deltar, c->m6_deltat, c->m6_fbias};
    if (c->sweep_exec && std::equal(key, key + 5, c->sweep_key)) return CHAMP_OK;
    if (c->sweep_exec) { HIPCHECK(hipGraphExecDestroy(c->sweep_exec)); c->sweep_exec = nullptr; }
    HIPCHECK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeRelaxed));
    const int rc = launch_sweep(c);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(c->stream, &g);
    if (rc) { if (g) hipGraphDestroy(g); return rc; }
    if (e != hipSuccess) { set_error(std::string("sweep capture: ") + hipGetErrorString(e)); return CHAMP_ERR_HIP; }
    const hipError_t ei = hipGraphInstantiate(&c->sweep_exec, g, nullptr, nullptr, 0);
    hipGraphDestroy(g);
    if (ei != hipSuccess) { c->sweep_exec = nullptr; set_error(std::string("sweep graph: ") + hipGetErrorString(ei)); return CHAMP_ERR_HIP; }
    std::copy(key, key + 5, c->sweep_key);
    return CHAMP_OK;
}

// ------------------------------------------------------------------------
// Autotuner: each candidate kernel form of a tunable launch is timed on the
// context's own walkers (HIP events around repeated launches on its stream)
// and the fastest is kept.  Every candidate computes the same quantities.
static int time_launch(champ_ctx *c, int (*fn)(champ_ctx *), int reps, float *ms)
{
    hipEvent_t e0, e1;
    HIPCHECK(hipEventCreate(&e0));
    HIPCHECK(hipEventCreate(&e1));
    int rc = fn(c);  // warm-up (code load, caches)
    if (!rc) {
        hipEventRecord(e0, c->stream);
        for (int r = 0; r < reps && !rc; r++) rc = fn(c);
        hipEventRecord(e1, c->stream);
        hipEventSynchronize(e1);
        hipEventElapsedTime(ms, e0, e1);
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    return rc;
}
// the per-move contraction timed in its setting: a compute-only move of
// electron 0 to its own position (move kernel A, the contraction, the
// multideterminant column, the ratio); the walkers' state is left unchanged
static int tune_move_gemm(champ_ctx *c)
{
    int rc = launch_move_a(c, 0, 1);
    if (!rc) rc = launch_move_gemm(c);
    if (!rc && c->md) rc = launch_md_post(c, 0);
    if (!rc) rc = launch_move_b(c, 0, 1);
    return rc;
}
static int tune_orb_gemm(champ_ctx *c)
{
    const int nw = std::min(c->hpsi_chunk, c->W);
    return launch_gemm(c, true, nw * c->NE * 5, c->NO, c->NB, c->aoall, c->NB, 0, c->coef, c->NB, 0, c->orball, c->NO,
                       0, 1, c->tune_gemm[1]);
}
static int tune_wmat(champ_ctx *c) { return launch_wmat(c); }
static int tune_nonloc(champ_ctx *c)
{
    HArgs a = hargs(c);
    return launch_nonloc(c, a);
}

extern "C" int champ_gpu_autotune(champ_ctx *c, int32_t *choices)
{
    if (!c) return CHAMP_ERR_ARG;
    const bool timing = c->timing;
    c->timing = false;
    int rc = launch_hpsi_impl(c, 0);  // every input of the tuned launches in place
    if (!rc) rc = launch_stay(c, 0);
    int (*site[3])(champ_ctx *) = {tune_move_gemm, tune_orb_gemm, tune_wmat};
    // the ECP coefficient rows (wmat) exist only with a non-local pseudopotential
    const int nsite = c->nloc > 0 ? 3 : 2;
    for (int s = 0; s < nsite && !rc; s++) {
        float best = 1e30f;
        int pick = GEMM_AUTO;
        // the per-move contraction may also run split over K, and on the mixed
        // path on FP32 operands
        int kinds[5] = {GEMM_TILE, GEMM_STRIP, GEMM_VALU}, nk = 3;
        if (s == 0) {
            if ((rc = gemm_reserve_splitk(c, 4 * c->W, c->NO))) break;
            kinds[nk++] = GEMM_SPLITK;
            if (c->prec == CHAMP_PREC_MIXED) kinds[nk++] = GEMM_F32;
        }
        for (int ik = 0; ik < nk; ik++) {
            c->tune_gemm[s] = kinds[ik];
            float ms = 0;
            if ((rc = time_launch(c, site[s], 10, &ms))) break;
            if (ms < best) { best = ms; pick = kinds[ik]; }
        }
        c->tune_gemm[s] = pick;
    }
    if (!rc && c->nloc > 0) {
        float best = 1e30f;
        int pick = 0;
        for (int kind : {0, 1}) {
            c->tune_nonloc = kind;
            float ms = 0;
            if ((rc = time_launch(c, tune_nonloc, 3, &ms))) break;
            if (ms < best) { best = ms; pick = kind; }
        }
        c->tune_nonloc = pick;
    }
    if (!rc) rc = launch_hpsi_impl(c, 0);  // state as before the trials
    if (!rc) HIPCHECK(hipStreamSynchronize(c->stream));
    c->timing = timing;
    if (c->sweep_exec) { hipGraphExecDestroy(c->sweep_exec); c->sweep_exec = nullptr; }
    if (choices) {
        for (int s = 0; s < 3; s++) choices[s] = c->tune_gemm[s];
        choices[3] = c->tune_nonloc;
    }
    return rc;
}

extern "C" int champ_gpu_set_tuning(champ_ctx *c, const int32_t *choices)
{
    if (!c || !choices) return CHAMP_ERR_ARG;
    for (int s = 0; s < 3; s++) {
        const int k = choices[s];
        const bool f32_ok = s == 0 && c->prec == CHAMP_PREC_MIXED;
        if (k < GEMM_AUTO || k > GEMM_SPLITK || (k == GEMM_F32 && !f32_ok) || (k == GEMM_SPLITK && s != 0)) {
            set_error("set_tuning: GEMM kind " + std::to_string(k) + " not available at site " + std::to_string(s));
            return CHAMP_ERR_ARG;
        }
    }
    if (choices[3] < -1 || choices[3] > 1) { set_error("set_tuning: quadrature form must be -1, 0 or 1"); return CHAMP_ERR_ARG; }
    if (c->nloc == 0 && (choices[2] != GEMM_AUTO || choices[3] != -1)) {
        set_error("set_tuning: no ECP sites on an all-electron system");
        return CHAMP_ERR_ARG;
    }
    if (choices[0] == GEMM_SPLITK) {
        const int rc = gemm_reserve_splitk(c, 4 * c->W, c->NO);
        if (rc) return rc;
    }
    for (int s = 0; s < 3; s++) c->tune_gemm[s] = choices[s];
    c->tune_nonloc = choices[3];
    if (c->sweep_exec) { HIPCHECK(hipGraphExecDestroy(c->sweep_exec)); c->sweep_exec = nullptr; }
    return CHAMP_OK;
}

extern "C" int champ_gpu_set_lu_every(champ_ctx *c, int32_t nmoves)
{
    if (!c || nmoves < 0) return CHAMP_ERR_ARG;
    c->lu_every = nmoves;
    if (c->sweep_exec) { HIPCHECK(hipGraphExecDestroy(c->sweep_exec)); c->sweep_exec = nullptr; }
    return CHAMP_OK;
}

extern "C" int champ_gpu_set_graphs(champ_ctx *c, int32_t on)
{
    if (!c) return CHAMP_ERR_ARG;
    c->graphs = on != 0;
    return CHAMP_OK;
}

// nsweep sweeps with the context's current mover (c->imetro)
static int run_sweeps(champ_ctx *c, int32_t nsweep, double tau, champ_sweep_stats *stats)
{
    c->tau = tau;
    HIPCHECK(hipMemsetAsync(c->esum_dev, 0, sizeof(double) * 4, c->stream));
    int rc;
    if (c->graphs && !c->timing && nsweep > 0) {
        if ((rc = sweep_graph(c, tau))) return rc;
        for (int isw = 0; isw < nsweep; isw++) HIPCHECK(hipGraphLaunch(c->sweep_exec, c->stream));
    } else {
        for (int isw = 0; isw < nsweep; isw++) {
            if ((rc = launch_sweep(c))) return rc;
            if (c->timing && (isw % 4) == 3) tflush(c);
        }
    }
    return sweep_stats(c, nsweep, stats);
}

extern "C" int champ_gpu_metrop1(champ_ctx *c, int32_t nsweep, double tau, champ_sweep_stats *stats)
{
    if (!c || nsweep < 0 || !(tau > 0)) return CHAMP_ERR_ARG;
    c->imetro = 1;
    return run_sweeps(c, nsweep, tau, stats);
}

extern "C" int champ_gpu_set_metrop6(champ_ctx *c, double deltar, double deltat, double fbias)
{
    if (!c || !std::isfinite(deltar) || !std::isfinite(deltat) || !std::isfinite(fbias)) return CHAMP_ERR_ARG;
    // the input checks of parser.f90:774-787: out-of-range steps are reset, fbias truncated to [0, 2]
    c->m6_deltar = deltar < 1.0 ? 2.0 : deltar;
    c->m6_deltat = (deltat < 0.0 || deltat > 2.0) ? 2.0 : deltat;
    c->m6_fbias = std::min(2.0, std::max(0.0, fbias));
    return CHAMP_OK;
}

extern "C" int champ_gpu_metrop6(champ_ctx *c, int32_t nsweep, champ_sweep_stats *stats)
{
    if (!c || nsweep < 0) return CHAMP_ERR_ARG;
    c->imetro = 6;
    const int rc = run_sweeps(c, nsweep, c->tau, stats);
    c->imetro = 1;
    return rc;
}

extern "C" int champ_gpu_vmc(champ_ctx *c, int32_t nblk, int32_t nstep, int32_t nblkeq, double tau, int32_t imetro,
                             double *enow, champ_vmc_result *res)
{
    if (!c || nblk < 1 || nstep < 1 || nblkeq < 0 || !(tau > 0) || (imetro != 1 && imetro != 6)) {
        set_error("vmc: nblk, nstep >= 1, nblkeq >= 0, tau > 0 and imetro 1 or 6 required");
        return CHAMP_ERR_ARG;
    }
    int rc;
    champ_sweep_stats st;
    // the mover of vmc.f90:120-127: metrop6 for imetro 6 (its parameters from
    // champ_gpu_set_metrop6; tau unused), metrop1 otherwise
    auto sweeps = [&](champ_sweep_stats *s) {
        return imetro == 6 ? champ_gpu_metrop6(c, nstep, s) : champ_gpu_metrop1(c, nstep, tau, s);
    };
    // equilibration blocks, not accumulated (vmc.f90: vmc_nblkeq before zerest)
    for (int b = 0; b < nblkeq; b++)
        if ((rc = sweeps(nullptr))) return rc;
    // acuest: ecum += esum, ecm2 += esum * enow, wcum += wsum per block
    double wcum = 0, ecum = 0, ecm2 = 0, e2cum = 0, acc = 0, prop = 0;
    const double wsum = (double)c->W * nstep;
    for (int b = 0; b < nblk; b++) {
        if ((rc = sweeps(&st))) return rc;
        const double en = st.esum / wsum;
        if (enow) enow[b] = en;
        wcum += wsum; ecum += st.esum; ecm2 += st.esum * en; e2cum += st.esum2;
        acc += st.accepted; prop += st.proposed;
    }
    if (res) {
        const double efin = ecum / wcum;
        res->efin = efin;
        res->eerr = std::sqrt(std::fabs(ecm2 / wcum - efin * efin) / nblk);
        res->sigma = std::sqrt(std::fabs(e2cum / wcum - efin * efin));
        const double err1 = res->sigma / std::sqrt(wcum);
        res->tcorr = err1 > 0 ? (res->eerr / err1) * (res->eerr / err1) : 0.0;
        res->accept = prop > 0 ? acc / prop : 0.0;
        res->nsamples = (int64_t)wcum;
        res->nflagged = st.nflagged;
    }
    return CHAMP_OK;
}

extern "C" int champ_gpu_metrop1_batch(champ_ctx *const *ctxs, int32_t n, int32_t nsweep, double tau,
                                       champ_sweep_stats *stats)
{
    if (!ctxs || n < 1 || nsweep < 0 || !(tau > 0)) return CHAMP_ERR_ARG;
    int nemax = 0;
    for (int m = 0; m < n; m++) {
        if (!ctxs[m]) return CHAMP_ERR_ARG;
        for (int k = 0; k < m; k++)
            if (ctxs[k] == ctxs[m]) { set_error("metrop1_batch: a context appears twice"); return CHAMP_ERR_ARG; }
        nemax = std::max(nemax, ctxs[m]->NE);
    }
    for (int m = 0; m < n; m++) { ctxs[m]->tau = tau; ctxs[m]->imetro = 1; }
    for (int m = 0; m < n; m++) HIPCHECK(hipMemsetAsync(ctxs[m]->esum_dev, 0, sizeof(double) * 4, ctxs[m]->stream));
    // Each population's kernels go to its own stream in the order of its
    // own metrop1 sweep; the host deals the move steps round-robin over the
    // populations so that all of them have work queued at once, and queues
    // a population's hpsi right after its last electron.
    int rc;
    for (int isw = 0; isw < nsweep; isw++) {
        for (int m = 0; m < n; m++)
            if (ctxs[m]->nloc > 0 && (rc = launch_rotqua(ctxs[m]))) return rc;
        for (int iel = 0; iel < nemax; iel++)
            for (int m = 0; m < n; m++) {
                champ_ctx *c = ctxs[m];
                if (iel >= c->NE) continue;
                if ((rc = launch_move_step(c, iel))) return rc;
                if (iel == c->NE - 1 && (rc = launch_hpsi_impl(c, 1))) return rc;
            }
        for (int m = 0; m < n; m++)
            if (ctxs[m]->timing && (isw % 4) == 3) tflush(ctxs[m]);
    }
    for (int m = 0; m < n; m++)
        if ((rc = sweep_stats(ctxs[m], nsweep, stats ? stats + m : nullptr))) return rc;
    return CHAMP_OK;
}

extern "C" int champ_gpu_get_walker_scalars(champ_ctx *c, double *psido, double *psijo, double *eold, double *psi2o)
{
    if (!c) return CHAMP_ERR_ARG;
    int rc;
    if ((rc = copy_out(c, psido, c->psido, c->W))) return rc;
    if ((rc = copy_out(c, psijo, c->psijo, c->W))) return rc;
    if ((rc = copy_out(c, eold, c->eold, c->W))) return rc;
    if ((rc = copy_out(c, psi2o, c->psi2o, c->W))) return rc;
    HIPCHECK(hipStreamSynchronize(c->stream));
    return CHAMP_OK;
}

extern "C" int champ_gpu_get_status(champ_ctx *c, int32_t *status)
{
    if (!c || !status) return CHAMP_ERR_ARG;
    HIPCHECK(hipStreamSynchronize(c->stream));
    HIPCHECK(hipMemcpy(status, c->status, sizeof(int) * c->W, hipMemcpyDeviceToHost));
    return CHAMP_OK;
}

extern "C" int64_t champ_gpu_array_size(champ_ctx *c, int32_t which)
{
    if (!c) return -1;
    const int64_t W = c->W, NE = c->NE, NM = c->NM, NO = c->NO;
    switch (which) {
    case CHAMP_ARR_SLMI: return W * 2 * NM * NM;
    case CHAMP_ARR_VJ: case CHAMP_ARR_DDX: case CHAMP_ARR_VJN: return W * NE * 3;
    case CHAMP_ARR_D2DX2: return W * NE;
    case CHAMP_ARR_DETIAB: case CHAMP_ARR_VPSP: return W * 2;
    case CHAMP_ARR_PELOC: case CHAMP_ARR_D2J: case CHAMP_ARR_NXQUAD: return W;
    case CHAMP_ARR_ORBN: return W * 4 * NO;
    case CHAMP_ARR_FSO: return W * NE * NE;
    default: return -1;
    }
}

extern "C" int champ_gpu_get_array(champ_ctx *c, int32_t which, double *out)
{
    if (!c || !out) return CHAMP_ERR_ARG;
    const int64_t n = champ_gpu_array_size(c, which);
    if (n < 0) return CHAMP_ERR_ARG;
    HIPCHECK(hipStreamSynchronize(c->stream));
    const size_t W = c->W, NE = c->NE;
    auto get = [&](const double *d, size_t cnt, std::vector<double> &h) -> int {
        h.resize(cnt);
        HIPCHECK(hipMemcpy(h.data(), d, cnt * 8, hipMemcpyDeviceToHost));
        return CHAMP_OK;
    };
    std::vector<double> h;
    int rc = CHAMP_OK;
    switch (which) {
    case CHAMP_ARR_SLMI: rc = get(c->slmi, n, h); if (!rc) std::memcpy(out, h.data(), n * 8); break;
    case CHAMP_ARR_VJ:
    case CHAMP_ARR_VJN:
        rc = get(which == CHAMP_ARR_VJ ? c->vjT : c->vjnT, W * 3 * NE, h);
        if (!rc) for (size_t w = 0; w < W; w++) for (size_t e = 0; e < NE; e++) for (int k = 0; k < 3; k++)
            out[(w * NE + e) * 3 + k] = h[(w * 3 + k) * NE + e];
        break;
    case CHAMP_ARR_DDX: rc = get(c->ddx, n, h); if (!rc) std::memcpy(out, h.data(), n * 8); break;
    case CHAMP_ARR_D2DX2: rc = get(c->d2dx2, n, h); if (!rc) std::memcpy(out, h.data(), n * 8); break;
    case CHAMP_ARR_DETIAB: rc = get(c->detiab, n, h); if (!rc) std::memcpy(out, h.data(), n * 8); break;
    case CHAMP_ARR_VPSP: rc = get(c->vpsp, n, h); if (!rc) std::memcpy(out, h.data(), n * 8); break;
    case CHAMP_ARR_PELOC: rc = get(c->pe_local, n, h); if (!rc) std::memcpy(out, h.data(), n * 8); break;
    case CHAMP_ARR_D2J: rc = get(c->d2o, n, h); if (!rc) std::memcpy(out, h.data(), n * 8); break;
    case CHAMP_ARR_FSO: rc = get(c->fso, n, h); if (!rc) std::memcpy(out, h.data(), n * 8); break;
    case CHAMP_ARR_ORBN: rc = get(c->orbn, n, h); if (!rc) std::memcpy(out, h.data(), n * 8); break;
    case CHAMP_ARR_NXQUAD: {
        std::vector<int> hi(W);
        HIPCHECK(hipMemcpy(hi.data(), c->nxquad, W * sizeof(int), hipMemcpyDeviceToHost));
        for (size_t w = 0; w < W; w++) out[w] = hi[w];
        break;
    }
    default: return CHAMP_ERR_ARG;
    }
    return rc;
}

extern "C" int champ_gpu_basis_fns(champ_ctx *c, int32_t npt, const double *pos, double *phin)
{
    if (!c || npt <= 0 || !pos || !phin) return CHAMP_ERR_ARG;
    double *dpos = nullptr, *dout = nullptr;
    HIPCHECK(hipMalloc(&dpos, (size_t)npt * 3 * 8));
    if (hipMalloc(&dout, (size_t)npt * 5 * c->NB * 8) != hipSuccess) { hipFree(dpos); set_error("basis_fns: alloc"); return CHAMP_ERR_NOMEM; }
    int rc = CHAMP_OK;
    hipError_t e = hipMemcpy(dpos, pos, (size_t)npt * 3 * 8, hipMemcpyHostToDevice);
    if (e == hipSuccess) rc = launch_basis_points(c, npt, dpos, dout);
    if (e == hipSuccess && !rc) e = hipStreamSynchronize(c->stream);
    if (e == hipSuccess && !rc) e = hipMemcpy(phin, dout, (size_t)npt * 5 * c->NB * 8, hipMemcpyDeviceToHost);
    if (e != hipSuccess) {
        set_error(std::string("basis_fns: ") + hipGetErrorString(e));
        rc = CHAMP_ERR_HIP;
    }
    hipFree(dpos);
    hipFree(dout);
    return rc;
}

extern "C" int champ_gpu_basis_fns_noalloc(champ_ctx *c, int32_t npt, const double *pos, double *phin)
{
    if (!c || npt <= 0 || !pos || !phin) return CHAMP_ERR_ARG;
    const size_t npos = ((size_t)npt * 3 + 31) / 32 * 32;  // the values start 256-byte aligned
    if (npos + (size_t)npt * 5 * c->NB > NOALLOC_DOUBLES) {
        set_error("basis_fns_noalloc: npt * (3 + 5 nbasis) doubles exceed the 16 MiB scratch (at most " +
                  std::to_string((NOALLOC_DOUBLES - 32) / (3 + 5 * (size_t)c->NB)) + " points here)");
        return CHAMP_ERR_ARG;
    }
    if (!c->scratch) { set_error("basis_fns_noalloc: no scratch address for this context's device"); return CHAMP_ERR_STATE; }
    double *dpos = c->scratch, *dout = c->scratch + npos;
    HIPCHECK(hipMemcpyAsync(dpos, pos, (size_t)npt * 3 * 8, hipMemcpyHostToDevice, c->stream));
    int rc = launch_basis_points(c, npt, dpos, dout);
    if (rc) return rc;
    HIPCHECK(hipMemcpyAsync(phin, dout, (size_t)npt * 5 * c->NB * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHECK(hipStreamSynchronize(c->stream));
    return CHAMP_OK;
}

static const char *kclass_name[K_NCLASS] = {"move_jastrow_ao", "move_mo_gemm", "move_ratio_commit", "hpsi_jastrow",
                                            "hpsi_ao", "hpsi_mo_gemm", "hpsi_slater_lu", "hpsi_ecp_prep", "hpsi_nonloc",
                                            "hpsi_nonloc_psi", "hpsi_eloc", "multidet", "other"};

extern "C" int champ_gpu_kernel_times(champ_ctx *c, int32_t *n, char *names, double *ms, int64_t *launches)
{
    if (!c || !n) return CHAMP_ERR_ARG;
    if (!names && !ms) {  // toggle / reset: *n > 0 enables timing, 0 disables
        c->timing = *n > 0;
        for (int k = 0; k < K_NCLASS; k++) { c->class_ms[k] = 0; c->class_n[k] = 0; }
        *n = K_NCLASS;
        return CHAMP_OK;
    }
    tflush(c);
    const int m = std::min<int>(*n, K_NCLASS);
    for (int k = 0; k < m; k++) {
        if (names) { std::strncpy(names + 32 * k, kclass_name[k], 31); names[32 * k + 31] = 0; }
        if (ms) ms[k] = c->class_ms[k];
        if (launches) launches[k] = c->class_n[k];
    }
    *n = m;
    return CHAMP_OK;
}
