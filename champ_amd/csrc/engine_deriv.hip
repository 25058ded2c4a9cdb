// engine_deriv.hip -- derivatives of ln psi with respect to the wave
// function's parameters at the walkers' current configurations (the gvalue
// of deriv_jastrow4.f90:3-667 for the linear Jastrow parameters, and the
// determinant ratios of optci.f90), and the population sums of stochastic
// reconfiguration (optwf_sr.f90:71-470: <O_k>, <E_L O_k>, <O_k O_l>).
//
// Parameters, in this order (kinds selects the groups):
//   e-n   a4[o] of each centre type, o = 2..norda (the polynomial part):
//         O = sum over electrons and centres of the type of R^o - R_inf^o
//   e-e   b[o] of each spin class, o = 2..nordb: O = sum over pairs of u^o - u_inf^o
//   e-e-n c_t of each centre type, every term t of the order-nordc
//         polynomial (jastrow4.f90:203-213 order, no cusp elimination):
//         O = sum over pairs and centres of the type of u^k (x^l + y^l) (x y)^m
//         with the switched distances u, x, y
//   CI    cdet_k of every determinant (ndet > 1): O = D_up,k D_dn,k / psi_D
// The sums use the distances hpsi left in the walker state, so hpsi (or a
// metrop1 sweep, which ends with one) must precede a call.
#include "een_poly.hpp"
#include "engine_internal.hpp"
#include "hpsi_args.hpp"

#include <algorithm>
#include <vector>

namespace champ {

struct DerLayout {
    int na, nb, nc, nci;  // parameters per group
    int n;                // all of them
};

static DerLayout der_layout(const champ_ctx *c, int kinds)
{
    const JasParams &j = c->hjas;
    DerLayout L;
    L.na = (kinds & CHAMP_DERIV_JAS_EN) ? j.nctype * std::max(0, j.norda - 1) : 0;
    L.nb = (kinds & CHAMP_DERIV_JAS_EE) ? j.nspin2b * std::max(0, j.nordb - 1) : 0;
    L.nc = (kinds & CHAMP_DERIV_JAS_EEN) ? j.nctype * j.nterms : 0;
    L.nci = ((kinds & CHAMP_DERIV_CI) && c->md) ? c->md_ndet : 0;
    L.n = L.na + L.nb + L.nc + L.nci;
    return L;
}

void deriv_free(champ_ctx *c)
{
    void *ps[] = {c->der_ot, c->der_s, c->der_v};
    for (void *p : ps) if (p) hipFree(p);
    c->der_ot = c->der_s = c->der_v = nullptr;
    c->der_cap = 0;
}

struct DerArgs {
    int W, NE, NUP, NC, na, nb, nc, nci, n;
    const double *xT, *rr_en;
    const int *iwctype;
    const JasParams *jas;
    double *ot;  // [n][W]
    // CI
    int NU;
    const int *du, *dd, *cur;
    const double *D, *S, *ph;
};

// Jastrow groups: one block per walker.  e-n over (electron, centre) pairs,
// e-e and e-e-n over electron pairs; fixed-order reductions.
template <int NORDC>
__global__ void __launch_bounds__(256) k_deriv_jastrow(DerArgs a)
{
    __shared__ double red[16];
    const int w = blockIdx.x, tid = threadIdx.x, bd = blockDim.x;
    const int NE = a.NE, NC = a.NC;
    const JasParams &jp = *a.jas;
    const double kap = jp.kap, ainf = jp.asymp_r;
    const double *rrw = a.rr_en + (size_t)w * NC * NE;
    const double *xw = a.xT + (size_t)w * 3 * NE;
    int off = 0;
    // e-n: a4[o] of type it
    if (a.na) {
        for (int it = 0; it < jp.nctype; it++)
            for (int o = 2; o <= jp.norda; o++) {
                double s = 0.0;
                const double pinf = powi(ainf, o);
                for (int t = tid; t < NC * NE; t += bd) {
                    const int cc = t / NE;
                    if (a.iwctype[cc] != it) continue;
                    s += powi(rrw[t], o) - pinf;
                }
                s = block_sum(s, red);
                if (tid == 0) a.ot[(size_t)(off + it * (jp.norda - 1) + (o - 2)) * a.W + w] = s;
            }
        off += a.na;
    }
    const int npair = NE * (NE - 1) / 2;
    // e-e: b[o] of spin class isb (like-spin pairs are class 1 when nspin2b = 2)
    if (a.nb) {
        for (int isb = 0; isb < jp.nspin2b; isb++)
            for (int o = 2; o <= jp.nordb; o++) {
                double s = 0.0;
                const double pinf = powi(ainf, o);
                for (int t = tid; t < npair; t += bd) {
                    // pair t = i (i - 1) / 2 + j, j < i
                    int i = (int)((1.0 + sqrt(1.0 + 8.0 * t)) * 0.5);
                    while (i * (i - 1) / 2 > t) i--;
                    while ((i + 1) * i / 2 <= t) i++;
                    const int j = t - i * (i - 1) / 2;
                    const bool like = (i < a.NUP) == (j < a.NUP);
                    if (jp.nspin2b == 2 && (like ? 1 : 0) != isb) continue;
                    const double d0 = xw[i] - xw[j], d1 = xw[NE + i] - xw[NE + j], d2 = xw[2 * NE + i] - xw[2 * NE + j];
                    double r2 = 0;
                    r2 = r2 + d0 * d0; r2 = r2 + d1 * d1; r2 = r2 + d2 * d2;
                    s += powi(scale_r(kap, sqrt(r2)), o) - pinf;
                }
                s = block_sum(s, red);
                if (tid == 0) a.ot[(size_t)(off + isb * (jp.nordb - 1) + (o - 2)) * a.W + w] = s;
            }
        off += a.nb;
    }
    // e-e-n: c_t of type it
    if constexpr (NORDC > 1) {
        if (a.nc) {
            constexpr EenTable<NORDC> ET{};
            for (int it = 0; it < jp.nctype; it++) {
                double acc[ET.n];
#pragma unroll
                for (int q = 0; q < ET.n; q++) acc[q] = 0.0;
                for (int t = tid; t < npair; t += bd) {
                    int i = (int)((1.0 + sqrt(1.0 + 8.0 * t)) * 0.5);
                    while (i * (i - 1) / 2 > t) i--;
                    while ((i + 1) * i / 2 <= t) i++;
                    const int j = t - i * (i - 1) / 2;
                    const double d0 = xw[i] - xw[j], d1 = xw[NE + i] - xw[NE + j], d2 = xw[2 * NE + i] - xw[2 * NE + j];
                    double r2 = 0;
                    r2 = r2 + d0 * d0; r2 = r2 + d1 * d1; r2 = r2 + d2 * d2;
                    const double rij = sqrt(r2);
                    const double u = switched_from<double>(kap, rij, scale_r(kap, rij));
                    double up[NORDC + 1];
                    up[0] = 1.0;
#pragma unroll
                    for (int p = 1; p <= NORDC; p++) up[p] = up[p - 1] * u;
                    for (int cc = 0; cc < NC; cc++) {
                        if (a.iwctype[cc] != it) continue;
                        const double x = 1.0 - kap * rrw[(size_t)cc * NE + i], y = 1.0 - kap * rrw[(size_t)cc * NE + j];
                        double xp[NORDC + 1], yp[NORDC + 1];
                        xp[0] = 1.0; yp[0] = 1.0;
#pragma unroll
                        for (int p = 1; p <= NORDC; p++) { xp[p] = xp[p - 1] * x; yp[p] = yp[p - 1] * y; }
#pragma unroll
                        for (int q = 0; q < ET.n; q++)
                            acc[q] += up[ET.k[q]] * (xp[ET.l[q]] + yp[ET.l[q]]) * (xp[ET.m[q]] * yp[ET.m[q]]);
                    }
                }
                for (int q = 0; q < ET.n; q++) {
                    const double s = block_sum(acc[q], red);
                    if (tid == 0) a.ot[(size_t)(off + it * jp.nterms + q) * a.W + w] = s;
                }
            }
        }
    }
}

// CI: O_k = phase_k D_up(u_k) D_dn(d_k) / S, thread per (walker, determinant)
__global__ void k_deriv_ci(DerArgs a, int off)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)a.nci * a.W) return;
    const int k = (int)(t / a.W), w = (int)(t % a.W);
    const double *Du = a.D + (((size_t)w * 2 + 0) * 2 + a.cur[2 * w]) * (size_t)a.NU;
    const double *Dd = a.D + (((size_t)w * 2 + 1) * 2 + a.cur[2 * w + 1]) * (size_t)a.NU;
    a.ot[(size_t)(off + k) * a.W + w] = a.ph[k] * Du[a.du[k]] * Dd[a.dd[k]] / a.S[w];
}

// sums over the walkers in walker order: v[k] = sum_w O_kw, v[n + k] = sum_w E_w O_kw
__global__ void k_deriv_sums(int n, int W, const double *ot, const double *eold, double *v)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const double *o = ot + (size_t)k * W;
    double s = 0.0, se = 0.0;
    for (int w = 0; w < W; w++) { s += o[w]; se += eold[w] * o[w]; }
    v[k] = s;
    v[n + k] = se;
}

static int deriv_compute(champ_ctx *c, int kinds, DerLayout &L)
{
    L = der_layout(c, kinds);
    if (L.n == 0) { set_error("deriv: no parameters selected (kinds) for this trial function"); return CHAMP_ERR_ARG; }
    if (c->hjas.nordc > 5) { set_error("deriv: e-e-n order above 5 is not supported"); return CHAMP_ERR_UNSUPPORTED; }
    if (L.n > c->der_cap) {
        deriv_free(c);
        HIPCHECK(hipMalloc(&c->der_ot, (size_t)L.n * c->W * sizeof(double)));
        HIPCHECK(hipMalloc(&c->der_s, (size_t)L.n * L.n * sizeof(double)));
        HIPCHECK(hipMalloc(&c->der_v, (size_t)2 * L.n * sizeof(double)));
        c->der_cap = L.n;
    }
    DerArgs a;
    a.W = c->W; a.NE = c->NE; a.NUP = c->NUP; a.NC = c->NC;
    a.na = L.na; a.nb = L.nb; a.nc = L.nc; a.nci = L.nci; a.n = L.n;
    a.xT = c->xT; a.rr_en = c->rr_en; a.iwctype = c->iwctype; a.jas = c->jas; a.ot = c->der_ot;
    a.NU = c->md_NU; a.du = c->md_du; a.dd = c->md_dd; a.cur = c->md_cur; a.D = c->md_D; a.S = c->md_S;
    a.ph = c->md_ph;
    if (L.na + L.nb + L.nc) {
        switch (c->hjas.nordc) {
        case 2: hipLaunchKernelGGL(k_deriv_jastrow<2>, dim3(c->W), dim3(256), 0, c->stream, a); break;
        case 3: hipLaunchKernelGGL(k_deriv_jastrow<3>, dim3(c->W), dim3(256), 0, c->stream, a); break;
        case 4: hipLaunchKernelGGL(k_deriv_jastrow<4>, dim3(c->W), dim3(256), 0, c->stream, a); break;
        case 5: hipLaunchKernelGGL(k_deriv_jastrow<5>, dim3(c->W), dim3(256), 0, c->stream, a); break;
        default: hipLaunchKernelGGL(k_deriv_jastrow<0>, dim3(c->W), dim3(256), 0, c->stream, a); break;
        }
        HIPCHECK(hipGetLastError());
    }
    if (L.nci) {
        const long long nt = (long long)L.nci * c->W;
        hipLaunchKernelGGL(k_deriv_ci, dim3((unsigned)((nt + 255) / 256)), dim3(256), 0, c->stream, a,
                           L.na + L.nb + L.nc);
        HIPCHECK(hipGetLastError());
    }
    return CHAMP_OK;
}

}  // namespace champ

using namespace champ;

extern "C" int champ_gpu_deriv_count(champ_ctx *c, int32_t kinds, int32_t *nparm)
{
    if (!c || !nparm) return CHAMP_ERR_ARG;
    *nparm = der_layout(c, kinds).n;
    return CHAMP_OK;
}

extern "C" int champ_gpu_deriv_psi(champ_ctx *c, int32_t kinds, double *o)
{
    if (!c) return CHAMP_ERR_ARG;
    DerLayout L;
    int rc = deriv_compute(c, kinds, L);
    if (rc) return rc;
    if (o) {
        std::vector<double> h((size_t)L.n * c->W);
        HIPCHECK(hipMemcpyAsync(h.data(), c->der_ot, h.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        HIPCHECK(hipStreamSynchronize(c->stream));
        for (int w = 0; w < c->W; w++)
            for (int k = 0; k < L.n; k++) o[(size_t)w * L.n + k] = h[(size_t)k * c->W + w];
    }
    return CHAMP_OK;
}

extern "C" int champ_gpu_sr_sums(champ_ctx *c, int32_t kinds, double *s, double *osum, double *eosum, double *esum)
{
    if (!c) return CHAMP_ERR_ARG;
    DerLayout L;
    int rc = deriv_compute(c, kinds, L);
    if (rc) return rc;
    const int n = L.n, W = c->W;
    // <O_k O_l> numerators: the parameter-major O times its transpose on the matrix cores
    if (s) {
        if ((rc = launch_gemm(c, true, n, n, W, c->der_ot, W, 0, c->der_ot, W, 0, c->der_s, n, 0, 1))) return rc;
        HIPCHECK(hipMemcpyAsync(s, c->der_s, (size_t)n * n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    }
    hipLaunchKernelGGL(k_deriv_sums, dim3((n + 63) / 64), dim3(64), 0, c->stream, n, W, c->der_ot, c->eold, c->der_v);
    HIPCHECK(hipGetLastError());
    std::vector<double> v((size_t)2 * n), e(W);
    HIPCHECK(hipMemcpyAsync(v.data(), c->der_v, v.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHECK(hipMemcpyAsync(e.data(), c->eold, (size_t)W * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHECK(hipStreamSynchronize(c->stream));
    if (osum) std::copy(v.begin(), v.begin() + n, osum);
    if (eosum) std::copy(v.begin() + n, v.end(), eosum);
    if (esum) {
        double t = 0.0;
        for (int w = 0; w < W; w++) t += e[w];
        *esum = t;
    }
    return CHAMP_OK;
}
