// engine_deriv.hip -- derivatives of ln psi with respect to the wave
// function's parameters at the walkers' current configurations (the gvalue
// of deriv_jastrow4.f90:3-667 for the Jastrow coefficients, and the
// determinant ratios of optci.f90:55-57), and the population sums of
// stochastic reconfiguration (optwf_sr.f90:504-620: <O_k>, <E_L O_k>,
// <O_k O_l>), and the SR parameter step of optwf_sr.f90:201-215 with the
// preconditioned conjugate gradients of optwf_sr_more.f90:31-160.
//
// Raw parameters, in the order of the system arrays (kinds selects groups):
//   e-n   a4[it][p], p = 0..norda, of each centre type it:
//         p = 0: sum of R/(1 + a4[1] R) minus its value at R_inf,
//         p = 1: sum of -a4[0] R^2/(1 + a4[1] R)^2 minus its value at R_inf,
//         p >= 2: sum of R^p - R_inf^p, over electrons and centres of type it
//   e-e   b[isb][p], p = 0..nordb, of each spin class, the same three forms in
//         the pair's scaled distance u, the p = 0 and 1 forms times the pair's
//         spin factor sspinn (1/2 for like spins when nspin2b = 1 and the
//         cusp is imposed, else 1)
//   e-e-n c[it][t] of each centre type, every term t of the order-nordc
//         polynomial (jastrow4.f90:203-213 order, no cusp elimination):
//         O = sum over pairs and centres of the type of u^k (x^l + y^l) (x y)^m
//         with the switched distances u, x, y
//   CI    cdet[k] of every determinant (ndet > 1): O = D_up,k D_dn,k / psi_D
//   ORB   the orbital terms (io, jo) set by champ_gpu_set_orb_terms, the
//         excitation phi_io += kappa phi_jo of compute_lcao
//         (optwf_handle_wf.f90:1128-1134): O = sum_e V[e][io] phi_jo(r_e)
//         with V[e][o] = d ln psi_D / d phi_o(r_e) (the reference inverse
//         for one determinant, the multideterminant V otherwise), the
//         orbprim of optorb_deriv (optorb.f90:5-97)
// A parameter map (champ_gpu_set_deriv_map) turns these into a caller's own
// parameters, e.g. CHAMP's iwjasa/iwjasb/iwjasc lists with the e-e-n cusp
// dependencies and the CSF combinations: O' = O M.
// The sums use the distances hpsi left in the walker state, so hpsi (or a
// sweep, which ends with one) must precede a call.
#include "een_poly.hpp"
#include "engine_internal.hpp"
#include "hpsi_args.hpp"

#include <algorithm>
#include <cmath>
#include <functional>
#include <vector>

namespace champ {

struct DerLayout {
    int na, nb, nc, nci, norb;  // parameters per group
    int n;                      // all of them
};

static DerLayout der_layout(const champ_ctx *c, int kinds)
{
    const JasParams &j = c->hjas;
    DerLayout L;
    L.na = (kinds & CHAMP_DERIV_JAS_EN) ? j.nctype * (j.norda + 1) : 0;
    L.nb = (kinds & CHAMP_DERIV_JAS_EE) ? j.nspin2b * (j.nordb + 1) : 0;
    L.nc = (kinds & CHAMP_DERIV_JAS_EEN) ? j.nctype * j.nterms : 0;
    L.nci = ((kinds & CHAMP_DERIV_CI) && c->md) ? c->md_ndet : 0;
    L.norb = (kinds & CHAMP_DERIV_ORB) ? c->der_norb : 0;
    L.n = L.na + L.nb + L.nc + L.nci + L.norb;
    return L;
}

static void deriv_free_scratch(champ_ctx *c)
{
    void *ps[] = {c->der_ot, c->der_s, c->der_v, c->der_om};
    for (void *p : ps) if (p) hipFree(p);
    c->der_ot = c->der_s = c->der_v = c->der_om = nullptr;
    c->der_cap = c->der_scap = c->der_om_cap = 0;
}

static void sr_free(champ_ctx *c)
{
    void *ps[] = {c->sr_acc, c->sr_accv, c->sr_z};
    for (void *p : ps) if (p) hipFree(p);
    c->sr_acc = c->sr_accv = c->sr_z = nullptr;
    c->sr_n = c->sr_kinds = 0;
    c->sr_wsum = c->sr_esum = 0.0;
}

static void map_free(champ_ctx *c)
{
    if (c->der_map) hipFree(c->der_map);
    if (c->der_src) hipFree(c->der_src);
    if (c->der_dt) hipFree(c->der_dt);
    c->der_map = c->der_dt = nullptr;
    c->der_src = nullptr;
    c->der_nd = c->der_dt_cap = 0;
    c->der_map_kinds = c->der_map_nraw = c->der_map_nout = 0;
}

void deriv_free(champ_ctx *c)
{
    deriv_free_scratch(c);
    sr_free(c);
    map_free(c);
    if (c->der_orb) hipFree(c->der_orb);
    if (c->der_occ) hipFree(c->der_occ);
    c->der_orb = c->der_occ = nullptr;
    c->der_norb = 0;
}

// d/dp of the e-n (e-e) form f(r) = q0 r / (1 + q1 r) + sum_{p>=2} q_p r^p
// at the scaled distance r: p = 0 and 1 are the Pade coefficients.
__device__ __forceinline__ double pade_deriv(int p, double q0, double q1, double r)
{
    if (p == 0) return r / (1.0 + q1 * r);
    if (p == 1) {
        const double bot = 1.0 + q1 * r;
        return -q0 * r * r / (bot * bot);
    }
    return powi(r, p);
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
    // e-n: a4[p] of type it
    if (a.na) {
        for (int it = 0; it < jp.nctype; it++) {
            const double a0 = jp.a4[it][0], a1 = jp.a4[it][1];
            for (int p = 0; p <= jp.norda; p++) {
                double s = 0.0;
                const double pinf = pade_deriv(p, a0, a1, ainf);
                for (int t = tid; t < NC * NE; t += bd) {
                    const int cc = t / NE;
                    if (a.iwctype[cc] != it) continue;
                    s += pade_deriv(p, a0, a1, rrw[t]) - pinf;
                }
                s = block_sum(s, red);
                if (tid == 0) a.ot[(size_t)(off + it * (jp.norda + 1) + p) * a.W + w] = s;
            }
        }
        off += a.na;
    }
    const int npair = NE * (NE - 1) / 2;
    // e-e: b[p] of spin class isb (like-spin pairs are class 1 when nspin2b = 2)
    if (a.nb) {
        const double slike = (jp.nspin2b == 1 && jp.nocuspb == 0) ? 0.5 : 1.0;
        for (int isb = 0; isb < jp.nspin2b; isb++) {
            const double b0 = jp.b[isb][0], b1 = jp.b[isb][1];
            for (int p = 0; p <= jp.nordb; p++) {
                double s = 0.0;
                const double pinf = pade_deriv(p, b0, b1, ainf);
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
                    const double g = pade_deriv(p, b0, b1, scale_r(kap, sqrt(r2))) - pinf;
                    s += (p < 2 && like) ? slike * g : g;
                }
                s = block_sum(s, red);
                if (tid == 0) a.ot[(size_t)(off + isb * (jp.nordb + 1) + p) * a.W + w] = s;
            }
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

// Orbital terms: a wave per term, lanes over electrons, for the walkers
// [w0, w0 + nw) whose orbital values are in orb ([nw][NE][NO]).
struct DerOrbArgs {
    int W, NE, NUP, NM, NO, w0, nterm, off, md;
    const int *terms, *occ;
    const double *orb, *slmi, *V;
    double *ot;
};

__global__ void __launch_bounds__(256) k_deriv_orb(DerOrbArgs a)
{
    const int wc = blockIdx.x, w = a.w0 + wc, lane = threadIdx.x % WAVE, wv = threadIdx.x / WAVE;
    const int nwv = blockDim.x / WAVE, NE = a.NE, NO = a.NO;
    const double *orb = a.orb + (size_t)wc * NE * NO;
    for (int k = wv; k < a.nterm; k += nwv) {
        const int io = a.terms[2 * k], jo = a.terms[2 * k + 1];
        double s = 0.0;
        for (int e = lane; e < NE; e += WAVE) {
            double v;
            if (a.md) {
                v = a.V[((size_t)w * NE + e) * NO + io];
            } else {
                const int sp = e < a.NUP ? 0 : 1, el = e - (sp ? a.NUP : 0), j = a.occ[sp * NO + io];
                // slmi row el holds (A^-1)[j][el] at column j: d ln det / d phi_{o_j}(r_e)
                v = j >= 0 ? a.slmi[((size_t)w * 2 + sp) * a.NM * a.NM + (size_t)el * a.NM + j] : 0.0;
            }
            s += v * orb[(size_t)e * NO + jo];
        }
        s = wave_sum(s);
        if (lane == 0) a.ot[(size_t)(a.off + k) * a.W + w] = s;
    }
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

// dst += src over n elements
__global__ void k_sr_add(size_t n, const double *__restrict__ src, double *__restrict__ dst)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        dst[i] += src[i];
}

// O' [nout][W] of the map: row j copies raw row src[j], or, for src[j] < 0,
// row -1 - src[j] of the dense columns' product dt
__global__ void k_deriv_assemble(int nout, int W, const int *__restrict__ src, const double *__restrict__ ot,
                                 const double *__restrict__ dt, double *__restrict__ om)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)nout * W) return;
    const int j = (int)(t / W), w = (int)(t % W), sj = src[j];
    om[t] = sj >= 0 ? ot[(size_t)sj * W + w] : dt[(size_t)(-1 - sj) * W + w];
}

// dst += src on the lower triangle (j <= i) of n x n: row i = blockIdx.y
__global__ void k_sr_add_lower(int n, const double *__restrict__ src, double *__restrict__ dst)
{
    const int i = blockIdx.y, j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j > i || j >= n) return;
    dst[(size_t)i * n + j] += src[(size_t)i * n + j];
}

// S[i][j] = S[j][i] for j > i, through a 32 x 32 LDS tile: block (bx, by)
// with bx >= by writes the upper tile (by, bx) from the lower tile (bx, by)
__global__ void __launch_bounds__(256) k_sr_symm(int n, double *S)
{
    __shared__ double tile[32][33];
    const int bx = blockIdx.x, by = blockIdx.y;
    if (bx < by) return;
    const int tx = threadIdx.x % 32, ty = threadIdx.x / 32;
    for (int r = ty; r < 32; r += 8) {
        const int gi = bx * 32 + r, gj = by * 32 + tx;
        tile[r][tx] = (gi < n && gj < n) ? S[(size_t)gi * n + gj] : 0.0;
    }
    __syncthreads();
    for (int r = ty; r < 32; r += 8) {
        const int gi = by * 32 + r, gj = bx * 32 + tx;  // upper element (gi, gj) = lower (gj, gi)
        if (gi < n && gj < n && gj > gi) S[(size_t)gi * n + gj] = tile[tx][r];
    }
}

// t = S z for the accumulated S [n][n] (symmetric): a wavefront per row, the
// row read once with consecutive lanes on consecutive columns
__global__ void __launch_bounds__(256) k_sr_gemv(int n, const double *__restrict__ S, const double *__restrict__ z,
                                                 double *__restrict__ t)
{
    const int lane = threadIdx.x % WAVE;
    const int row = blockIdx.x * (blockDim.x / WAVE) + threadIdx.x / WAVE;
    if (row >= n) return;
    const double *r = S + (size_t)row * n;
    double acc = 0.0;
    for (int l = lane; l < n; l += WAVE) acc += r[l] * z[l];
    acc = wave_sum(acc);
    if (lane == 0) t[row] = acc;
}

__global__ void k_sr_diag(int n, const double *__restrict__ S, double *__restrict__ d)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) d[k] = S[(size_t)k * n + k];
}

template <typename T>
static int grow(T **p, int *cap, int need, size_t per)
{
    if (need <= *cap) return CHAMP_OK;
    if (*p) hipFree(*p);
    *p = nullptr;
    *cap = 0;
    HIPCHECK(hipMalloc((void **)p, (size_t)need * per));
    *cap = need;
    return CHAMP_OK;
}

// O of the selected parameters, parameter-major: *ot = [n][W] on the device.
// With CHAMP_DERIV_MAPPED the raw parameters are the map's, and the result is
// Mt [nout][nraw] x O [nraw][W] (a matrix-core GEMM).
static int deriv_compute(champ_ctx *c, int kinds, const double **ot, int *n)
{
    const bool mapped = (kinds & CHAMP_DERIV_MAPPED) != 0;
    if (mapped && !c->der_map_nout) { set_error("deriv: CHAMP_DERIV_MAPPED without a parameter map"); return CHAMP_ERR_ARG; }
    const int raw = mapped ? c->der_map_kinds : kinds;
    DerLayout L = der_layout(c, raw);
    if (L.n == 0) { set_error("deriv: no parameters selected (kinds) for this trial function"); return CHAMP_ERR_ARG; }
    if (c->hjas.nordc > 5) { set_error("deriv: e-e-n order above 5 is not supported"); return CHAMP_ERR_UNSUPPORTED; }
    const int nout = mapped ? c->der_map_nout : L.n;
    int rc;
    if ((rc = grow(&c->der_ot, &c->der_cap, L.n, (size_t)c->W * sizeof(double)))) return rc;
    if (nout > c->der_scap) {
        if (c->der_s) hipFree(c->der_s);
        if (c->der_v) hipFree(c->der_v);
        c->der_s = c->der_v = nullptr;
        c->der_scap = 0;
        HIPCHECK(hipMalloc(&c->der_s, (size_t)nout * nout * sizeof(double)));
        HIPCHECK(hipMalloc(&c->der_v, (size_t)2 * nout * sizeof(double)));
        c->der_scap = nout;
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
    if (L.norb) {
        DerOrbArgs o;
        o.W = c->W; o.NE = c->NE; o.NUP = c->NUP; o.NM = c->NM; o.NO = c->NO; o.nterm = L.norb;
        o.off = L.na + L.nb + L.nc + L.nci; o.md = c->md;
        o.terms = c->der_orb; o.occ = c->der_occ; o.orb = c->orball; o.slmi = c->slmi; o.V = c->md_V;
        o.ot = c->der_ot;
        // orbital values chunk by chunk in the hpsi scratch (hpsi rewrites it)
        for (int w0 = 0; w0 < c->W; w0 += c->hpsi_chunk) {
            const int nw = std::min(c->hpsi_chunk, c->W - w0);
            if ((rc = launch_orbital_values(c, w0, nw))) return rc;
            o.w0 = w0;
            hipLaunchKernelGGL(k_deriv_orb, dim3(nw), dim3(256), 0, c->stream, o);
            HIPCHECK(hipGetLastError());
        }
    }
    if (mapped) {
        // dense columns on the matrix cores, unit columns copied
        if (c->der_nd > 0) {
            if ((rc = grow(&c->der_dt, &c->der_dt_cap, c->der_nd, (size_t)c->W * sizeof(double)))) return rc;
            if ((rc = launch_gemm(c, false, c->der_nd, c->W, L.n, c->der_map, L.n, 0, c->der_ot, c->W, 0, c->der_dt,
                                  c->W, 0, 1)))
                return rc;
        }
        if ((rc = grow(&c->der_om, &c->der_om_cap, nout, (size_t)c->W * sizeof(double)))) return rc;
        const long long nt = (long long)nout * c->W;
        hipLaunchKernelGGL(k_deriv_assemble, dim3((unsigned)((nt + 255) / 256)), dim3(256), 0, c->stream, nout, c->W,
                           c->der_src, c->der_ot, c->der_dt, c->der_om);
        HIPCHECK(hipGetLastError());
        *ot = c->der_om;
    } else {
        *ot = c->der_ot;
    }
    *n = nout;
    return CHAMP_OK;
}

// ---- SR step (host): sr_hs, optwf_sr.f90:504-620 with one state and the
// walker weights summed in wsum; pcg + asolve + atimes_n,
// optwf_sr_more.f90:31-160, on the dense S of the sums.
struct SrSystem {
    int n;
    std::vector<double> o, s_diag, s_ii_inv, h;  // o = <O_k>
    // t = <O O> z (the sums divided by the weight): a dense host product or a
    // matrix-vector product over the sums on the GPU
    std::function<int(const double *, double *)> oo_times;
};

// oo_diag = <O_k O_k>, eo = <E O_k>, e = <E>
static void sr_hs(SrSystem &A, double adiag, const std::vector<double> &oo_diag, const std::vector<double> &eo,
                  double e)
{
    const int n = A.n;
    double smax = 0.0;
    for (int k = 0; k < n; k++) {
        const double aux = oo_diag[k] - A.o[k] * A.o[k];
        A.s_diag[k] = aux * adiag;
        A.s_ii_inv[k] = aux + A.s_diag[k];
        A.h[k] = -2.0 * (eo[k] - A.o[k] * e);
        if (A.s_ii_inv[k] > smax) smax = A.s_ii_inv[k];
    }
    for (int k = 0; k < n; k++)
        A.s_ii_inv[k] = (A.s_ii_inv[k] / smax > 1e-14) ? 1.0 / A.s_ii_inv[k] : 0.0;
}

// r = S z with S_kl = <O_k O_l> - <O_k><O_l> + s_diag_k delta_kl
static int sr_atimes(const SrSystem &A, const double *z, double *r)
{
    const int n = A.n;
    double oz = 0.0;
    for (int k = 0; k < n; k++) oz += A.o[k] * z[k];
    int rc = A.oo_times(z, r);
    if (rc) return rc;
    for (int k = 0; k < n; k++) r[k] = r[k] - A.o[k] * oz + A.s_diag[k] * z[k];
    return CHAMP_OK;
}

static int sr_pcg(const SrSystem &A, double eps, std::vector<double> &x, int *iters)
{
    const int n = A.n, imax = n, imod = 50;
    std::vector<double> r(n), d(n), q(n), sv(n);
    auto dot = [n](const double *u, const double *v) {
        double t = 0.0;
        for (int k = 0; k < n; k++) t += u[k] * v[k];
        return t;
    };
    int rc;
    x.assign(n, 0.0);
    if ((rc = sr_atimes(A, x.data(), r.data()))) return rc;
    for (int k = 0; k < n; k++) r[k] = A.h[k] - r[k];
    for (int k = 0; k < n; k++) d[k] = r[k] * A.s_ii_inv[k];
    double delta_new = dot(d.data(), r.data());
    const double delta_0 = delta_new * eps * eps;
    int i = 0;
    for (; i < imax; i++) {
        if (delta_new < delta_0) break;
        if ((rc = sr_atimes(A, d.data(), q.data()))) return rc;
        const double alpha = delta_new / dot(d.data(), q.data());
        for (int k = 0; k < n; k++) x[k] += alpha * d[k];
        if (i % imod == 0) {
            if ((rc = sr_atimes(A, x.data(), r.data()))) return rc;
            for (int k = 0; k < n; k++) r[k] = A.h[k] - r[k];
        } else {
            for (int k = 0; k < n; k++) r[k] -= alpha * q[k];
        }
        for (int k = 0; k < n; k++) sv[k] = r[k] * A.s_ii_inv[k];
        const double delta_old = delta_new;
        delta_new = dot(r.data(), sv.data());
        const double beta = delta_new / delta_old;
        for (int k = 0; k < n; k++) d[k] = sv[k] + beta * d[k];
    }
    *iters = i;
    return CHAMP_OK;
}

// optwf_sr.f90:201-231: solve, dparm = -sr_tau x, and test_solution_parm
// (optwf_handle_wf.f90:1338-1365): a step whose rms exceeds dparm_norm_min is
// redone with ten times the diagonal shift
static int sr_step(SrSystem &A, const std::vector<double> &oo_diag, const std::vector<double> &eo, double e,
                   double sr_tau, double sr_eps, double sr_adiag, double dparm_norm_min, double *dparm,
                   champ_sr_info *info)
{
    const int n = A.n;
    A.s_diag.resize(n);
    A.s_ii_inv.resize(n);
    A.h.resize(n);
    double adiag = sr_adiag, norm = 0.0;
    int iters = 0, tries = 0, rc;
    std::vector<double> x;
    for (;; tries++) {
        sr_hs(A, adiag, oo_diag, eo, e);
        if ((rc = sr_pcg(A, sr_eps, x, &iters))) return rc;
        norm = 0.0;
        for (int k = 0; k < n; k++) { dparm[k] = -sr_tau * x[k]; norm += dparm[k] * dparm[k]; }
        norm = std::sqrt(norm / n);
        if (!(adiag > 0.0) || norm <= dparm_norm_min) break;
        if (tries == 30) { set_error("sr_solve: no step below dparm_norm_min"); return CHAMP_ERR_NUMERIC; }
        adiag *= 10.0;
    }
    if (info) { info->adiag = adiag; info->dparm_norm = norm; info->cg_iters = iters; info->retries = tries; }
    return CHAMP_OK;
}

}  // namespace champ

using namespace champ;

extern "C" int champ_gpu_deriv_count(champ_ctx *c, int32_t kinds, int32_t *nparm)
{
    if (!c || !nparm) return CHAMP_ERR_ARG;
    if (kinds & CHAMP_DERIV_MAPPED) {
        if (!c->der_map_nout) { set_error("deriv: CHAMP_DERIV_MAPPED without a parameter map"); return CHAMP_ERR_ARG; }
        *nparm = c->der_map_nout;
    } else {
        *nparm = der_layout(c, kinds).n;
    }
    return CHAMP_OK;
}

extern "C" int champ_gpu_set_deriv_map(champ_ctx *c, int32_t kinds, int32_t nout, const double *map)
{
    if (!c) return CHAMP_ERR_ARG;
    map_free(c);
    if (!map || nout == 0) return CHAMP_OK;
    const int raw_bits = CHAMP_DERIV_JAS_EN | CHAMP_DERIV_JAS_EE | CHAMP_DERIV_JAS_EEN | CHAMP_DERIV_CI | CHAMP_DERIV_ORB;
    if (nout < 0 || (kinds & ~raw_bits)) {
        set_error("set_deriv_map: bad kinds or nout");
        return CHAMP_ERR_ARG;
    }
    const int nraw = der_layout(c, kinds).n;
    if (nraw == 0) { set_error("set_deriv_map: no parameters selected (kinds) for this trial function"); return CHAMP_ERR_ARG; }
    // a column with a single non-zero equal to 1 copies that raw parameter;
    // the others keep their coefficients, transposed ([nd][nraw])
    std::vector<int> nnz(nout, 0), one(nout, -1);
    for (int k = 0; k < nraw; k++) {
        const double *row = map + (size_t)k * nout;
        for (int j = 0; j < nout; j++)
            if (row[j] != 0.0) {
                nnz[j]++;
                one[j] = row[j] == 1.0 ? k : -2;
            }
    }
    std::vector<int> src(nout);
    int nd = 0;
    for (int j = 0; j < nout; j++) src[j] = (nnz[j] == 1 && one[j] >= 0) ? one[j] : -1 - nd++;
    std::vector<double> mt((size_t)nd * nraw);
    for (int k = 0; k < nraw; k++) {
        const double *row = map + (size_t)k * nout;
        for (int j = 0; j < nout; j++)
            if (src[j] < 0) mt[(size_t)(-1 - src[j]) * nraw + k] = row[j];
    }
    HIPCHECK(hipMalloc(&c->der_src, (size_t)nout * sizeof(int)));
    HIPCHECK(hipMemcpy(c->der_src, src.data(), (size_t)nout * sizeof(int), hipMemcpyHostToDevice));
    if (nd > 0) {
        HIPCHECK(hipMalloc(&c->der_map, mt.size() * sizeof(double)));
        HIPCHECK(hipMemcpy(c->der_map, mt.data(), mt.size() * sizeof(double), hipMemcpyHostToDevice));
    }
    c->der_nd = nd;
    c->der_map_kinds = kinds;
    c->der_map_nraw = nraw;
    c->der_map_nout = nout;
    return CHAMP_OK;
}

extern "C" int champ_gpu_set_orb_terms(champ_ctx *c, int32_t nterm, const int32_t *ideriv)
{
    if (!c || nterm < 0 || (nterm > 0 && !ideriv)) return CHAMP_ERR_ARG;
    for (int k = 0; k < 2 * nterm; k++)
        if (ideriv[k] < 0 || ideriv[k] >= c->NO) { set_error("set_orb_terms: orbital index outside 0..norb-1"); return CHAMP_ERR_ARG; }
    if (c->der_orb) hipFree(c->der_orb);
    if (c->der_occ) hipFree(c->der_occ);
    c->der_orb = c->der_occ = nullptr;
    c->der_norb = 0;
    if (nterm == 0) return CHAMP_OK;
    // the slot of each orbital in the reference determinant, per spin
    std::vector<int> iw(c->NE), occ((size_t)2 * c->NO, -1);
    HIPCHECK(hipMemcpy(iw.data(), c->iworbd, (size_t)c->NE * sizeof(int), hipMemcpyDeviceToHost));
    for (int e = 0; e < c->NE; e++) {
        const int sp = e < c->NUP ? 0 : 1;
        occ[(size_t)sp * c->NO + iw[e]] = e - (sp ? c->NUP : 0);
    }
    HIPCHECK(hipMalloc(&c->der_orb, (size_t)2 * nterm * sizeof(int)));
    HIPCHECK(hipMemcpy(c->der_orb, ideriv, (size_t)2 * nterm * sizeof(int), hipMemcpyHostToDevice));
    HIPCHECK(hipMalloc(&c->der_occ, occ.size() * sizeof(int)));
    HIPCHECK(hipMemcpy(c->der_occ, occ.data(), occ.size() * sizeof(int), hipMemcpyHostToDevice));
    c->der_norb = nterm;
    // a map over the previous raw layout no longer fits
    if (c->der_map_nout && (c->der_map_kinds & CHAMP_DERIV_ORB)) map_free(c);
    return CHAMP_OK;
}

extern "C" int champ_gpu_deriv_psi(champ_ctx *c, int32_t kinds, double *o)
{
    if (!c) return CHAMP_ERR_ARG;
    const double *ot = nullptr;
    int n = 0;
    int rc = deriv_compute(c, kinds, &ot, &n);
    if (rc) return rc;
    if (o) {
        std::vector<double> h((size_t)n * c->W);
        HIPCHECK(hipMemcpyAsync(h.data(), ot, h.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        HIPCHECK(hipStreamSynchronize(c->stream));
        for (int w = 0; w < c->W; w++)
            for (int k = 0; k < n; k++) o[(size_t)w * n + k] = h[(size_t)k * c->W + w];
    }
    return CHAMP_OK;
}

extern "C" int champ_gpu_sr_sums(champ_ctx *c, int32_t kinds, double *s, double *osum, double *eosum, double *esum)
{
    if (!c) return CHAMP_ERR_ARG;
    const double *ot = nullptr;
    int n = 0;
    int rc = deriv_compute(c, kinds, &ot, &n);
    if (rc) return rc;
    const int W = c->W;
    // <O_k O_l> numerators: the parameter-major O times its transpose on the matrix cores
    if (s) {
        if ((rc = launch_gemm(c, true, n, n, W, ot, W, 0, ot, W, 0, c->der_s, n, 0, 1))) return rc;
        HIPCHECK(hipMemcpyAsync(s, c->der_s, (size_t)n * n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    }
    hipLaunchKernelGGL(k_deriv_sums, dim3((n + 63) / 64), dim3(64), 0, c->stream, n, W, ot, c->eold, c->der_v);
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

extern "C" int champ_sr_solve(int32_t nparm, double wsum, const double *s, const double *osum, const double *eosum,
                              double esum, double sr_tau, double sr_eps, double sr_adiag, double dparm_norm_min,
                              double *dparm, champ_sr_info *info)
{
    if (nparm <= 0 || !(wsum > 0.0) || !s || !osum || !eosum || !dparm) { set_error("sr_solve: bad arguments"); return CHAMP_ERR_ARG; }
    const int n = nparm;
    SrSystem A;
    A.n = n;
    A.o.resize(n);
    std::vector<double> eo(n), diag(n), oo((size_t)n * n);
    for (size_t k = 0; k < (size_t)n * n; k++) oo[k] = s[k] / wsum;
    for (int k = 0; k < n; k++) { A.o[k] = osum[k] / wsum; eo[k] = eosum[k] / wsum; diag[k] = oo[(size_t)k * n + k]; }
    A.oo_times = [&oo, n](const double *z, double *t) {
        for (int k = 0; k < n; k++) {
            double u = 0.0;
            const double *row = &oo[(size_t)k * n];
            for (int l = 0; l < n; l++) u += row[l] * z[l];
            t[k] = u;
        }
        return (int)CHAMP_OK;
    };
    return sr_step(A, diag, eo, esum / wsum, sr_tau, sr_eps, sr_adiag, dparm_norm_min, dparm, info);
}

extern "C" int champ_gpu_sr_accumulate(champ_ctx *c, int32_t kinds, int32_t reset, double *wsum, double *esum)
{
    if (!c) return CHAMP_ERR_ARG;
    const double *ot = nullptr;
    int n = 0;
    int rc = deriv_compute(c, kinds, &ot, &n);
    if (rc) return rc;
    if (n > 65535) { set_error("sr_accumulate: more than 65535 parameters"); return CHAMP_ERR_UNSUPPORTED; }
    if (!reset && c->sr_acc && (c->sr_n != n || c->sr_kinds != kinds)) {
        set_error("sr_accumulate: the parameters differ from the accumulated ones (start again with reset)");
        return CHAMP_ERR_ARG;
    }
    if (reset || !c->sr_acc) {
        if (c->sr_n != n) {
            sr_free(c);
            HIPCHECK(hipMalloc(&c->sr_acc, (size_t)n * n * sizeof(double)));
            HIPCHECK(hipMalloc(&c->sr_accv, (size_t)2 * n * sizeof(double)));
            HIPCHECK(hipMalloc(&c->sr_z, (size_t)2 * n * sizeof(double)));
        }
        HIPCHECK(hipMemsetAsync(c->sr_acc, 0, (size_t)n * n * sizeof(double), c->stream));
        HIPCHECK(hipMemsetAsync(c->sr_accv, 0, (size_t)2 * n * sizeof(double), c->stream));
        c->sr_n = n;
        c->sr_kinds = kinds;
        c->sr_wsum = c->sr_esum = 0.0;
    }
    const int W = c->W;
    // S is symmetric: its lower triangle here, the upper one mirrored by champ_gpu_sr_solve
    if ((rc = launch_syrk_lower(c, n, W, ot, W, c->der_s, n))) return rc;
    hipLaunchKernelGGL(k_sr_add_lower, dim3((n + 255) / 256, n), dim3(256), 0, c->stream, n, c->der_s, c->sr_acc);
    HIPCHECK(hipGetLastError());
    hipLaunchKernelGGL(k_deriv_sums, dim3((n + 63) / 64), dim3(64), 0, c->stream, n, W, ot, c->eold, c->der_v);
    HIPCHECK(hipGetLastError());
    hipLaunchKernelGGL(k_sr_add, dim3((2 * n + 255) / 256), dim3(256), 0, c->stream, (size_t)2 * n, c->der_v,
                       c->sr_accv);
    HIPCHECK(hipGetLastError());
    std::vector<double> e(W);
    HIPCHECK(hipMemcpyAsync(e.data(), c->eold, (size_t)W * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHECK(hipStreamSynchronize(c->stream));
    double t = 0.0;
    for (int w = 0; w < W; w++) t += e[w];
    c->sr_esum += t;
    c->sr_wsum += W;
    if (wsum) *wsum = c->sr_wsum;
    if (esum) *esum = c->sr_esum;
    return CHAMP_OK;
}

extern "C" int champ_gpu_sr_solve(champ_ctx *c, double sr_tau, double sr_eps, double sr_adiag, double dparm_norm_min,
                                  double *dparm, champ_sr_info *info)
{
    if (!c || !dparm) return CHAMP_ERR_ARG;
    if (!c->sr_acc || !(c->sr_wsum > 0.0)) { set_error("sr_solve: no sums accumulated"); return CHAMP_ERR_ARG; }
    const int n = c->sr_n;
    const double wsum = c->sr_wsum;
    SrSystem A;
    A.n = n;
    A.o.resize(n);
    std::vector<double> v((size_t)2 * n), eo(n), diag(n);
    const int nt = (n + 31) / 32;
    hipLaunchKernelGGL(k_sr_symm, dim3(nt, nt), dim3(256), 0, c->stream, n, c->sr_acc);
    HIPCHECK(hipGetLastError());
    hipLaunchKernelGGL(k_sr_diag, dim3((n + 255) / 256), dim3(256), 0, c->stream, n, c->sr_acc, c->sr_z);
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipMemcpyAsync(diag.data(), c->sr_z, (size_t)n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHECK(hipMemcpyAsync(v.data(), c->sr_accv, v.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHECK(hipStreamSynchronize(c->stream));
    for (int k = 0; k < n; k++) { A.o[k] = v[k] / wsum; eo[k] = v[(size_t)n + k] / wsum; diag[k] /= wsum; }
    A.oo_times = [c, n, wsum](const double *z, double *t) {
        HIPCHECK(hipMemcpyAsync(c->sr_z, z, (size_t)n * sizeof(double), hipMemcpyHostToDevice, c->stream));
        hipLaunchKernelGGL(k_sr_gemv, dim3((n + 3) / 4), dim3(256), 0, c->stream, n, c->sr_acc, c->sr_z, c->sr_z + n);
        HIPCHECK(hipGetLastError());
        HIPCHECK(hipMemcpyAsync(t, c->sr_z + n, (size_t)n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        HIPCHECK(hipStreamSynchronize(c->stream));
        for (int k = 0; k < n; k++) t[k] /= wsum;
        return (int)CHAMP_OK;
    };
    return sr_step(A, diag, eo, c->sr_esum / wsum, sr_tau, sr_eps, sr_adiag, dparm_norm_min, dparm, info);
}
