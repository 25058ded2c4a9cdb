// engine_hpsi.hip -- FP64 once-per-sweep kernels: hpsi (hpsi.f90:3-273) over the
// whole population.  distances + jastrow_factor4, AO values for all
// electrons, the AO->MO GEMM, batched LU inverse of both Slater matrices,
// nonloc_pot quadrature and the local-energy assembly.
#include "een_poly.hpp"
#include "hpsi_args.hpp"

#include <algorithm>
#include <cmath>

namespace champ {

// ------------------------------------------------------------------------
// distances(0) + jastrow_factor4 (jastrow4.f90:3-415, iforce_analy = 0),
// one 256-thread block per walker.
// Tuning of k_hpsi_jastrow: partner slots per lane in the pair loop and the
// minimum waves per SIMD asked of the register allocator.
#ifndef CHAMP_HJB
#define CHAMP_HJB 1
#endif
#ifndef CHAMP_HJAS_WPE
#define CHAMP_HJAS_WPE 1
#endif
constexpr int HJB = CHAMP_HJB;

// T: arithmetic type of the pair loop and its e-e-n tables (double on the
// FP64 path, float on the mixed path); distances, e-n terms and all sums stay
// double.
template <int NORDC, typename T>
__global__ void __launch_bounds__(512) CHAMP_WAVES_PER_EU(CHAMP_HJAS_WPE) k_hpsi_jastrow(HArgs a)
{
    HIP_DYNAMIC_SHARED(double, lds)
    __shared__ double red[16];
    const int w = blockIdx.x, tid = threadIdx.x, bd = blockDim.x;
    const int NE = a.NE, NC = a.NC;
    const JasParams &jp = *a.jas;
    const double kap = jp.kap, asymp_r = jp.asymp_r;
    const double *xw = a.xT + (size_t)w * 3 * NE;
    double *rw = a.r_en + (size_t)w * NC * NE, *rrw = a.rr_en + (size_t)w * NC * NE;
    double *F = a.fso + (size_t)w * NE * NE;
    double *G1 = a.g1 + (size_t)w * NE * 3 * NE, *G2 = a.g2 + (size_t)w * NE * 3 * NE;

    for (int t = tid; t < NC * NE; t += bd) {
        const int c = t / NE, e = t % NE;
        const double r = en_distance(xw, NE, e, a.cent, c);
        const double rr = scale_r(kap, r);
        rw[t] = r;
        rrw[t] = rr;
        if (a.enp) a.enp[(size_t)w * NC * NE + t] = enp_pack(kap, asymp_r, r, rr);
    }
    __syncthreads();

    double fsp = 0, d2p = 0;
    // e-n terms (jastrow4.f90:300-370)
    for (int e = tid; e < NE; e += bd) {
        double fs = 0, g[3] = {0, 0, 0}, d2 = 0;
        for (int c = 0; c < NC; c++) {
            const double r = rw[(size_t)c * NE + e], rr = rrw[(size_t)c * NE + e];
            if (rr == asymp_r) continue;
            const double dd7 = 1.0 - kap * rr, dd9 = -kap * dd7;
            const int it = a.iwctype[c];
            const double *a4 = jp.a4[it];
            double top = a4[0] * rr, topi = a4[0], topii = 0;
            double bot = 1 + a4[1] * rr, boti = a4[1], botii = 0, bot2 = bot * bot;
            double fen = top / bot - jp.ja[it];
            double feni = topi / bot - boti * top / bot2;
            double fenii = topii - (botii * top + 2 * boti * topi) / bot + 2 * boti * boti * top / bot2;
            fenii = fenii / bot;
            double rp1 = rr, rp2 = 1.0;  // rri(iord-1), rri(iord-2)
            for (int iord = 2; iord <= jp.norda; iord++) {
                const double rp = powi(rr, iord);
                fen = fen + a4[iord] * rp;
                feni = feni + a4[iord] * iord * rp1;
                fenii = fenii + a4[iord] * iord * (iord - 1) * rp2;
                rp2 = rp1; rp1 = rp;
            }
            fenii = fenii * dd7 * dd7 + feni * dd9;
            feni = feni * dd7 / r;
            fs = fs + fen;
            g[0] += feni * (xw[e] - a.cent[3 * c]);
            g[1] += feni * (xw[NE + e] - a.cent[3 * c + 1]);
            g[2] += feni * (xw[2 * NE + e] - a.cent[3 * c + 2]);
            d2 = d2 + fenii + 2 * feni;
        }
        F[(size_t)e * NE + e] = fs;
        for (int k = 0; k < 3; k++) { G1[((size_t)e * 3 + k) * NE + e] = g[k]; G2[((size_t)e * 3 + k) * NE + e] = g[k]; }
        fsp += fs; d2p += d2;
    }

    // e-e and e-e-n pair terms (jastrow4.f90:79-296).  One wave per row i
    // (rows dealt longest first), lanes over the partners j < i in slots of
    // WAVE.  The row's per-centre e-e-n tables (value and two x-derivatives,
    // een_poly.hpp) and scalars are built by the wave into its LDS region.
    constexpr int NA = een_naz<NORDC>();
    constexpr int RS = 3 * NA + 8;  // doubles per centre: tables, r, rr, dd7, dd9, rvec (3), 1/r
    const int lane = tid % WAVE, wv = tid / WAVE, nwv = bd / WAVE;
    T *rowt = (T *)lds + (size_t)wv * NC * RS;
    const T tkap = (T)kap;
    for (int rr0 = wv; rr0 < NE - 1; rr0 += nwv) {
        const int i = NE - 1 - rr0;
        if constexpr (NORDC > 1) {
            __syncwarp();  // the previous row's table is no longer read
            for (int c = lane; c < NC; c += WAVE) {
                T *rt = rowt + (size_t)c * RS;
                const double ri = rw[(size_t)c * NE + i], rri = rrw[(size_t)c * NE + i];
                const T sw = (T)(1.0 - kap * rri);
                een_tables<NORDC, 2, T>(jp.c[a.iwctype[c]], sw, rt);
                rt[3 * NA] = (T)ri;
                rt[3 * NA + 1] = rri == asymp_r ? T(1) : T(0);  // electron i asymptotic for centre c
                rt[3 * NA + 2] = -tkap * sw;           // dd7
                rt[3 * NA + 3] = -tkap * (-tkap * sw);  // dd9
                for (int k = 0; k < 3; k++) rt[3 * NA + 4 + k] = (T)(xw[k * NE + i] - a.cent[3 * c + k]);
                rt[3 * NA + 7] = T(1) / (T)ri;
            }
            __syncwarp();
        }
        for (int j0 = 0; j0 < i; j0 += WAVE * HJB) {
            T fs[HJB], d2[HJB], fI[HJB][3], fJ[HJB][3], rv[HJB][3], xj[HJB][3], sfu[HJB];
            T uu1[HJB], dd1e[HJB], dd2e[HJB], irij[HJB], rij2[HJB];
            bool on[HJB];
#pragma unroll
            for (int s = 0; s < HJB; s++) {
                const int j = j0 + s * WAVE + lane;
                on[s] = j < i;
                const int jq = on[s] ? j : 0;
                fs[s] = 0; d2[s] = 0; sfu[s] = 0;
                for (int k = 0; k < 3; k++) {
                    fI[s][k] = 0; fJ[s][k] = 0;
                    const double xjd = xw[k * NE + jq];
                    xj[s][k] = (T)xjd;
                    rv[s][k] = (T)(xw[k * NE + i] - xjd);
                }
                T r2 = 0;
                r2 = r2 + rv[s][0] * rv[s][0]; r2 = r2 + rv[s][1] * rv[s][1]; r2 = r2 + rv[s][2] * rv[s][2];
                const T rij = sqrt(r2);
                rij2[s] = rij * rij;
                irij[s] = T(1) / rij;
                T sspinn = 1; int isb = 0, ipar = 0;
                if (i < a.NUP || jq >= a.NUP) {
                    if (jp.nspin2b == 2) isb = 1;
                    else if (jp.nocuspb == 0) sspinn = T(0.5);
                    ipar = 1;
                }
                const T u = scale_r<T>(tkap, rij);
                const T dd1 = switched_from<T>(kap, rij, u), dd2 = -tkap * dd1;
                uu1[s] = dd1;
                dd1e[s] = -tkap * dd1;
                dd2e[s] = -tkap * dd2;
                if (on[s] && (sizeof(T) == sizeof(float) || u != (T)asymp_r)) {
                    const double *b = jp.b[isb];
                    T top = sspinn * (T)b[0] * u, topu = sspinn * (T)b[0], topuu = 0;
                    T bot = 1 + (T)b[1] * u, botu = (T)b[1], botuu = 0, bot2 = bot * bot;
                    T fee = top / bot - (T)jp.jb[ipar];
                    T feeu = topu / bot - botu * top / bot2;
                    T feeuu = topuu - (botuu * top + 2 * botu * topu) / bot + 2 * botu * botu * top / bot2;
                    feeuu = feeuu / bot;
                    T u1 = u, u2 = 1;  // uu(iord-1), uu(iord-2)
                    for (int iord = 2; iord <= jp.nordb; iord++) {
                        const T un = u * u1;
                        fee = fee + (T)b[iord] * un;
                        feeu = feeu + (T)b[iord] * (T)iord * u1;
                        feeuu = feeuu + (T)b[iord] * (T)(iord * (iord - 1)) * u2;
                        u2 = u1; u1 = un;
                    }
                    feeuu = feeuu * dd1 * dd1 + feeu * dd2;
                    feeu = feeu * dd1 / rij;
                    if constexpr (sizeof(T) == sizeof(float))  // value minus asymptote without cancellation
                        fee = pade_minus_asym<T>(b, jp.nordb, sspinn, u, asym_offset<T>(asymp_r, dd1), (T)asymp_r);
                    fs[s] += fee;
                    for (int k = 0; k < 3; k++) { fI[s][k] += feeu * rv[s][k]; fJ[s][k] -= feeu * rv[s][k]; }
                    d2[s] += 2 * (feeuu + 2 * feeu);
                }
            }
            if constexpr (NORDC > 1) {
                // partner distances to centre c with one centre of lookahead
                // (slots past the row read electron 0 and are not used)
                double rjn[HJB], rrjn[HJB];
                int jcl[HJB];
#pragma unroll
                for (int s = 0; s < HJB; s++) {
                    const int j = j0 + s * WAVE + lane;
                    jcl[s] = j < i ? j : 0;
                    rjn[s] = rw[jcl[s]];
                    rrjn[s] = rrw[jcl[s]];
                }
                for (int c = 0; c < NC; c++) {
                    const T *rt = rowt + (size_t)c * RS;
                    const T ri = rt[3 * NA], dd7 = rt[3 * NA + 2], dd9 = rt[3 * NA + 3];
                    const bool iasym = rt[3 * NA + 1] != T(0);
                    const T rvi0 = rt[3 * NA + 4], rvi1 = rt[3 * NA + 5], rvi2 = rt[3 * NA + 6], iri = rt[3 * NA + 7];
                    const T cx = (T)a.cent[3 * c], cy = (T)a.cent[3 * c + 1], cz = (T)a.cent[3 * c + 2];
                    double rjc[HJB], rrjc[HJB];
#pragma unroll
                    for (int s = 0; s < HJB; s++) {
                        rjc[s] = rjn[s]; rrjc[s] = rrjn[s];
                        if (c + 1 < NC) {
                            rjn[s] = rw[(size_t)(c + 1) * NE + jcl[s]];
                            rrjn[s] = rrw[(size_t)(c + 1) * NE + jcl[s]];
                        }
                    }
#pragma unroll
                    for (int s = 0; s < HJB; s++) {
                        if (j0 + s * WAVE >= i) break;
                        if (!on[s]) continue;
                        const T rj = (T)rjc[s];
                        if (iasym && rrjc[s] == asymp_r) continue;
                        const T swj = (T)(1.0 - kap * rrjc[s]);
                        const T dd8 = -tkap * swj, dd10 = -tkap * (-tkap * swj);
                        const EenOut<T> e = een_eval_tab<NORDC, 2, T>(rt, uu1[s], swj);
                        const T d1 = dd1e[s];
                        const T fuu = e.fuu * d1 * d1 + e.fu * dd2e[s];
                        const T fu = e.fu * d1 * irij[s];
                        const T fui = e.fux * d1 * dd7, fuj = e.fuy * d1 * dd8;
                        const T fii = e.fxx * dd7 * dd7 + e.fx * dd9, fjj = e.fyy * dd8 * dd8 + e.fy * dd10;
                        const T irj = T(1) / rj;
                        const T fi = e.fx * dd7 * iri, fj = e.fy * dd8 * irj;
                        fs[s] += e.f;
                        sfu[s] += fu;
                        const T st = (ri + rj) * (ri - rj);
                        const T u2pst = rij2[s] + st, u2mst = rij2[s] - st;
                        fI[s][0] += fi * rvi0; fI[s][1] += fi * rvi1; fI[s][2] += fi * rvi2;
                        fJ[s][0] += fj * (xj[s][0] - cx); fJ[s][1] += fj * (xj[s][1] - cy); fJ[s][2] += fj * (xj[s][2] - cz);
                        d2[s] = d2[s] + 2 * (fuu + 2 * fu) + fui * u2pst * (iri * irij[s]) + fuj * u2mst * (irj * irij[s]) + fii +
                                2 * fi + fjj + 2 * fj;
                    }
                }
            }
#pragma unroll
            for (int s = 0; s < HJB; s++) {
                const int j = j0 + s * WAVE + lane;
                if (!on[s]) continue;
                for (int k = 0; k < 3; k++) { fI[s][k] += sfu[s] * rv[s][k]; fJ[s][k] -= sfu[s] * rv[s][k]; }
                const double fsd = (double)fs[s];
                F[(size_t)i * NE + j] = fsd;
                F[(size_t)j * NE + i] = fsd;
                for (int k = 0; k < 3; k++) {
                    const double gI = (double)fI[s][k], gJ = (double)fJ[s][k];
                    G1[((size_t)i * 3 + k) * NE + j] = gI;
                    G2[((size_t)i * 3 + k) * NE + j] = gJ;
                    G1[((size_t)j * 3 + k) * NE + i] = gJ;
                    G2[((size_t)j * 3 + k) * NE + i] = gI;
                }
                fsp += fsd; d2p += (double)d2[s];
            }
        }
    }
    __syncthreads();
    double *vj = a.vjT + (size_t)w * 3 * NE;
    for (int t = tid; t < 3 * NE; t += bd) {
        const int e = t % NE, k = t / NE;
        const double *row = G1 + ((size_t)e * 3 + k) * NE;
        double s = 0;
        for (int j = 0; j < NE; j++) s += row[j];
        vj[(size_t)k * NE + e] = s;
    }
    const double fsum = block_sum(fsp, red);
    const double d2sum = block_sum(d2p, red);
    if (tid == 0) { a.fsumo[w] = fsum; a.d2o[w] = d2sum; }
}

// ------------------------------------------------------------------------
// AO values, gradients and Laplacians of every electron (basis_fns with
// ider=2 inside orbitals, orbitals_no_qmckl.f90:46): one wavefront per
// (walker, electron), lanes over AOs.  Chunk of walkers [w0, w0+nw).
// T: arithmetic type of the radial functions and AO values (double on the
// FP64 path, float on the mixed path); the AO block is stored as double.
// NCOMP 5: value, gradient and Laplacian; 1: values only (the periodic
// recompute of the Slater inverse inside a sweep).
template <typename T, int NCOMP>
__global__ void __launch_bounds__(256) k_hpsi_ao(HArgs a)
{
    HIP_DYNAMIC_SHARED(double, lds)
    const int lane = threadIdx.x % WAVE, wv = threadIdx.x / WAVE;
    const long long gw = (long long)blockIdx.x * (blockDim.x / WAVE) + wv;
    if (gw >= (long long)a.nw * a.NE) return;  // whole wavefront
    const int wc = (int)(gw / a.NE), e = (int)(gw % a.NE), w = a.w0 + wc;
    const int NE = a.NE, NRI = a.NRI;
    const double *xw = a.xT + (size_t)w * 3 * NE;
    double *out = a.aoall + (size_t)gw * NCOMP * a.NB;
    const double x0 = xw[e], x1 = xw[NE + e], x2 = xw[2 * NE + e];
    // radial functions (R, R', R'') once per (centre, shell), in the wave's LDS.
    // The distances are formed here (the expression that wrote r_en): faster
    // than reading them back from r_en.
    T *R = (T *)lds + (size_t)wv * 3 * NRI;
    for (int ri = lane; ri < NRI; ri += WAVE) {
        const double r = en_distance(xw, NE, e, a.cent, a.ri_cent[ri]);
        T wfv[3] = {T(0), T(0), T(0)};
        if (r <= a.ri_rmax[ri])
            splfit_jx<T>(a.tab, a.ri_type[ri], a.ri_rad[ri], (T)r, grid_jx(a.tab, r), wfv, NCOMP == 1 ? 0 : 2);
        R[3 * ri] = wfv[0]; R[3 * ri + 1] = wfv[1]; R[3 * ri + 2] = wfv[2];
    }
    __syncwarp();
    for (int m = lane; m < a.NB; m += WAVE) {
        const int c = a.ao_cent[m];
        const T *Rm = R + 3 * a.ao_ri[m];
        if constexpr (NCOMP == 1) {
            T o0 = T(0);
            if (Rm[0] != T(0))
                o0 = slm_value<T>(slm_decode(a.ao_slm[m]), (T)(x0 - a.cent[3 * c]), (T)(x1 - a.cent[3 * c + 1]),
                                  (T)(x2 - a.cent[3 * c + 2])) * Rm[0];
            out[m] = (double)o0;
            continue;
        }
        T o[5] = {T(0), T(0), T(0), T(0), T(0)};
        if (Rm[0] != T(0) || Rm[1] != T(0) || Rm[2] != T(0))  // else beyond the cut radius: exactly zero
            ao_combine<T>(a.ao_slm[m], (T)(x0 - a.cent[3 * c]), (T)(x1 - a.cent[3 * c + 1]),
                          (T)(x2 - a.cent[3 * c + 2]), (T)en_distance(xw, NE, e, a.cent, c), Rm, 2, o);
        for (int k = 0; k < 5; k++) out[(size_t)k * a.NB + m] = (double)o[k];
    }
}

// ------------------------------------------------------------------------
// determinant (determinant.f90:9-140): per (walker, spin) block, Slater
// matrix in LDS, in-place Gauss-Jordan inverse with partial (row) pivoting,
// determinant from the pivots with matinv's mantissa/exponent loop
// (matinv.f90:65-84), then ddx / d2dx2 and the dorb rows for the moves.
// GLB: the work matrix of more than ~140 electrons of one spin does not fit
// in LDS; it then lives in the block's slice of a global work buffer (L2
// resident during the elimination), with the same arithmetic in the same order.
// R > 0 (up to 16 R electrons per spin, LDS form only): the elimination runs
// on register tiles.  Thread (ti, tj) of a 16 x 16 grid holds the elements
// (ti + 16 a, tj + 16 b), a, b < R; each pivot step exchanges only the pivot
// column and the two swapped rows through LDS (double-buffered), so the
// updates are register FMAs instead of dependent LDS round trips.  Same
// arithmetic and order per element as the LDS form.
template <bool GLB, int R>
__global__ void __launch_bounds__(256) k_hpsi_slater(HArgs a)
{
    // static arrays: SLATER_STATIC_LDS bytes (hpsi_args.hpp)
    HIP_DYNAMIC_SHARED(double, lds)
    __shared__ int perm[512];
    __shared__ double piv[512];
    __shared__ int s_p;
    __shared__ int s_bad;
    __shared__ double s_akk;
    const int wc = blockIdx.x / 2, iab = blockIdx.x % 2, w = a.w0 + wc;
    const int tid = threadIdx.x, bd = blockDim.x, lane = tid % WAVE, wv = tid / WAVE;
    const int NE = a.NE, NO = a.NO;
    const int nel = iab ? NE - a.NUP : a.NUP, ish = iab ? a.NUP : 0;
    const int ld = nel + 1;
    double *L = GLB ? a.lu_ws + (size_t)blockIdx.x * a.NM * (a.NM + 1) : lds;
    double *col = GLB ? lds : L + (size_t)nel * ld;
    const int oc = a.ocomp;
    const double *ob = a.orball + (size_t)wc * NE * oc * NO;
    if (tid == 0) s_bad = 0;
    for (int t = tid; t < nel * nel; t += bd) {
        const int i = t / nel, j = t % nel;
        L[i * ld + j] = ob[((size_t)(ish + i) * oc) * NO + a.iworbd[ish + j]];
    }
    __syncthreads();
    if constexpr (R > 0) {
        static_assert(!GLB, "register tiles with the LDS work matrix only");
        const int ti = tid / 16, tj = tid % 16, NMX = a.NM;
        double m[R][R];
#pragma unroll
        for (int ra = 0; ra < R; ra++)
#pragma unroll
            for (int rb = 0; rb < R; rb++) {
                const int i = ti + 16 * ra, j = tj + 16 * rb;
                m[ra][rb] = (i < nel && j < nel) ? L[i * ld + j] : 0.0;
            }
        double *XB = col;  // [2 parities][C, RK, RP][NM]
        for (int k = 0; k < nel; k++) {
            double *Cb = XB + (size_t)(k & 1) * 3 * NMX, *Rk = Cb + NMX, *Rp = Rk + NMX;
            // the pivot column
            if (tj == k % 16) {
#pragma unroll
                for (int ra = 0; ra < R; ra++) {
                    double v = 0.0;
#pragma unroll
                    for (int rb = 0; rb < R; rb++) v = (rb == k / 16) ? m[ra][rb] : v;
                    const int i = ti + 16 * ra;
                    if (i < nel) Cb[i] = v;
                }
            }
            __syncthreads();
            if (wv == 0) {
                double best = -1.0; int bi = k;
                for (int i = k + lane; i < nel; i += WAVE) {
                    const double v = fabs(Cb[i]);
                    if (v > best) { best = v; bi = i; }
                }
                for (int off = 32; off > 0; off >>= 1) {
                    const double ob2 = __shfl_xor(best, off, WAVE);
                    const int oi = __shfl_xor(bi, off, WAVE);
                    if (ob2 > best || (ob2 == best && oi < bi)) { best = ob2; bi = oi; }
                }
                if (lane == 0) {
                    s_p = bi; perm[k] = bi;
                    piv[k] = Cb[bi]; s_akk = Cb[k];
                    if (piv[k] == 0.0) s_bad = 1;
                }
            }
            __syncthreads();
            const int p = s_p;
            // rows k and p before the swap
            if (ti == k % 16 || ti == p % 16) {
#pragma unroll
                for (int rb = 0; rb < R; rb++) {
                    const int j = tj + 16 * rb;
                    if (j >= nel) continue;
                    double vk = 0.0, vp = 0.0;
#pragma unroll
                    for (int ra = 0; ra < R; ra++) {
                        vk = (ti + 16 * ra == k) ? m[ra][rb] : vk;
                        vp = (ti + 16 * ra == p) ? m[ra][rb] : vp;
                    }
                    if (ti == k % 16) Rk[j] = vk;
                    if (ti == p % 16) Rp[j] = vp;
                }
            }
            __syncthreads();
            const double pinv = 1.0 / piv[k], akk = s_akk;
            double rj[R], rkj[R];
#pragma unroll
            for (int rb = 0; rb < R; rb++) {
                const int j = tj + 16 * rb;
                rj[rb] = j < nel ? ((j == k) ? pinv : Rp[j] * pinv) : 0.0;
                rkj[rb] = j < nel ? Rk[j] : 0.0;
            }
#pragma unroll
            for (int ra = 0; ra < R; ra++) {
                const int i = ti + 16 * ra;
                if (i >= nel) continue;
                if (i == k) {
#pragma unroll
                    for (int rb = 0; rb < R; rb++) m[ra][rb] = rj[rb];
                    continue;
                }
                const double ci = (i == p) ? akk : Cb[i];
#pragma unroll
                for (int rb = 0; rb < R; rb++) {
                    const int j = tj + 16 * rb;
                    double base = (i == p) ? rkj[rb] : m[ra][rb];
                    base = (j == k) ? 0.0 : base;
                    m[ra][rb] = base - rj[rb] * ci;
                }
            }
        }
        // undo the column order of the row swaps: the final inverse's column c is
        // the eliminated matrix's column sig(c), sig = swaps (k, perm[k]) applied
        // in reverse order; each element goes to its place in the LDS matrix
        int *sinv = (int *)XB;  // [NM] (the exchange buffers are free now)
        __syncthreads();
        if (tid == 0) {
            int *sig = sinv + NMX;
            for (int c = 0; c < nel; c++) sig[c] = c;
            for (int k = nel - 1; k >= 0; k--) { const int t = sig[k]; sig[k] = sig[perm[k]]; sig[perm[k]] = t; }
            for (int c = 0; c < nel; c++) sinv[sig[c]] = c;
        }
        __syncthreads();
#pragma unroll
        for (int ra = 0; ra < R; ra++)
#pragma unroll
            for (int rb = 0; rb < R; rb++) {
                const int i = ti + 16 * ra, j = tj + 16 * rb;
                if (i < nel && j < nel) L[i * ld + sinv[j]] = m[ra][rb];
            }
        __syncthreads();
    }
    const int nd = nel > 0 ? nel : 1;  // element stepping of the elimination (no elimination when nel = 0)
    const int i0 = tid / nd, j0 = tid % nd, si = bd / nd, sj = bd % nd;
    for (int k = 0; k < (R > 0 ? 0 : nel); k++) {
        if (wv == 0) {
            double best = -1.0; int bi = k;
            for (int i = k + lane; i < nel; i += WAVE) {
                const double v = fabs(L[i * ld + k]);
                if (v > best) { best = v; bi = i; }
            }
            for (int off = 32; off > 0; off >>= 1) {
                const double ob2 = __shfl_xor(best, off, WAVE);
                const int oi = __shfl_xor(bi, off, WAVE);
                if (ob2 > best || (ob2 == best && oi < bi)) { best = ob2; bi = oi; }
            }
            if (lane == 0) {
                s_p = bi; perm[k] = bi;
                // pivot and old diagonal, read before any thread writes this step
                piv[k] = L[bi * ld + k]; s_akk = L[k * ld + k];
                if (piv[k] == 0.0) s_bad = 1;
            }
        }
        __syncthreads();
        // row swap, scaling of the pivot row and the copy of column k in one pass
        const int p = s_p;
        const double pinv = 1.0 / piv[k], akk = s_akk;
        for (int j = tid; j < nel; j += bd) {
            const double rk = L[k * ld + j], rp = L[p * ld + j];
            L[k * ld + j] = (j == k) ? pinv : rp * pinv;
            if (p != k) L[p * ld + j] = rk;
        }
        // column k after the swap (row k is skipped by the elimination)
        for (int i = tid; i < nel; i += bd) col[i] = (i == p) ? akk : (i == k ? 0.0 : L[i * ld + k]);
        __syncthreads();
        // elements t = tid, tid + bd, ... of the nel x nel block, (i, j) stepped
        // incrementally instead of divided out per element
        for (int i = i0, j = j0; i < nel;) {
            if (i != k) {
                const double base = (j == k) ? 0.0 : L[i * ld + j];
                L[i * ld + j] = base - L[k * ld + j] * col[i];
            }
            i += si; j += sj;
            if (j >= nel) { j -= nel; i++; }
        }
        __syncthreads();
    }
    for (int k = (R > 0 ? -1 : nel - 1); k >= 0; k--) {
        const int p = perm[k];
        if (p != k)
            for (int i = tid; i < nel; i += bd) { const double t = L[i * ld + k]; L[i * ld + k] = L[i * ld + p]; L[i * ld + p] = t; }
        __syncthreads();
    }
    // slmi(j + (i-1) n) = A^-1(j, i): row i of slmi holds column i of the inverse
    double *S = a.slmi + ((size_t)w * 2 + iab) * a.NM * a.NM;
    for (int t = tid; t < nel * nel; t += bd) {
        const int i = t / nel, j = t % nel;
        S[(size_t)i * a.NM + j] = L[j * ld + i];
    }
    if (tid == 0) {
        // mantissa/exponent loops only on finite values: a NaN or Inf pivot
        // would never leave them (it marks the walker instead)
        double det1 = 1.0, det2 = 0.0;
        int nonfinite = 0;
        for (int i = 0; i < nel; i++) {
            if (perm[i] != i) det1 = -det1;
            det1 = piv[i] * det1;
            if (det1 == 0.0) break;
            if (!isfinite(det1)) { nonfinite = 1; break; }
            for (int it = 0; it < 800 && fabs(det1) < 1.0; it++) { det1 = 10.0 * det1; det2 = det2 - 1.0; }
            for (int it = 0; it < 800 && fabs(det1) >= 10.0; it++) { det1 = det1 / 10.0; det2 = det2 + 1.0; }
        }
        a.detiab[2 * w + iab] = nonfinite ? det1 : det1 * pow(10.0, det2);
        if (s_bad) atomicOr(&a.status[w], CHAMP_WALKER_SINGULAR);
        if (nonfinite) atomicOr(&a.status[w], CHAMP_WALKER_NONFINITE);
    }
    if (oc < 5) return;  // orbital values only: the inverse and determinants are the whole result
    // ddx, d2dx2 (determinant.f90:81-88)
    for (int i = tid; i < nel; i += bd) {
        const double *orbi = ob + (size_t)(ish + i) * 5 * NO;
        double d0 = 0, d1 = 0, d2 = 0, dd = 0;
        for (int j = 0; j < nel; j++) {
            const double s = L[j * ld + i];
            const int jo = a.iworbd[ish + j];
            d0 = d0 + s * orbi[NO + jo];
            d1 = d1 + s * orbi[2 * NO + jo];
            d2 = d2 + s * orbi[3 * NO + jo];
            dd = dd + s * orbi[4 * NO + jo];
        }
        double *dx = a.ddx + ((size_t)w * NE + ish + i) * 3;
        dx[0] = d0; dx[1] = d1; dx[2] = d2;
        a.d2dx2[(size_t)w * NE + ish + i] = dd;
    }
    // orbital gradients of this spin's electrons for the moves (orbval dorb)
    for (int t = tid; t < nel * 3 * NO; t += bd) {
        const int i = t / (3 * NO), r = t % (3 * NO);
        a.dorb[((size_t)w * NE + ish + i) * 3 * NO + r] = ob[(size_t)(ish + i) * 5 * NO + NO + r];
    }
}

// ------------------------------------------------------------------------
// gauss_pot (readps_gauss.f90:57-92), l is the 1-based component
__device__ double gauss_pot(const EcpParams &ep, int ict, int l, double r)
{
    double v = 0, rsq = r * r;
    const int nt = ep.nterm[ict][l - 1];
    for (int i = 0; i < nt; i++) {
        const int pw = ep.power[ict][l - 1][i];
        const double ex = ep.expo[ict][l - 1][i];
        double p = (pw != 2) ? powi(r, pw - 2) : 1.0;
        double e = (ex * rsq > 7.0e2) ? 0.0 : ep.coef[ict][l - 1][i] * exp(-ex * rsq);
        v = v + p * e;
    }
    return v;
}

// getvps_gauss + local part of nonloc_pot + active (electron, centre) list
// in the order of nonloc.f90:126-186.  One block per walker, threads over
// electrons, deterministic order by a block-wide scan.
__global__ void __launch_bounds__(256) k_hpsi_vps(HArgs a)
{
    __shared__ int cnt[1024];
    __shared__ double red[16];
    const int w = blockIdx.x, tid = threadIdx.x, bd = blockDim.x;
    const int NE = a.NE, NC = a.NC;
    const EcpParams &ep = *a.ecp;
    const double *rw = a.r_en + (size_t)w * NC * NE;
    double pe = 0;
    for (int e = tid; e < 1024; e += bd) cnt[e] = 0;
    __syncthreads();
    for (int e = tid; e < NE; e += bd) {
        int n = 0;
        for (int c = 0; c < NC; c++) {
            const int ict = a.iwctype[c], lp = ep.lpot[ict];
            double r = rw[(size_t)c * NE + e];
            r = r > 1.0e-10 ? r : 1.0e-10;
            pe += gauss_pot(ep, ict, lp, r) - ep.znuc[ict] * (1.0 / r);
            for (int l = 1; l <= lp - 1; l++)
                if (fabs(gauss_pot(ep, ict, l, r)) > 1.0e-4) { n++; break; }
        }
        cnt[e] = n;
    }
    __syncthreads();
    if (tid == 0) {  // exclusive scan (NE <= 1024)
        int s = 0;
        for (int e = 0; e < NE; e++) { const int v = cnt[e]; cnt[e] = s; s += v; }
        a.pair_count[w] = s;
        a.nxquad[w] = s * a.NQ;
    }
    __syncthreads();
    for (int e = tid; e < NE; e += bd) {
        int off = cnt[e];
        for (int c = 0; c < NC; c++) {
            const int ict = a.iwctype[c], lp = ep.lpot[ict];
            double r = rw[(size_t)c * NE + e];
            r = r > 1.0e-10 ? r : 1.0e-10;
            double v[CHAMP_MPS_L] = {0, 0, 0, 0, 0};
            int act = 0;
            for (int l = 1; l <= lp - 1; l++) {
                v[l - 1] = gauss_pot(ep, ict, l, r);
                if (fabs(v[l - 1]) > 1.0e-4) act = 1;
            }
            if (!act) continue;
            const size_t slot = (size_t)w * a.cap + off;
            a.pair_e[slot] = e;
            a.pair_c[slot] = c;
            for (int l = 0; l < CHAMP_MPS_L; l++) a.pair_vps[slot * CHAMP_MPS_L + l] = v[l];
            off++;
        }
    }
    // rr_en with the centre index innermost for the quadrature kernel
    const double *rrw = a.rr_en + (size_t)w * NC * NE;
    double *rrt = a.rrT + (size_t)w * NE * NC;
    for (int t = tid; t < NE * NC; t += bd) rrt[t] = rrw[(size_t)(t % NC) * NE + t / NC];
    const double pes = block_sum(pe, red);
    if (tid == 0) a.pe_local[w] = pes;
}

__global__ void k_pair_scan(int W, const int *count, int *off, int *total, int div)
{
    // single block: exclusive scan of the per-walker counts of pairs
    // (div = 1) or of groups of div pairs
    __shared__ int part[1024];
    const int tid = threadIdx.x, bd = blockDim.x;
    const int per = (W + bd - 1) / bd;
    int s = 0;
    for (int i = 0; i < per; i++) { const int w = tid * per + i; if (w < W) s += (count[w] + div - 1) / div; }
    part[tid] = s;
    __syncthreads();
    if (tid == 0) { int acc = 0; for (int t = 0; t < bd; t++) { const int v = part[t]; part[t] = acc; acc += v; } total[0] = acc; }
    __syncthreads();
    int acc = part[tid];
    for (int i = 0; i < per; i++) {
        const int w = tid * per + i;
        if (w < W) { off[w] = acc; acc += (count[w] + div - 1) / div; }
    }
    if (tid == 0) off[W] = total[0];
}



// ------------------------------------------------------------------------
// Local energy assembly, pot_local and psi (hpsi.f90:177-246,
// pot_local.f90:3-78, determinant_psit.f90:3-26).  One block per walker.
__global__ void __launch_bounds__(256) k_hpsi_eloc(HArgs a)
{
    __shared__ double red[16];
    const int w = blockIdx.x, tid = threadIdx.x, bd = blockDim.x;
    const int NE = a.NE, NC = a.NC;
    const double *xw = a.xT + (size_t)w * 3 * NE;
    const double *rw = a.r_en + (size_t)w * NC * NE;
    double pe = 0;
    if (a.nloc == 0)
        for (int t = tid; t < NC * NE; t += bd) pe = pe + (-a.ecp->znuc[a.iwctype[t / NE]] / rw[t]);
    const int NP = NE * (NE - 1) / 2;
    for (int p = tid; p < NP; p += bd) {
        int i = (int)((1.0 + sqrt(1.0 + 8.0 * p)) * 0.5);
        while (i * (i - 1) / 2 > p) i--;
        while ((i + 1) * i / 2 <= p) i++;
        const int j = p - i * (i - 1) / 2;
        const double d0 = xw[i] - xw[j], d1 = xw[NE + i] - xw[NE + j], d2 = xw[2 * NE + i] - xw[2 * NE + j];
        double r2 = 0;
        r2 = r2 + d0 * d0; r2 = r2 + d1 * d1; r2 = r2 + d2 * d2;
        pe = pe + 1 / sqrt(r2);
    }
    const double *vj = a.vjT + (size_t)w * 3 * NE;
    double kin[2] = {0, 0}, vv = 0;
    for (int e = tid; e < NE; e += bd) {
        const double *d = a.ddx + ((size_t)w * NE + e) * 3;
        const double v0 = vj[e], v1 = vj[NE + e], v2 = vj[2 * NE + e];
        const double tmp = -HB * (a.d2dx2[(size_t)w * NE + e] + 2.0 * (v0 * d[0] + v1 * d[1] + v2 * d[2]));
        kin[e < a.NUP ? 0 : 1] += tmp;
        vv = vv + (-HB * (v0 * v0 + v1 * v1 + v2 * v2));
    }
    double vp[2] = {0, 0};
    if (a.nloc > 0) {
        const int np = a.pair_count[w];
        for (int p = tid; p < np; p += bd) {
            const size_t slot = (size_t)w * a.cap + p;
            vp[a.pair_e[slot] < a.NUP ? 0 : 1] += a.pair_val[slot];
        }
    }
    const double pes = block_sum(pe, red);
    const double k0 = block_sum(kin[0], red), k1 = block_sum(kin[1], red);
    const double vvs = block_sum(vv, red);
    const double vp0 = block_sum(vp[0], red), vp1 = block_sum(vp[1], red);
    if (tid == 0) {
        const double pe_local = a.pecent + pes + (a.nloc > 0 ? a.pe_local[w] : 0.0);
        const double ekin_other = -HB * a.d2o[w] + vvs;
        const double e_other = ekin_other + pe_local;
        const double eloc0 = k0 + vp0, eloc1 = k1 + vp1;
        const double energy = 0.0 + eloc0 + eloc1 + e_other;
        const double ek = 0.0 + k0 + k1 + ekin_other;
        const double psid = 0.0 + a.detiab[2 * w] * a.detiab[2 * w + 1] * (a.md ? a.md_S[w] : a.cdet0);
        const double psij = a.fsumo[w];
        a.psido[w] = psid; a.psijo[w] = psij; a.eold[w] = energy; a.ekin[w] = ek;
        a.psi2o[w] = 2 * (log(fabs(psid)) + psij);
        a.pe_local[w] = pe_local;
        a.vpsp[2 * w] = vp0; a.vpsp[2 * w + 1] = vp1;
        if (!isfinite(energy) || !isfinite(psij) || psid == 0.0) atomicOr(&a.status[w], CHAMP_WALKER_NONFINITE);
        if (a.accumulate) {
            atomicAdd(&a.esum_dev[1], energy);
            atomicAdd(&a.esum_dev[2], energy * energy);
        }
    }
}

// AO values, orbitals and the LU inverse of both Slater matrices of every
// walker, in chunks of walkers (orbitals of all electrons are large): the
// determinant part of hpsi, also used on its own for the periodic FP64
// refresh of the inverse on the mixed path.
// Register tiles of the Slater inverse: 16 R >= electrons per spin, R <= 8
// (64 doubles per thread); 0 above that (the LDS elimination)
#ifndef CHAMP_SLATER_TILES
#define CHAMP_SLATER_TILES 1
#endif
static int slater_tiles(int nmx) { return (CHAMP_SLATER_TILES && nmx <= 128) ? (nmx + 15) / 16 : 0; }

template <int R>
static int launch_slater_tiles_r(champ_ctx *c, const HArgs &a, size_t lds)
{
    if (lds + SLATER_STATIC_LDS > 64 * 1024)
        HIPCHECK(hipFuncSetAttribute((const void *)k_hpsi_slater<false, R>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                     (int)lds));
    hipLaunchKernelGGL((k_hpsi_slater<false, R>), dim3(2 * a.nw), dim3(256), lds, c->stream, a);
    HIPCHECK(hipGetLastError());
    return CHAMP_OK;
}

static int launch_slater_tiles(champ_ctx *c, const HArgs &a, int rt, size_t lds)
{
    switch (rt) {
    case 1: return launch_slater_tiles_r<1>(c, a, lds);
    case 2: return launch_slater_tiles_r<2>(c, a, lds);
    case 3: return launch_slater_tiles_r<3>(c, a, lds);
    case 4: return launch_slater_tiles_r<4>(c, a, lds);
    case 5: return launch_slater_tiles_r<5>(c, a, lds);
    case 6: return launch_slater_tiles_r<6>(c, a, lds);
    case 7: return launch_slater_tiles_r<7>(c, a, lds);
    default: return launch_slater_tiles_r<8>(c, a, lds);
    }
}

// values_only: orbital values alone, and from them the inverse and the spin
// determinants (the periodic recompute inside a sweep of a single-determinant
// trial function); otherwise also gradients and Laplacians, ddx, d2dx2, the
// orbital gradients of the moves and the multideterminant state.
static int launch_orbitals_slater(champ_ctx *c, HArgs a, bool values_only = false)
{
    int slot;
    a.ocomp = values_only ? 1 : 5;
    for (int w0 = 0; w0 < c->W; w0 += c->hpsi_chunk) {
        a.w0 = w0;
        a.nw = std::min(c->hpsi_chunk, c->W - w0);
        const long long nwave = (long long)a.nw * c->NE;
        tbegin(c, K_HPSI_AO, &slot);
        // FP64 on both paths: these AOs feed the LU inverse that hpsi and the
        // periodic refresh recompute from scratch, and the Slater matrices of a
        // trial function can be ill-conditioned enough to turn FP32 AO rounding
        // into 1e-3 relative errors of the determinant
        if (values_only)
            hipLaunchKernelGGL((k_hpsi_ao<double, 1>), dim3((unsigned)((nwave + 3) / 4)), dim3(256),
                               (size_t)4 * 3 * c->NRI * sizeof(double), c->stream, a);
        else
            hipLaunchKernelGGL((k_hpsi_ao<double, 5>), dim3((unsigned)((nwave + 3) / 4)), dim3(256),
                               (size_t)4 * 3 * c->NRI * sizeof(double), c->stream, a);
        tend(c, slot);
        HIPCHECK(hipGetLastError());
        tbegin(c, K_HPSI_GEMM, &slot);
        int rcg = launch_gemm(c, true, (int)(nwave * a.ocomp), c->NO, c->NB, c->aoall, c->NB, 0, c->coef, c->NB, 0,
                              c->orball, c->NO, 0, 1, c->tune_gemm[1]);
        tend(c, slot);
        if (rcg) return rcg;
        const int nmx = std::max(c->NUP, c->NDN);
        tbegin(c, K_HPSI_SLATER, &slot);
        if (a.lu_ws) {
            hipLaunchKernelGGL((k_hpsi_slater<true, 0>), dim3(2 * a.nw), dim3(256), (size_t)nmx * sizeof(double),
                               c->stream, a);
        } else {
            // dynamic work matrix; with SLATER_STATIC_LDS it fits one block's LDS
            // (otherwise init allocated lu_ws and the branch above is taken)
            const int rt = slater_tiles(nmx);
            if (rt > 0) {
                // register tiles: the work matrix plus the double-buffered pivot column and rows
                const size_t lds = (size_t)(nmx * (nmx + 1) + 6 * nmx) * sizeof(double);
                const int rc = launch_slater_tiles(c, a, rt, lds);
                if (rc) return rc;
            } else {
                const size_t lds = (size_t)(nmx * (nmx + 1) + nmx) * sizeof(double);
                if (lds + SLATER_STATIC_LDS > 64 * 1024)
                    HIPCHECK(hipFuncSetAttribute((const void *)k_hpsi_slater<false, 0>,
                                                 hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
                hipLaunchKernelGGL((k_hpsi_slater<false, 0>), dim3(2 * a.nw), dim3(256), lds, c->stream, a);
            }
        }
        tend(c, slot);
        HIPCHECK(hipGetLastError());
        if (c->md && !values_only) {
            tbegin(c, K_MD, &slot);
            const int rcm = launch_md_hpsi(c, w0, a.nw);
            tend(c, slot);
            if (rcm) return rcm;
        }
    }
    return CHAMP_OK;
}

// Orbital values of every electron of walkers [w0, w0 + nw) into orball
// ([nw][NE][NO]): values-only AOs and one GEMM (the orbital-rotation
// derivatives read them).
int launch_orbital_values(champ_ctx *c, int w0, int nw)
{
    HArgs a = hargs(c);
    a.w0 = w0;
    a.nw = nw;
    a.ocomp = 1;
    const long long nwave = (long long)nw * c->NE;
    hipLaunchKernelGGL((k_hpsi_ao<double, 1>), dim3((unsigned)((nwave + 3) / 4)), dim3(256),
                       (size_t)4 * 3 * c->NRI * sizeof(double), c->stream, a);
    HIPCHECK(hipGetLastError());
    return launch_gemm(c, true, (int)nwave, c->NO, c->NB, c->aoall, c->NB, 0, c->coef, c->NB, 0, c->orball, c->NO, 0, 1,
                       c->tune_gemm[1]);
}

// psido / psi2o from refreshed determinants (the Jastrow part is unchanged)
__global__ void k_refresh_psi(HArgs a)
{
    const int w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= a.W) return;
    const double psid = 0.0 + a.detiab[2 * w] * a.detiab[2 * w + 1] * (a.md ? a.md_S[w] : a.cdet0);
    a.psido[w] = psid;
    a.psi2o[w] = 2 * (log(fabs(psid)) + a.psijo[w]);
}

int launch_slater_refresh(champ_ctx *c)
{
    HArgs a = hargs(c);
    // single determinant: the inverse and determinants from orbital values
    // alone (the orbital gradients of the moves are kept as the moves wrote
    // them); multideterminant: the full recompute, which rebuilds its state
    const int rc = launch_orbitals_slater(c, a, c->md == 0);
    if (rc) return rc;
    hipLaunchKernelGGL(k_refresh_psi, dim3((c->W + 255) / 256), dim3(256), 0, c->stream, a);
    HIPCHECK(hipGetLastError());
    return CHAMP_OK;
}

template <typename T>
static int launch_hpsi_jastrow_t(champ_ctx *c, const HArgs &a, int nordc, int naz)
{
    // up to 8 wavefronts per walker (one LDS row-table region each), so that
    // few walkers are in flight per XCD and their distance tables stay in L2
    const size_t per_wave = nordc > 1 ? (size_t)c->NC * (3 * naz + 8) * sizeof(T) : 8;
    const int nwv = (int)std::max<size_t>(1, std::min<size_t>(8, (160 * 1024 - 1024) / per_wave));
    const size_t lds = nwv * per_wave;
    const void *fn = nordc == 5 ? (const void *)k_hpsi_jastrow<5, T> : nordc == 4 ? (const void *)k_hpsi_jastrow<4, T>
                   : nordc == 3 ? (const void *)k_hpsi_jastrow<3, T> : nordc == 2 ? (const void *)k_hpsi_jastrow<2, T>
                                                                            : (const void *)k_hpsi_jastrow<0, T>;
    if (lds > 160 * 1024) {
        set_error("hpsi: the Jastrow kernel's LDS tables exceed 160 KB for this system");
        return CHAMP_ERR_UNSUPPORTED;
    }
    if (lds > 64 * 1024) HIPCHECK(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    const dim3 block(WAVE * nwv);
    switch (nordc) {
    case 5: hipLaunchKernelGGL((k_hpsi_jastrow<5, T>), dim3(c->W), block, lds, c->stream, a); break;
    case 4: hipLaunchKernelGGL((k_hpsi_jastrow<4, T>), dim3(c->W), block, lds, c->stream, a); break;
    case 3: hipLaunchKernelGGL((k_hpsi_jastrow<3, T>), dim3(c->W), block, lds, c->stream, a); break;
    case 2: hipLaunchKernelGGL((k_hpsi_jastrow<2, T>), dim3(c->W), block, lds, c->stream, a); break;
    default: hipLaunchKernelGGL((k_hpsi_jastrow<0, T>), dim3(c->W), block, lds, c->stream, a); break;
    }
    HIPCHECK(hipGetLastError());
    return CHAMP_OK;
}

// wmat[w][i][m] = sum_j slmi[w][spin][i][j] coef[iworbd(j)][m], per spin;
// multideterminant: wmat[w][i][m] = sum_o V[w][i][o] coef[o][m]
int launch_wmat(champ_ctx *c)
{
    if (!c->wmat) {
        set_error("wmat: no ECP coefficient rows (all-electron system)");
        return CHAMP_ERR_STATE;
    }
    if (c->md)
        return launch_gemm(c, false, c->NE, c->NB, c->NO, c->md_V, c->NO, (long long)c->NE * c->NO, c->coef, c->NB, 0,
                           c->wmat, c->NB, (long long)c->NE * c->NB, c->W, c->tune_gemm[2]);
    for (int iab = 0; iab < 2; iab++) {
        const int nel = iab ? c->NDN : c->NUP, ish = iab ? c->NUP : 0;
        if (nel == 0) continue;
        int rcg = launch_gemm(c, false, nel, c->NB, nel, c->slmi + (size_t)iab * c->NM * c->NM, c->NM,
                              (long long)2 * c->NM * c->NM, c->cocc[iab], c->NB, 0, c->wmat + (size_t)ish * c->NB,
                              c->NB, (long long)c->NE * c->NB, c->W, c->tune_gemm[2]);
        if (rcg) return rcg;
    }
    return CHAMP_OK;
}

// ------------------------------------------------------------------------
int launch_hpsi_impl(champ_ctx *c, int accumulate)
{
    HArgs a = hargs(c);
    a.accumulate = accumulate;
    int slot;
    tbegin(c, K_HPSI_JAS, &slot);
    {
        const int nordc = c->hjas.nordc;
        const int naz[] = {0, 0, een_naz<2>(), een_naz<3>(), een_naz<4>(), een_naz<5>()};
        if (nordc > 5) {
            set_error("hpsi: e-e-n order above 5 is not supported");
            return CHAMP_ERR_UNSUPPORTED;
        }
        const int rcj = c->prec == CHAMP_PREC_MIXED ? launch_hpsi_jastrow_t<float>(c, a, nordc, naz[nordc])
                                                     : launch_hpsi_jastrow_t<double>(c, a, nordc, naz[nordc]);
        if (rcj) return rcj;
    }
    tend(c, slot);
    HIPCHECK(hipGetLastError());
    const int rco = launch_orbitals_slater(c, a);
    if (rco) return rco;
    a.w0 = 0; a.nw = c->W;
    if (c->nloc > 0) {
        // vps and the active-pair list, their offsets, the coefficient rows
        tbegin(c, K_HPSI_ECP, &slot);
        hipLaunchKernelGGL(k_hpsi_vps, dim3(c->W), dim3(256), 0, c->stream, a);
        HIPCHECK(hipGetLastError());
        hipLaunchKernelGGL(k_pair_scan, dim3(1), dim3(1024), 0, c->stream, c->W, c->pair_count, c->pair_off, c->pair_total, 1);
        HIPCHECK(hipGetLastError());
        hipLaunchKernelGGL(k_pair_scan, dim3(1), dim3(1024), 0, c->stream, c->W, c->pair_count, c->grp_off,
                           c->pair_total + 1, NL_GROUP);
        HIPCHECK(hipGetLastError());
        const int rcw = launch_wmat(c);
        tend(c, slot);
        if (rcw) return rcw;
        // timed inside: the psi-ratio kernel and the quadrature kernel apart
        const int rcn = launch_nonloc(c, a);
        if (rcn) return rcn;
    }
    tbegin(c, K_HPSI_ELOC, &slot);
    hipLaunchKernelGGL(k_hpsi_eloc, dim3(c->W), dim3(256), 0, c->stream, a);
    tend(c, slot);
    HIPCHECK(hipGetLastError());
    return CHAMP_OK;
}

int launch_hpsi(champ_ctx *c) { return launch_hpsi_impl(c, 0); }

// basis_fns at arbitrary points (checking entry point): one wave per point.
__global__ void k_basis_points(int npt, const double *pos, HArgs a, double *out)
{
    const int lane = threadIdx.x % WAVE;
    const int p = blockIdx.x * (blockDim.x / WAVE) + threadIdx.x / WAVE;
    if (p >= npt) return;
    for (int m = lane; m < a.NB; m += WAVE) {
        const int c = a.ao_cent[m];
        const double d0 = pos[3 * p] - a.cent[3 * c], d1 = pos[3 * p + 1] - a.cent[3 * c + 1], d2 = pos[3 * p + 2] - a.cent[3 * c + 2];
        double r2 = 0;
        r2 = r2 + d0 * d0; r2 = r2 + d1 * d1; r2 = r2 + d2 * d2;
        double o[5];
        ao_eval<double>(a.tab, a.ao_type[m], a.ao_rad[m], a.ao_slm[m], d0, d1, d2, sqrt(r2), a.rmax_ao[m], 2, o);
        for (int k = 0; k < 5; k++) out[((size_t)p * 5 + k) * a.NB + m] = o[k];
    }
}

int launch_basis_points(champ_ctx *c, int npt, const double *dpos, double *dout)
{
    HArgs a = hargs(c);
    hipLaunchKernelGGL(k_basis_points, dim3((npt + 3) / 4), dim3(256), 0, c->stream, npt, dpos, a, dout);
    HIPCHECK(hipGetLastError());
    return CHAMP_OK;
}

}  // namespace champ
