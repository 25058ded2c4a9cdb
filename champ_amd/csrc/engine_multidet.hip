// engine_multidet.hip -- multideterminant trial functions
//   psi_D = sum_k c_k D_up,k D_dn,k
// on top of the reference determinant's inverse (determinant 0, CHAMP's kref).
//
// Excitation tables (multideterminants_define, set_input_data.f90:46-414):
// per spin, every determinant is the reference with holes h_1..h_r (reference
// slots) filled by particles p_1..p_r (other orbitals).  With
//   aa[j][m] = sum_i Ainv[j][i] phi_m(r_i)          (multideterminant.f90:66-85)
// the determinant ratio is det(alpha), alpha[a][b] = aa[h_a][p_b].  The spin
// determinants are de-duplicated (iwundet); particles are kept in increasing
// orbital order and the parity of the reordering multiplies c_k, as CHAMP
// multiplies cdet by (-1)**itotphase.
//
// Every one-electron quantity comes from one matrix per electron.  Row i of
// every determinant is linear in phi(r_i), so
//   d ln psi_D / d phi_o(r_i) = V[i][o],
//   V[i][o_j] = Ainv[j][i] - sum_v aa[j][v] Z[v][i]  (o_j: reference orbital of slot j)
//   V[i][v]   = Z[v][i]                               (v: a particle orbital)
//   Z = Y Ainv,  Y[v][j] = sum_u W_u C_u[a][b] over the (u, a, b) with h_a = j, p_b = v,
// where C_u is the cofactor matrix of alpha_u and W_u = sum_{k using u} c_k
// D_other,k / S, S = psi_D / (D_up,ref D_dn,ref).  Drift and Laplacian of
// electron i are sum_o V[i][o] grad/lap phi_o(r_i), its ECP quadrature ratio
// sum_o V[i][o] phi_o(x_q).  This is the content of CHAMP's ymat / tildem
// contraction (multideterminant.f90:258-362, hpsi.f90:230-246) with the
// operator applied last.
//
// A move of electron iel (reference ratio R0) changes aa by a rank-one term,
//   aa'[j][v] = aa[j][v] + Ainv[j][iel] / R0 * (phi'_v - sum_j' phi'_{o_j'} aa[j'][v]),
// after which the spin's determinants, S, Y and the new electron's column of
// V follow (multideterminante.f90:3-222).  aa and the spin determinants are
// double-buffered per walker and spin; an accepted move flips the buffer.
// The moves never store cofactors: they need only C_u^T (Ainv column) per
// unique determinant; hpsi writes the weighted cofactors once into scratch.
#include "engine_internal.hpp"

#include <algorithm>
#include <map>
#include <vector>

namespace champ {

constexpr int MD_MAXR = 4;  // excitation rank per spin handled by the closed-form cofactors

struct MdArgs {
    int W, NE, NUP, NM, NO, NV, NU, ndet, w0, iel, mode, lds_aa;  // lds_aa: the per-move kernels read aa from LDS
    int nu[2], nv[2];
    const int *du, *dd;
    const double *cdet;
    const int *rank[2], *hole[2], *part[2], *virt[2], *woff[2], *woth[2], *yoff[2], *yent[2], *kslot[2], *vidx[2];
    const double *wcf[2];
    const int *zoff[2], *zent[2];
    const int *iworbd;
    const double *slmi, *orball, *orbn;
    double *ddx, *d2dx2;
    double *aa, *D, *tb, *G, *S, *Sn, *vcol, *vcoln, *V, *Y, *Z, *Wt, *g;
    int *cur;
};

static MdArgs md_args(champ_ctx *c, int iel)
{
    MdArgs m;
    m.W = c->W; m.NE = c->NE; m.NUP = c->NUP; m.NM = c->NM; m.NO = c->NO; m.NV = c->md_NV; m.NU = c->md_NU;
    m.ndet = c->md_ndet; m.w0 = 0; m.iel = iel; m.mode = 0; m.lds_aa = 0;
    for (int s = 0; s < 2; s++) {
        m.nu[s] = c->md_nu[s]; m.nv[s] = c->md_nv[s];
        m.rank[s] = c->md_rank[s]; m.hole[s] = c->md_hole[s]; m.part[s] = c->md_part[s]; m.virt[s] = c->md_virt[s];
        m.woff[s] = c->md_woff[s]; m.woth[s] = c->md_woth[s]; m.wcf[s] = c->md_wcf[s]; m.yoff[s] = c->md_yoff[s]; m.yent[s] = c->md_yent[s];
        m.kslot[s] = c->md_kslot[s]; m.vidx[s] = c->md_vidx[s];
        m.zoff[s] = c->md_zoff[s]; m.zent[s] = c->md_zent[s];
    }
    m.du = c->md_du; m.dd = c->md_dd; m.cdet = c->md_c; m.iworbd = c->iworbd;
    m.slmi = c->slmi; m.orball = c->orball; m.orbn = c->orbn; m.ddx = c->ddx; m.d2dx2 = c->d2dx2;
    m.aa = c->md_aa; m.D = c->md_D; m.tb = c->md_tb; m.G = c->md_G; m.S = c->md_S; m.Sn = c->md_Sn; m.vcol = c->md_vcol;
    m.vcoln = c->md_vcoln; m.V = c->md_V; m.Y = c->md_Y; m.Z = c->md_Z; m.Wt = c->md_Wt; m.g = c->md_g;
    m.cur = c->md_cur;
    return m;
}

// per-walker, per-spin, per-buffer slices
__device__ __forceinline__ double *md_aa(const MdArgs &m, int w, int s, int b)
{
    return m.aa + (((size_t)w * 2 + s) * 2 + b) * (size_t)m.NM * m.NV;
}
__device__ __forceinline__ double *md_D(const MdArgs &m, int w, int s, int b)
{
    return m.D + (((size_t)w * 2 + s) * 2 + b) * (size_t)m.NU;
}

// det(alpha) and all cofactors C[a][b] (at a * 4 + b) of a 4 x 4 matrix.
// alpha_u of rank r < 4 is padded with an identity block, which leaves its
// determinant and the cofactors of its r x r part unchanged; every index is
// a compile-time constant, so the matrices stay in registers.
__device__ __forceinline__ double det_cof4(const double *al /*[16]*/, double *C /*[16]*/)
{
#define M(i, j) al[(i) * 4 + (j)]
#pragma unroll
    for (int a = 0; a < 4; a++)
#pragma unroll
        for (int b = 0; b < 4; b++) {
            const int r0 = a == 0 ? 1 : 0, r1 = a <= 1 ? 2 : 1, r2 = a <= 2 ? 3 : 2;
            const int c0 = b == 0 ? 1 : 0, c1 = b <= 1 ? 2 : 1, c2 = b <= 2 ? 3 : 2;
            const double mn = M(r0, c0) * (M(r1, c1) * M(r2, c2) - M(r1, c2) * M(r2, c1))
                            - M(r0, c1) * (M(r1, c0) * M(r2, c2) - M(r1, c2) * M(r2, c0))
                            + M(r0, c2) * (M(r1, c0) * M(r2, c1) - M(r1, c1) * M(r2, c0));
            C[a * 4 + b] = ((a + b) & 1) ? -mn : mn;
        }
    return M(0, 0) * C[0] + M(0, 1) * C[1] + M(0, 2) * C[2] + M(0, 3) * C[3];
#undef M
}

__device__ __forceinline__ double det4(const double *al)
{
#define M(i, j) al[(i) * 4 + (j)]
    const double s0 = M(0, 0) * M(1, 1) - M(1, 0) * M(0, 1), s1 = M(0, 0) * M(1, 2) - M(1, 0) * M(0, 2);
    const double s2 = M(0, 0) * M(1, 3) - M(1, 0) * M(0, 3), s3 = M(0, 1) * M(1, 2) - M(1, 1) * M(0, 2);
    const double s4 = M(0, 1) * M(1, 3) - M(1, 1) * M(0, 3), s5 = M(0, 2) * M(1, 3) - M(1, 2) * M(0, 3);
    const double c5 = M(2, 2) * M(3, 3) - M(3, 2) * M(2, 3), c4 = M(2, 1) * M(3, 3) - M(3, 1) * M(2, 3);
    const double c3 = M(2, 1) * M(3, 2) - M(3, 1) * M(2, 2), c2 = M(2, 0) * M(3, 3) - M(3, 0) * M(2, 3);
    const double c1 = M(2, 0) * M(3, 2) - M(3, 0) * M(2, 2), c0 = M(2, 0) * M(3, 1) - M(3, 0) * M(2, 1);
    return s0 * c5 - s1 * c4 + s2 * c3 + s3 * c2 - s4 * c1 + s5 * c0;
#undef M
}

// det(alpha) of the leading R x R block of alpha (row stride R) and, with
// COF, its cofactors C[a][b] at a * R + b.  Every index is a compile-time
// constant, so the matrices stay in registers.
template <int R, bool COF>
__device__ __forceinline__ double det_cof_r(const double *al, double *C)
{
#define M(i, j) al[(i) * R + (j)]
    if constexpr (R == 1) {
        if (COF) C[0] = 1.0;
        return M(0, 0);
    } else if constexpr (R == 2) {
        if (COF) { C[0] = M(1, 1); C[1] = -M(1, 0); C[2] = -M(0, 1); C[3] = M(0, 0); }
        return M(0, 0) * M(1, 1) - M(0, 1) * M(1, 0);
    } else if constexpr (R == 3) {
        double c[9];
#pragma unroll
        for (int a = 0; a < 3; a++)
#pragma unroll
            for (int b = 0; b < 3; b++) {
                const int r0 = a == 0 ? 1 : 0, r1 = a == 2 ? 1 : 2, c0 = b == 0 ? 1 : 0, c1 = b == 2 ? 1 : 2;
                const double mn = M(r0, c0) * M(r1, c1) - M(r0, c1) * M(r1, c0);
                c[a * 3 + b] = ((a + b) & 1) ? -mn : mn;
            }
        if (COF)
#pragma unroll
            for (int t = 0; t < 9; t++) C[t] = c[t];
        return M(0, 0) * c[0] + M(0, 1) * c[1] + M(0, 2) * c[2];
    } else {
        if (COF) return det_cof4(al, C);
        return det4(al);
    }
#undef M
}

// alpha_u (R x R, R = rank of u) gathered from aa
template <int R>
__device__ __forceinline__ void md_alpha_r(const MdArgs &m, int s, int u, const double *aa, double *al)
{
    const int *h = m.hole[s] + 4 * u, *p = m.part[s] + 4 * u;
#pragma unroll
    for (int a = 0; a < R; a++)
#pragma unroll
        for (int b = 0; b < R; b++) al[a * R + b] = aa[(size_t)h[a] * m.NV + p[b]];
}

// determinant of unique determinant u of rank R
template <int R>
__device__ __forceinline__ double md_det_u(const MdArgs &m, int s, int u, const double *aa)
{
    double al[R * R];
    md_alpha_r<R>(m, s, u, aa, al);
    return det_cof_r<R, false>(al, nullptr);
}

// t_u[b] = sum_a C_u[a][b] Ainv[h_a][i] (b < R) into tbu; Ai: row i of the
// spin's slmi (Ai[j] = Ainv[j][i]).  With D, also det(alpha_u) into *D.
template <int R>
__device__ __forceinline__ void md_tb_u(const MdArgs &m, int s, int u, const double *aa, const double *Ai, double *tbu,
                                        double *D)
{
    const int *h = m.hole[s] + 4 * u;
    double al[R * R], C[R * R], x[R];
    md_alpha_r<R>(m, s, u, aa, al);
    const double d = det_cof_r<R, true>(al, C);
    if (D) *D = d;
#pragma unroll
    for (int a = 0; a < R; a++) x[a] = Ai[h[a]];
#pragma unroll
    for (int b = 0; b < R; b++) {
        double t = 0;
#pragma unroll
        for (int a = 0; a < R; a++) t += C[a * R + b] * x[a];
        tbu[b] = t;
    }
}

// alpha_u[a][b] = aa[h_a][p_b], padded to 4 x 4 with the identity
__device__ __forceinline__ void md_alpha(const MdArgs &m, int s, int u, const double *aa, double *al)
{
    const int r = m.rank[s][u];
    const int *h = m.hole[s] + 4 * u, *p = m.part[s] + 4 * u;
#pragma unroll
    for (int a = 0; a < 4; a++)
#pragma unroll
        for (int b = 0; b < 4; b++)
            al[a * 4 + b] = (a < r && b < r) ? aa[(size_t)h[a] * m.NV + p[b]] : (a == b ? 1.0 : 0.0);
}

// spin determinants of spin s from aa (one thread per unique determinant)
__device__ void md_dets(const MdArgs &m, int s, const double *aa, double *D)
{
    for (int u = threadIdx.x; u < m.nu[s]; u += blockDim.x) {
        double d = 1.0;  // rank 0: the reference itself
        switch (m.rank[s][u]) {
        case 1: d = md_det_u<1>(m, s, u, aa); break;
        case 2: d = md_det_u<2>(m, s, u, aa); break;
        case 3: d = md_det_u<3>(m, s, u, aa); break;
        case 4: d = md_det_u<4>(m, s, u, aa); break;
        default: break;
        }
        D[u] = d;
    }
}

__device__ double md_block_sum(double v, double *red)
{
    const int lane = threadIdx.x % WAVE, wv = threadIdx.x / WAVE, nwv = blockDim.x / WAVE;
    v = wave_sum(v);
    __syncthreads();
    if (lane == 0) red[wv] = v;
    __syncthreads();
    double t = 0;
    for (int i = 0; i < nwv; i++) t += red[i];
    __syncthreads();
    return t;  // on every thread
}

// S = sum_k c_k D_up[du_k] D_dn[dd_k]
__device__ double md_sum(const MdArgs &m, const double *Dup, const double *Ddn, double *red)
{
    double s = 0;
    for (int k = threadIdx.x; k < m.ndet; k += blockDim.x) s += m.cdet[k] * Dup[m.du[k]] * Ddn[m.dd[k]];
    return md_block_sum(s, red);
}

// weights W_u = sum_{k using u} c_k D_other,k / S of spin s (Do: the other
// spin's determinants) into the per-walker scratch Wt.  The table holds c_k
// and the other spin's determinant index of each (u, k) entry, so that one
// gather stands between an entry and its product.
__device__ void md_weights(const MdArgs &m, int w, int s, const double *Do, double S)
{
    double *Wt = m.Wt + (size_t)w * m.NU;
    const double sinv = 1.0 / S;
    const double *cf = m.wcf[s];
    const int *oth = m.woth[s];
    for (int u = threadIdx.x; u < m.nu[s]; u += blockDim.x) {
        double t = 0;
        for (int e = m.woff[s][u]; e < m.woff[s][u + 1]; e++) t += cf[e] * Do[oth[e]];
        Wt[u] = t * sinv;
    }
    __syncthreads();
}

// Y[v][j] = sum W_u C_u[a][b] over (h_a, p_b) = (j, v): the weighted
// cofactors of each unique determinant into the scratch G (one thread per
// determinant), then one gather per cell
__device__ void md_y(const MdArgs &m, int w, int s, const double *aa, int nel)
{
    const double *Wt = m.Wt + (size_t)w * m.NU;
    double *G = m.G + (size_t)w * m.NU * 16;
    for (int u = threadIdx.x; u < m.nu[s]; u += blockDim.x) {
        double al[16], C[16];
        md_alpha(m, s, u, aa, al);
        det_cof4(al, C);
        const double wu = Wt[u];
#pragma unroll
        for (int t = 0; t < 16; t++) G[(size_t)u * 16 + t] = wu * C[t];
    }
    __syncthreads();
    double *Y = m.Y + (size_t)w * m.NV * m.NM;
    for (int t = threadIdx.x; t < m.nv[s] * nel; t += blockDim.x) {
        const int v = t / nel, j = t % nel, cell = v * m.NM + j;
        double y = 0;
        for (int e = m.yoff[s][cell]; e < m.yoff[s][cell + 1]; e++) y += G[m.yent[s][e]];
        Y[cell] = y;
    }
    __syncthreads();
}

// Column i of Z = Y Ainv without forming Y:
//   z[v] = sum_u W_u sum_b [p_b = v] t_u[b],  t_u[b] = sum_a C_u[a][b] Ainv[h_a][i],
// in two passes.  md_tb: C_u from aa in registers, t_u into the per-walker
// scratch tb (and, with D, the determinants of aa as they come out of the
// same cofactors).  md_zgather: one gather per particle v.  A: the spin's
// slmi (A[i][j] = Ainv[j][i]).
__device__ void md_tb(const MdArgs &m, int w, int s, const double *aa, const double *A, int i, double *D)
{
    double *tb = m.tb + (size_t)w * m.NU * 4;
    const double *Ai = A + (size_t)i * m.NM;
    for (int u = threadIdx.x; u < m.nu[s]; u += blockDim.x) {
        double *tbu = tb + (size_t)u * 4, *Du = D ? D + u : nullptr;
        switch (m.rank[s][u]) {
        case 1: md_tb_u<1>(m, s, u, aa, Ai, tbu, Du); break;
        case 2: md_tb_u<2>(m, s, u, aa, Ai, tbu, Du); break;
        case 3: md_tb_u<3>(m, s, u, aa, Ai, tbu, Du); break;
        case 4: md_tb_u<4>(m, s, u, aa, Ai, tbu, Du); break;
        default: if (Du) *Du = 1.0; break;  // the reference: no particles
        }
    }
    __syncthreads();
}

__device__ void md_zgather(const MdArgs &m, int w, int s, double scale, double *z)
{
    // one wavefront per particle v, lanes over its (u, b) entries (about
    // 1.5 ndet / nv of them), then a fixed-order wave reduction
    const double *tb = m.tb + (size_t)w * m.NU * 4;
    const double *Wt = m.Wt + (size_t)w * m.NU;
    const int lane = threadIdx.x % WAVE, wv = threadIdx.x / WAVE, nwv = blockDim.x / WAVE;
    for (int v = wv; v < m.nv[s]; v += nwv) {
        double x = 0;
        for (int e = m.zoff[s][v] + lane; e < m.zoff[s][v + 1]; e += WAVE) {
            const int ent = m.zent[s][e];
            x += Wt[ent >> 2] * tb[ent];
        }
        x = wave_sum(x);
        if (lane == 0) z[v] = x * scale;
    }
    __syncthreads();
}

// ------------------------------------------------------------------------
// hpsi: aa, spin determinants, S and V of every electron, then ddx / d2dx2
// from V; one block per walker of the current chunk (orbitals in orball).
__global__ void __launch_bounds__(256) k_md_hpsi(MdArgs m)
{
    __shared__ double red[8];
    const int wc = blockIdx.x, w = m.w0 + wc, tid = threadIdx.x, bd = blockDim.x;
    const int NE = m.NE, NO = m.NO, NV = m.NV, NM = m.NM;
    const double *ob = m.orball + (size_t)wc * NE * 5 * NO;
    for (int s = 0; s < 2; s++) {
        const int nel = s ? NE - m.NUP : m.NUP, ish = s ? m.NUP : 0, b = m.cur[2 * w + s];
        const double *A = m.slmi + ((size_t)w * 2 + s) * NM * NM;  // A[i][j] = Ainv[j][i]
        double *aa = md_aa(m, w, s, b);
        for (int t = tid; t < nel * m.nv[s]; t += bd) {
            const int j = t / m.nv[s], v = t % m.nv[s], o = m.virt[s][v];
            double x = 0;
            for (int i = 0; i < nel; i++) x += A[(size_t)i * NM + j] * ob[(size_t)(ish + i) * 5 * NO + o];
            aa[(size_t)j * NV + v] = x;
        }
    }
    __syncthreads();
    for (int s = 0; s < 2; s++) {
        const int b = m.cur[2 * w + s];
        md_dets(m, s, md_aa(m, w, s, b), md_D(m, w, s, b));
    }
    __syncthreads();
    const double *D0 = md_D(m, w, 0, m.cur[2 * w]), *D1 = md_D(m, w, 1, m.cur[2 * w + 1]);
    const double S = md_sum(m, D0, D1, red);
    if (tid == 0) m.S[w] = S;
    double *Zw = m.Z + (size_t)w * NV * NM;
    const double *Yw = m.Y + (size_t)w * NV * NM;
    for (int s = 0; s < 2; s++) {
        const int nel = s ? NE - m.NUP : m.NUP, ish = s ? m.NUP : 0, b = m.cur[2 * w + s], nv = m.nv[s];
        const double *A = m.slmi + ((size_t)w * 2 + s) * NM * NM;
        const double *aa = md_aa(m, w, s, b);
        md_weights(m, w, s, s ? D0 : D1, S);
        md_y(m, w, s, aa, nel);
        for (int t = tid; t < nv * nel; t += bd) {  // Z[v][i] = sum_j Y[v][j] Ainv[j][i]
            const int v = t / nel, i = t % nel;
            double z = 0;
            for (int j = 0; j < nel; j++) z += Yw[(size_t)v * NM + j] * A[(size_t)i * NM + j];
            Zw[(size_t)v * NM + i] = z;
        }
        __syncthreads();
        double *V = m.V + ((size_t)w * NE + ish) * NO;
        for (int t = tid; t < nel * NO; t += bd) {
            const int i = t / NO, o = t % NO, j = m.kslot[s][o], v = m.vidx[s][o];
            double x = 0;
            if (j >= 0) {
                double c = 0;
                for (int vv = 0; vv < nv; vv++) c += aa[(size_t)j * NV + vv] * Zw[(size_t)vv * NM + i];
                x = A[(size_t)i * NM + j] - c;
            } else if (v >= 0) {
                x = Zw[(size_t)v * NM + i];
            }
            V[(size_t)i * NO + o] = x;
        }
        __syncthreads();
        for (int i = tid; i < nel; i += bd) {
            const double *orbi = ob + (size_t)(ish + i) * 5 * NO, *Vi = V + (size_t)i * NO;
            double d0 = 0, d1 = 0, d2 = 0, dd = 0;
            for (int o = 0; o < NO; o++) {
                const double x = Vi[o];
                d0 += x * orbi[NO + o]; d1 += x * orbi[2 * NO + o]; d2 += x * orbi[3 * NO + o]; dd += x * orbi[4 * NO + o];
            }
            double *dx = m.ddx + ((size_t)w * NE + ish + i) * 3;
            dx[0] = d0; dx[1] = d1; dx[2] = d2;
            m.d2dx2[(size_t)w * NE + ish + i] = dd;
        }
        __syncthreads();  // Y / Z scratch reused by the other spin
    }
}

// ------------------------------------------------------------------------
// Column iel of V at the current configuration (drift at the old position).
__global__ void __launch_bounds__(256) k_md_pre(MdArgs m)
{
    const int w = blockIdx.x, tid = threadIdx.x, bd = blockDim.x;
    const int NE = m.NE, NO = m.NO, NV = m.NV, NM = m.NM, iel = m.iel;
    const int s = iel < m.NUP ? 0 : 1, nel = s ? NE - m.NUP : m.NUP, ish = s ? m.NUP : 0, i = iel - ish;
    const int b = m.cur[2 * w + s], bo = m.cur[2 * w + 1 - s], nv = m.nv[s];
    HIP_DYNAMIC_SHARED(double, sAA)
    const double *A = m.slmi + ((size_t)w * 2 + s) * NM * NM;
    const double *aa = md_aa(m, w, s, b);
    if (m.lds_aa) {  // the nel rows of aa, gathered below once per unique determinant
        for (int t = tid; t < nel * NV; t += bd) sAA[t] = aa[t];
        aa = sAA;  // md_weights ends with a barrier
    }
    md_weights(m, w, s, md_D(m, w, 1 - s, bo), m.S[w]);
    double *z = m.Z + (size_t)w * NV * NM;
    md_tb(m, w, s, aa, A, i, nullptr);
    md_zgather(m, w, s, 1.0, z);
    double *vc = m.vcol + (size_t)w * NO;
    for (int o = tid; o < NO; o += bd) {
        const int j = m.kslot[s][o], v = m.vidx[s][o];
        double x = 0;
        if (j >= 0) {
            double c = 0;
            for (int vv = 0; vv < nv; vv++) c += aa[(size_t)j * NV + vv] * z[vv];
            x = A[(size_t)i * NM + j] - c;
        } else if (v >= 0) {
            x = z[v];
        }
        vc[o] = x;
    }
}

// ------------------------------------------------------------------------
// Electron iel proposed (orbitals orbn at the new position): aa', the spin's
// determinants and cofactors into the spare buffer, S' and column iel of the
// new V (drift at the new position).  The reference ratio R0 = 0 leaves S' = 0
// (psi_D' = 0: the move is refused and the walker flagged by the ratio kernel).
__global__ void __launch_bounds__(256) k_md_post(MdArgs m)
{
    __shared__ double red[8];
    const int w = blockIdx.x, tid = threadIdx.x, bd = blockDim.x;
    const int NE = m.NE, NO = m.NO, NV = m.NV, NM = m.NM, iel = m.iel;
    const int s = iel < m.NUP ? 0 : 1, nel = s ? NE - m.NUP : m.NUP, ish = s ? m.NUP : 0, i = iel - ish;
    const int b = m.cur[2 * w + s], bn = b ^ 1, bo = m.cur[2 * w + 1 - s], nv = m.nv[s];
    const double *A = m.slmi + ((size_t)w * 2 + s) * NM * NM;
    const double *on = m.orbn + (size_t)w * 4 * NO;
    const int *ko = m.iworbd + ish;
    double r0 = 0;
    for (int j = tid; j < nel; j += bd) r0 += A[(size_t)i * NM + j] * on[ko[j]];
    const double R0 = md_block_sum(r0, red);
    double *vcn = m.vcoln + (size_t)w * NO;
    if (R0 == 0.0) {
        for (int o = tid; o < NO; o += bd) vcn[o] = 0.0;
        if (tid == 0) m.Sn[w] = 0.0;
        return;
    }
    const double *aa = md_aa(m, w, s, b);
    double *aan = md_aa(m, w, s, bn);
    double *gw = m.g + (size_t)w * NV;
    for (int v = tid; v < nv; v += bd) {
        double x = on[m.virt[s][v]];
        for (int j = 0; j < nel; j++) x -= on[ko[j]] * aa[(size_t)j * NV + v];
        gw[v] = x;
    }
    __syncthreads();
    const double rinv = 1.0 / R0;
    // elements t = tid, tid + bd, ... of the nel x nv block, (j, v) stepped
    // incrementally (a spin without particle orbitals has no block); with
    // lds_aa also into LDS, where the gathers below read it
    HIP_DYNAMIC_SHARED(double, sAA)
    if (nv > 0) {
        for (int j = tid / nv, v = tid % nv; j < nel;) {
            const double x = aa[(size_t)j * NV + v] + A[(size_t)i * NM + j] * rinv * gw[v];
            aan[(size_t)j * NV + v] = x;
            if (m.lds_aa) sAA[(size_t)j * NV + v] = x;
            j += bd / nv; v += bd % nv;
            if (v >= nv) { v -= nv; j++; }
        }
    }
    __syncthreads();
    const double *aar = m.lds_aa ? sAA : aan;
    // the spin's determinants and t_u from the same cofactors of aa'
    md_tb(m, w, s, aar, A, i, md_D(m, w, s, bn));
    const double *Dn = md_D(m, w, s, bn), *Do = md_D(m, w, 1 - s, bo);
    const double Sn = s == 0 ? md_sum(m, Dn, Do, red) : md_sum(m, Do, Dn, red);
    if (tid == 0) m.Sn[w] = Sn;
    if (Sn == 0.0) {
        for (int o = tid; o < NO; o += bd) vcn[o] = 0.0;
        return;
    }
    md_weights(m, w, s, Do, Sn);
    double *z = m.Z + (size_t)w * NV * NM;
    md_zgather(m, w, s, rinv, z);
    for (int o = tid; o < NO; o += bd) {
        const int j = m.kslot[s][o], v = m.vidx[s][o];
        double x = 0;
        if (j >= 0) {
            double c = 0;
            for (int vv = 0; vv < nv; vv++) c += aar[(size_t)j * NV + vv] * z[vv];
            x = A[(size_t)i * NM + j] * rinv - c;
        } else if (v >= 0) {
            x = z[v];
        }
        vcn[o] = x;
    }
}

// ------------------------------------------------------------------------
// launchers
int launch_md_hpsi(champ_ctx *c, int w0, int nw)
{
    if (!c->md) return CHAMP_OK;
    MdArgs m = md_args(c, 0);
    m.w0 = w0;
    hipLaunchKernelGGL(k_md_hpsi, dim3(nw), dim3(256), 0, c->stream, m);
    HIPCHECK(hipGetLastError());
    return CHAMP_OK;
}

// The per-move kernels stage the moved spin's aa (nel x NV doubles) in LDS
// when it fits in 64 KiB (two or more blocks per CU); else they read it from HBM.
static MdArgs md_move_args(champ_ctx *c, int iel, size_t *lds)
{
    MdArgs m = md_args(c, iel);
    const size_t bytes = (size_t)c->NM * c->md_NV * sizeof(double);
    m.lds_aa = bytes > 0 && bytes <= 64 * 1024;
    *lds = m.lds_aa ? bytes : 0;
    return m;
}

int launch_md_pre(champ_ctx *c, int iel)
{
    if (!c->md) return CHAMP_OK;
    size_t lds;
    const MdArgs m = md_move_args(c, iel, &lds);
    hipLaunchKernelGGL(k_md_pre, dim3(c->W), dim3(256), lds, c->stream, m);
    HIPCHECK(hipGetLastError());
    return CHAMP_OK;
}

int launch_md_post(champ_ctx *c, int iel)
{
    if (!c->md) return CHAMP_OK;
    size_t lds;
    const MdArgs m = md_move_args(c, iel, &lds);
    hipLaunchKernelGGL(k_md_post, dim3(c->W), dim3(256), lds, c->stream, m);
    HIPCHECK(hipGetLastError());
    return CHAMP_OK;
}

// ------------------------------------------------------------------------
// Excitation tables (multideterminants_define, set_input_data.f90:46-414) and
// the per-walker multideterminant state.
template <typename T>
static int md_upload(T **p, const std::vector<T> &h)
{
    const size_t n = std::max<size_t>(h.size(), 1);
    hipError_t e = hipMalloc((void **)p, n * sizeof(T));
    if (e != hipSuccess) { set_error(std::string("hipMalloc: ") + hipGetErrorString(e)); return CHAMP_ERR_NOMEM; }
    if (!h.empty()) HIPCHECK(hipMemcpy(*p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
    return CHAMP_OK;
}
template <typename T>
static int md_zeros(T **p, size_t n)
{
    hipError_t e = hipMalloc((void **)p, std::max<size_t>(n, 1) * sizeof(T));
    if (e != hipSuccess) { set_error(std::string("hipMalloc: ") + hipGetErrorString(e)); return CHAMP_ERR_NOMEM; }
    HIPCHECK(hipMemset(*p, 0, std::max<size_t>(n, 1) * sizeof(T)));
    return CHAMP_OK;
}

int md_init(champ_ctx *c, const champ_system_t *s)
{
    c->md = s->ndet > 1;
    if (!c->md) return CHAMP_OK;
    const int NE = c->NE, NO = c->NO, ndet = s->ndet;
    c->md_ndet = ndet;
    std::vector<double> cd(ndet), ph(ndet, 1.0);
    for (int k = 0; k < ndet; k++) cd[k] = s->cdet[k];
    std::vector<int> uidx[2];
    struct SpinTab {
        std::vector<int> rank, hole, part_orb, virt, kslot, vidx;
    } T[2];
    for (int sp = 0; sp < 2; sp++) {
        const int nel = sp ? NE - c->NUP : c->NUP, ish = sp ? c->NUP : 0;
        const int *ko = s->iworbd + ish;  // the reference determinant (kref = 1)
        std::vector<int> kslot(NO, -1);
        for (int j = 0; j < nel; j++) {
            if (kslot[ko[j]] >= 0) { set_error("init: an orbital appears twice in the reference determinant"); return CHAMP_ERR_ARG; }
            kslot[ko[j]] = j;
        }
        std::map<std::vector<int>, int> seen;  // (holes, particles) -> unique spin determinant
        {
            std::vector<int> key0;
            seen[key0] = 0;
            T[sp].rank.push_back(0);
            for (int t = 0; t < MD_MAXR; t++) { T[sp].hole.push_back(0); T[sp].part_orb.push_back(0); }
        }
        uidx[sp].resize(ndet);
        for (int k = 0; k < ndet; k++) {
            const int *ok = s->iworbd + (size_t)k * NE + ish;
            std::vector<int> in(NO, 0);
            for (int i = 0; i < nel; i++) {
                if (ok[i] < 0 || ok[i] >= NO) { set_error("init: iworbd outside 0..norb-1"); return CHAMP_ERR_ARG; }
                if (in[ok[i]]) { set_error("init: an orbital appears twice in a determinant"); return CHAMP_ERR_ARG; }
                in[ok[i]] = 1;
            }
            std::vector<int> holes, parts;
            for (int j = 0; j < nel; j++) if (!in[ko[j]]) holes.push_back(j);
            for (int i = 0; i < nel; i++) if (kslot[ok[i]] < 0) parts.push_back(ok[i]);
            std::sort(parts.begin(), parts.end());
            if ((int)holes.size() > MD_MAXR) {
                set_error("init: an excitation of more than 4 orbitals of one spin");
                return CHAMP_ERR_UNSUPPORTED;
            }
            // the reference with the holes filled in order, against determinant k's order
            std::vector<int> sw(ko, ko + nel);
            for (size_t a = 0; a < holes.size(); a++) sw[holes[a]] = parts[a];
            std::vector<int> pos(NO, -1);
            for (int i = 0; i < nel; i++) pos[sw[i]] = i;
            std::vector<int> perm(nel), vis(nel, 0);
            for (int i = 0; i < nel; i++) perm[i] = pos[ok[i]];
            int parity = 0;
            for (int i = 0; i < nel; i++) {
                if (vis[i]) continue;
                int len = 0;
                for (int j = i; !vis[j]; j = perm[j]) { vis[j] = 1; len++; }
                parity ^= (len - 1) & 1;
            }
            if (parity) { cd[k] = -cd[k]; ph[k] = -ph[k]; }
            std::vector<int> key(holes);
            key.insert(key.end(), parts.begin(), parts.end());
            auto it = seen.find(key);
            int u;
            if (it == seen.end()) {
                u = (int)T[sp].rank.size();
                seen[key] = u;
                T[sp].rank.push_back((int)holes.size());
                for (int t = 0; t < MD_MAXR; t++) {
                    T[sp].hole.push_back(t < (int)holes.size() ? holes[t] : 0);
                    T[sp].part_orb.push_back(t < (int)parts.size() ? parts[t] : 0);
                }
            } else {
                u = it->second;
            }
            uidx[sp][k] = u;
        }
        // unique determinants in order of rank (stable, the reference stays
        // first): the lanes of a wavefront then mostly share one rank, and the
        // per-rank cofactor code of the moves runs without divergence
        {
            const int nu = (int)T[sp].rank.size();
            std::vector<int> order(nu), pos(nu);
            for (int u = 0; u < nu; u++) order[u] = u;
            std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return T[sp].rank[x] < T[sp].rank[y]; });
            std::vector<int> rk(nu), ho((size_t)nu * MD_MAXR), po((size_t)nu * MD_MAXR);
            for (int n = 0; n < nu; n++) {
                const int u = order[n];
                pos[u] = n;
                rk[n] = T[sp].rank[u];
                for (int t = 0; t < MD_MAXR; t++) {
                    ho[(size_t)n * MD_MAXR + t] = T[sp].hole[(size_t)u * MD_MAXR + t];
                    po[(size_t)n * MD_MAXR + t] = T[sp].part_orb[(size_t)u * MD_MAXR + t];
                }
            }
            T[sp].rank.swap(rk); T[sp].hole.swap(ho); T[sp].part_orb.swap(po);
            for (int k = 0; k < ndet; k++) uidx[sp][k] = pos[uidx[sp][k]];
        }
        // virtual orbitals: every particle of this spin
        std::vector<int> isv(NO, 0);
        for (size_t u = 1; u < T[sp].rank.size(); u++)
            for (int b = 0; b < T[sp].rank[u]; b++) isv[T[sp].part_orb[4 * u + b]] = 1;
        T[sp].vidx.assign(NO, -1);
        for (int o = 0; o < NO; o++)
            if (isv[o]) { T[sp].vidx[o] = (int)T[sp].virt.size(); T[sp].virt.push_back(o); }
        T[sp].kslot = kslot;
    }
    c->md_NV = std::max<int>(1, (int)std::max(T[0].virt.size(), T[1].virt.size()));
    c->md_NU = (int)std::max(T[0].rank.size(), T[1].rank.size());
    const int NV = c->md_NV, NM = c->NM, W = c->W;
    for (int sp = 0; sp < 2; sp++) {
        const int nu = (int)T[sp].rank.size(), nv = (int)T[sp].virt.size();
        c->md_nu[sp] = nu; c->md_nv[sp] = nv;
        std::vector<int> part(T[sp].part_orb.size());
        for (size_t t = 0; t < part.size(); t++) part[t] = 0;
        for (int u = 1; u < nu; u++)
            for (int b = 0; b < T[sp].rank[u]; b++) part[4 * u + b] = T[sp].vidx[T[sp].part_orb[4 * u + b]];
        // dets of each unique spin determinant (CSR)
        std::vector<int> woff(nu + 1, 0), wdet;
        for (int k = 0; k < ndet; k++) woff[uidx[sp][k] + 1]++;
        for (int u = 0; u < nu; u++) woff[u + 1] += woff[u];
        wdet.resize(ndet);
        std::vector<int> fill(woff.begin(), woff.end() - 1);
        for (int k = 0; k < ndet; k++) wdet[fill[uidx[sp][k]]++] = k;
        // per entry: c_k and the other spin's unique determinant of det k
        std::vector<double> wcf(ndet);
        std::vector<int> woth(ndet);
        for (int e = 0; e < ndet; e++) { wcf[e] = cd[wdet[e]]; woth[e] = uidx[1 - sp][wdet[e]]; }
        c->md_hwdet[sp] = wdet;
        // cofactor entries gathered into each cell (v, j) of Y (CSR over v * NM + j)
        std::vector<std::vector<int>> cell((size_t)NV * NM);
        for (int u = 1; u < nu; u++)
            for (int a = 0; a < T[sp].rank[u]; a++)
                for (int b = 0; b < T[sp].rank[u]; b++)
                    cell[(size_t)part[4 * u + b] * NM + T[sp].hole[4 * u + a]].push_back(u * 16 + a * 4 + b);
        std::vector<int> yoff((size_t)NV * NM + 1, 0), yent;
        for (size_t t = 0; t < cell.size(); t++) {
            yoff[t + 1] = yoff[t] + (int)cell[t].size();
            yent.insert(yent.end(), cell[t].begin(), cell[t].end());
        }
        // (u, b) with particle p_b = v, per v (CSR): the column gathers of the moves
        std::vector<std::vector<int>> pv(NV);
        for (int u = 1; u < nu; u++)
            for (int b = 0; b < T[sp].rank[u]; b++) pv[part[4 * u + b]].push_back(u * 4 + b);
        std::vector<int> zoff(NV + 1, 0), zent;
        for (int v = 0; v < NV; v++) {
            zoff[v + 1] = zoff[v] + (int)pv[v].size();
            zent.insert(zent.end(), pv[v].begin(), pv[v].end());
        }
        int rc;
        if ((rc = md_upload(&c->md_rank[sp], T[sp].rank))) return rc;
        if ((rc = md_upload(&c->md_hole[sp], T[sp].hole))) return rc;
        if ((rc = md_upload(&c->md_part[sp], part))) return rc;
        if ((rc = md_upload(&c->md_virt[sp], T[sp].virt))) return rc;
        if ((rc = md_upload(&c->md_woff[sp], woff))) return rc;
        if ((rc = md_upload(&c->md_woth[sp], woth))) return rc;
        if ((rc = md_upload(&c->md_wcf[sp], wcf))) return rc;
        if ((rc = md_upload(&c->md_yoff[sp], yoff))) return rc;
        if ((rc = md_upload(&c->md_yent[sp], yent))) return rc;
        if ((rc = md_upload(&c->md_kslot[sp], T[sp].kslot))) return rc;
        if ((rc = md_upload(&c->md_vidx[sp], T[sp].vidx))) return rc;
        if ((rc = md_upload(&c->md_zoff[sp], zoff))) return rc;
        if ((rc = md_upload(&c->md_zent[sp], zent))) return rc;
    }
    int rc;
    if ((rc = md_upload(&c->md_du, uidx[0]))) return rc;
    if ((rc = md_upload(&c->md_dd, uidx[1]))) return rc;
    if ((rc = md_upload(&c->md_c, cd))) return rc;
    if ((rc = md_upload(&c->md_ph, ph))) return rc;
    c->md_hph = ph;
    const size_t NU = c->md_NU;
    if ((rc = md_zeros(&c->md_aa, (size_t)W * 4 * NM * NV))) return rc;
    if ((rc = md_zeros(&c->md_D, (size_t)W * 4 * NU))) return rc;
    if ((rc = md_zeros(&c->md_tb, (size_t)W * NU * 4))) return rc;
    if ((rc = md_zeros(&c->md_G, (size_t)W * NU * 16))) return rc;
    if ((rc = md_zeros(&c->md_S, (size_t)W))) return rc;
    if ((rc = md_zeros(&c->md_Sn, (size_t)W))) return rc;
    if ((rc = md_zeros(&c->md_cur, (size_t)W * 2))) return rc;
    if ((rc = md_zeros(&c->md_vcol, (size_t)W * NO))) return rc;
    if ((rc = md_zeros(&c->md_vcoln, (size_t)W * NO))) return rc;
    if ((rc = md_zeros(&c->md_V, (size_t)W * NE * NO))) return rc;
    if ((rc = md_zeros(&c->md_Y, (size_t)W * NV * NM))) return rc;
    if ((rc = md_zeros(&c->md_Z, (size_t)W * NV * NM))) return rc;
    if ((rc = md_zeros(&c->md_Wt, (size_t)W * NU))) return rc;
    if ((rc = md_zeros(&c->md_g, (size_t)W * NV))) return rc;
    return CHAMP_OK;
}

// New determinant coefficients: the reordering phases and the w-matrix
// order of init stay, only the coefficient tables are rewritten.
int md_set_cdet(champ_ctx *c, const double *cdet)
{
    const int ndet = c->md_ndet;
    std::vector<double> cd(ndet);
    for (int k = 0; k < ndet; k++) cd[k] = cdet[k] * c->md_hph[k];
    std::vector<double> wcf[2];
    for (int sp = 0; sp < 2; sp++) {
        wcf[sp].resize(ndet);
        for (int e = 0; e < ndet; e++) wcf[sp][e] = cd[c->md_hwdet[sp][e]];
    }
    HIPCHECK(hipMemcpyAsync(c->md_c, cd.data(), ndet * sizeof(double), hipMemcpyHostToDevice, c->stream));
    for (int sp = 0; sp < 2; sp++)
        HIPCHECK(hipMemcpyAsync(c->md_wcf[sp], wcf[sp].data(), ndet * sizeof(double), hipMemcpyHostToDevice,
                                c->stream));
    HIPCHECK(hipStreamSynchronize(c->stream));
    return CHAMP_OK;
}

void md_free(champ_ctx *c)
{
    void *ps[] = {c->md_du, c->md_dd, c->md_c, c->md_ph, c->md_aa, c->md_D, c->md_tb, c->md_G, c->md_S, c->md_Sn, c->md_cur,
                  c->md_vcol, c->md_vcoln, c->md_V, c->md_Y, c->md_Z, c->md_Wt, c->md_g};
    for (void *p : ps) if (p) hipFree(p);
    for (int sp = 0; sp < 2; sp++) {
        void *qs[] = {c->md_rank[sp], c->md_hole[sp], c->md_part[sp], c->md_virt[sp], c->md_woff[sp], c->md_woth[sp],
                      c->md_wcf[sp],
                      c->md_yoff[sp], c->md_yent[sp], c->md_kslot[sp], c->md_vidx[sp], c->md_zoff[sp],
                      c->md_zent[sp]};
        for (void *q : qs) if (q) hipFree(q);
    }
}

}  // namespace champ
