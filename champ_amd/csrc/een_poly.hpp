// een_poly.hpp -- the electron-electron-nucleus part of the ijas=4 Jastrow
// (jastrow4.f90:203-262, nonlpsi.f90:4-86) as a polynomial with its structure
// fixed at compile time by the order nordc.
//
// With switched scaled distances u (electron-electron), x and y (the two
// electrons to one centre) the term sum is
//     P(u, x, y) = sum_t c_t u^k (x^l + y^l) (x y)^m
//                = sum_{k,n} A_kn(x) u^k y^n,
// where term t adds c_t x^(l+m) to A[k][m] and c_t x^m to A[k][l+m] (for l = 0
// both land on A[k][m], giving the factor 2 of ss(0) = 2).  For one electron
// and one centre the tables A, dA/dx and d2A/dx2 are built once; every partner
// electron then costs a Horner evaluation over the non-zero pattern.  All
// loop bounds and indices below are compile-time constants after unrolling.
#pragma once
#include "device_common.hpp"

namespace champ {

template <int N>
struct EenTable {
    int n;                    // number of terms (13 at N = 4)
    int k[40], l[40], m[40];  // (k, l, m) in the order of jastrow4.f90:203-213
    int idx[N + 1][N + 1];    // packed position of A[k][n], -1 where zero
    int naz;                  // number of non-zero A entries (12 at N = 4)
    int nmax[N + 1];          // highest non-zero power of y in row k, -1 if none
    int kmax;                 // highest power of u, -1 if no terms
    constexpr EenTable() : n(0), k{}, l{}, m{}, idx{}, naz(0), nmax{}, kmax(-1)
    {
        bool nz[N + 1][N + 1] = {};
        for (int nn = 2; nn <= N; nn++)
            for (int kk = nn - 1; kk >= 0; kk--) {
                const int lhi = kk == 0 ? nn - kk - 2 : nn - kk;
                for (int ll = lhi; ll >= 0; ll--) {
                    const int mm = (nn - kk - ll) / 2;
                    if (2 * mm == nn - kk - ll) {
                        k[n] = kk; l[n] = ll; m[n] = mm; n++;
                        nz[kk][mm] = true;
                        nz[kk][ll + mm] = true;
                    }
                }
            }
        for (int kk = 0; kk <= N; kk++) {
            nmax[kk] = -1;
            for (int nn = 0; nn <= N; nn++) {
                idx[kk][nn] = nz[kk][nn] ? naz++ : -1;
                if (nz[kk][nn]) nmax[kk] = nn;
            }
            if (nmax[kk] >= 0) kmax = kk;
        }
    }
};

// Switched scaled distance 1 - kap R(r) = exp(-kap r) of isc = 2
// (switch_scale2, scale_dist.f90:314-331).  In double the reference's form
// 1 - kap R; in float that form cancels catastrophically far from a centre
// (R -> 1/kap), so the float path evaluates exp(-kap r) directly.
// kap is taken in double: rounding it to float would shift every exponent the
// same way and bias sums over many pairs.
template <typename T>
__device__ __forceinline__ T switched_from(double kap, T r, T R)
{
    if constexpr (sizeof(T) == sizeof(float)) return exp((float)(-kap * (double)r));
    else return T(1) - (T)kap * R;
}

// ds = s - 1/kap = -x / kap for the float form of pade_minus_asym, with 1/kap
// applied in double (see switched_from)
template <typename T>
__device__ __forceinline__ T asym_offset(double asymp_r, T x)
{
    return (T)(-(double)x * asymp_r);
}

// Pade + polynomial of a scaled distance s minus its asymptote (s -> U = 1/kap):
//   scale0 p0 s/(1 + p1 s) + sum_{o>=2} p_o s^o  -  (same at s = U),
// with ds = s - U passed in exactly (-exp(-kap r)/kap).  Written as
// ds [scale0 p0/((1 + p1 s)(1 + p1 U)) + sum_o p_o g_o], g_o = s g_(o-1) + U^(o-1),
// so nothing cancels when s is near the asymptote: the form the float path
// uses for psibnl / psianl (nonlpsi.f90:88-133) minus asymp_jasb / asymp_jasa.
template <typename T>
__device__ __forceinline__ T pade_minus_asym(const double *p, int nord, T scale0, T s, T ds, T U)
{
    T v = scale0 * (T)p[0] / ((T(1) + (T)p[1] * s) * (T(1) + (T)p[1] * U));
    T g = T(1), Up = T(1);
    for (int o = 2; o <= nord; o++) {
        Up = Up * U;
        g = s * g + Up;
        v = v + (T)p[o] * g;
    }
    return ds * v;
}

template <int N>
constexpr int een_naz() { return EenTable<N>{}.naz > 0 ? EenTable<N>{}.naz : 1; }

// Coefficient tables of one (electron, centre) at switched distance x:
// tab[d * naz + idx] = d^d A / dx^d for d = 0..D.  c = parameters of the
// centre's type in term order.
template <int N, int D, typename T = double>
__device__ __forceinline__ void een_tables(const double *c, T x, T *tab)
{
    constexpr EenTable<N> ET{};
    constexpr int NA = een_naz<N>();
    T xp[N + 1];
    xp[0] = T(1);
#pragma unroll
    for (int p = 1; p <= N; p++) xp[p] = p == 1 ? x : xp[1] * xp[p - 1];
    // one entry at a time: sum over the (term, half) pairs that land on it
#pragma unroll
    for (int d = 0; d <= D; d++)
#pragma unroll
        for (int e = 0; e < NA; e++) {
            T v = T(0);
#pragma unroll
            for (int t = 0; t < ET.n; t++)
#pragma unroll
                for (int h = 0; h < 2; h++) {
                    const int n = h == 0 ? ET.m[t] : ET.l[t] + ET.m[t];
                    const int p = h == 0 ? ET.l[t] + ET.m[t] : ET.m[t];
                    if (ET.idx[ET.k[t]][n] != e || p < d) continue;
                    const T f = d == 0 ? T(1) : (d == 1 ? (T)p : (T)(p * (p - 1)));
                    v += (T)c[t] * f * xp[p - d];
                }
            tab[d * NA + e] = v;
        }
}

// P and its derivatives at (u, y) from the tables of x:
//   D = 0: f;  D = 1: f, fu, fx, fy;  D = 2: also fuu, fxx, fyy, fux, fuy.
template <typename T = double>
struct EenOut {
    T f, fu, fx, fy, fuu, fxx, fyy, fux, fuy;
};

template <int N, int D, typename T = double>
__device__ __forceinline__ EenOut<T> een_eval_tab(const T *tab, T u, T y)
{
    constexpr EenTable<N> ET{};
    constexpr int NA = een_naz<N>();
    EenOut<T> o{0, 0, 0, 0, 0, 0, 0, 0, 0};
    // Horner in u over the row polynomials of y.  The derivative recurrences
    // (d2 = d2 y + d1, d1 = d1 y + p, p = p y + a) start from zero, so their
    // first updates are plain copies: the step counters are compile-time and
    // no multiply by a zero accumulator is emitted.  *2 names hold p''/2.
    T P = 0, Pu = 0, Puu2 = 0, Py = 0, Puy = 0, Pyy2 = 0, Px = 0, Pux = 0, Pxx = 0;
#pragma unroll
    for (int kk = N; kk >= 0; kk--) {
        if (kk > ET.kmax) continue;
        const int su = ET.kmax - kk;  // Horner step in u
        T q = 0, q1 = 0, q2 = 0, r = 0, s = 0;  // row k: value, d/dy, (d2/dy2)/2; dA/dx row; d2A/dx2 row
        const int nm = ET.nmax[kk];
        if (nm >= 0) {
            q = tab[ET.idx[kk][nm]];
            if (D >= 1) r = tab[NA + ET.idx[kk][nm]];
            if (D >= 2) s = tab[2 * NA + ET.idx[kk][nm]];
#pragma unroll
            for (int nn = N - 1; nn >= 0; nn--) {
                if (nn >= nm) continue;
                const int st = nm - 1 - nn;  // Horner step in y
                const int e = ET.idx[kk][nn];
                if (D >= 2 && st >= 1) q2 = st == 1 ? q1 : q2 * y + q1;
                if (D >= 1) q1 = st == 0 ? q : q1 * y + q;
                q = e >= 0 ? q * y + tab[e] : q * y;
                if (D >= 1) r = e >= 0 ? r * y + tab[NA + e] : r * y;
                if (D >= 2) s = e >= 0 ? s * y + tab[2 * NA + e] : s * y;
            }
        }
        if (su == 0) {
            P = q; Py = q1; Pyy2 = q2; Px = r; Pxx = s;
        } else {
            if (D >= 2 && su >= 2) Puu2 = su == 2 ? Pu : Puu2 * u + Pu;
            if (D >= 1) Pu = su == 1 ? P : Pu * u + P;
            P = P * u + q;
            if (D >= 2) Puy = su == 1 ? Py : Puy * u + Py;
            if (D >= 1) Py = Py * u + q1;
            if (D >= 2) Pyy2 = Pyy2 * u + q2;
            if (D >= 2) Pux = su == 1 ? Px : Pux * u + Px;
            if (D >= 1) Px = Px * u + r;
            if (D >= 2) Pxx = Pxx * u + s;
        }
    }
    o.f = P;
    if (D >= 1) { o.fu = Pu; o.fx = Px; o.fy = Py; }
    if (D >= 2) { o.fuu = T(2) * Puu2; o.fyy = T(2) * Pyy2; o.fxx = Pxx; o.fux = Pux; o.fuy = Puy; }
    return o;
}

}  // namespace champ
