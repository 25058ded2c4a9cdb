// device_common.hpp -- device-side building blocks shared by the engine
// kernels: wavefront reductions, xoshiro256**, the grid-3 spline, Cartesian
// monomials, scaled distances and the e-e-n polynomial.  Each routine cites
// the CHAMP routine whose arithmetic it reproduces.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define WAVE 64

namespace champ {

// ---------------------------------------------------------------- reductions
template <typename T>
__device__ __forceinline__ T wave_sum(T v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, WAVE);
    return v;
}

// --------------------------------------------------------------- xoshiro256**
// Blackman & Vigna (2018), bound by CHAMP through random.f90:3-55.
__device__ __forceinline__ uint64_t rotl64(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
__device__ __forceinline__ uint64_t xo_next(uint64_t s[4])
{
    const uint64_t result = rotl64(s[1] * 5, 7) * 9;
    const uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl64(s[3], 45);
    return result;
}
__device__ __forceinline__ double xo_uniform(uint64_t s[4])
{
    return (double)(xo_next(s) >> 11) * 0x1.0p-53;
}
// gauss.f90:3-22: cos(2 pi u1) * sqrt(-2 ln u2)
__device__ __forceinline__ double xo_gauss(uint64_t s[4])
{
    const double pi = 3.141592653589793;  // 4*atan(1) in double
    double g = cos(2 * pi * xo_uniform(s));
    return g * sqrt(-2 * log(xo_uniform(s)));
}

// ------------------------------------------------------------ integer power
// gfortran's real**integer (libgcc __powidf2): binary exponentiation.
template <typename T>
__device__ __forceinline__ T powi(T x, int m)
{
    unsigned n = m < 0 ? (unsigned)(-m) : (unsigned)m;
    T y = (n % 2) ? x : T(1);
    while (n >>= 1) {
        x = x * x;
        if (n % 2) y = y * x;
    }
    return m < 0 ? T(1) / y : y;
}

// Occupancy hint for the compiler's register allocation: at least n waves of
// a kernel per SIMD (n = 1 leaves the choice to the compiler).
#ifndef CHAMP_WAVES_PER_EU
#define CHAMP_WAVES_PER_EU(n) __attribute__((amdgpu_waves_per_eu(n)))
#endif

// four doubles / floats: the accumulator operands of the FP64 / FP32 MFMA instructions
#ifndef CHAMP_DBL4_DEFINED
typedef double dbl4 __attribute__((ext_vector_type(4)));
typedef float flt4 __attribute__((ext_vector_type(4)));
#endif

// ------------------------------------------------------------------- tables
struct SplineTab {
    const double *rwf, *d2rwf, *ce, *ae, *rmaxwf;
    const double *gref;  // [nr][2]: the grid-3 node radii (ref0, ref1) of interval jx
    int nr, nrbas_max;
    double r0, arg, dlogag;
};

// The node radii of interval jx exactly as splfit.f90 forms them; stored per
// interval at set-up so the evaluation below needs no power loop.
__device__ __forceinline__ void grid_refs(double r0, double arg, int jx, double &ref0, double &ref1)
{
    ref0 = r0 * powi(arg, jx - 1) - r0;
    ref1 = (ref0 + r0) * arg - r0;
}

// Grid-3 interval of radius r (splfit.f90:22-24): 0 below the grid, nr past it.
__device__ __forceinline__ int grid_jx(const SplineTab &tab, double r)
{
    const double xr = log((r + tab.r0) / tab.r0) / tab.dlogag + 1.0;
    // compare before converting: a NaN or huge r must never become a table index
    return (xr >= 1.0 && xr < (double)tab.nr) ? (int)xr : (xr >= (double)tab.nr ? tab.nr : 0);
}

// splfit.f90:3-94, grid type 3, with the interval jx of r already known.
// ider 0 -> wfv[0]; ider >= 1 -> wfv[0..2] (the third derivative is only
// needed by forces and is not produced).
template <typename T>
__device__ __forceinline__ void splfit_jx(const SplineTab &tab, int it, int irb, T r, int jx, T wfv[3], int ider)
{
    const size_t base = ((size_t)it * tab.nrbas_max + irb);
    if (jx < 1) {
        const double *ce = tab.ce + base * 5;
        T v = (T)ce[0];
        for (int ii = 2; ii <= 5; ii++) v += (T)ce[ii - 1] * powi(r, ii - 1);
        wfv[0] = v;
        if (ider >= 1) {
            T d1 = 0, d2 = 0;
            for (int ii = 2; ii <= 5; ii++) {
                d1 += (T)((ii - 1) * ce[ii - 1]) * powi(r, ii - 2);
                d2 += (T)((ii - 1) * (ii - 2) * ce[ii - 1]) * powi(r, ii - 3);
            }
            wfv[1] = d1;
            wfv[2] = d2;
        }
    } else if (jx >= tab.nr) {
        const double *ae = tab.ae + base * 2;
        T a1 = (T)ae[1];
        wfv[0] = (T)ae[0] * exp(-a1 * r);
        wfv[1] = -a1 * wfv[0];
        wfv[2] = -a1 * wfv[1];
    } else {
        const double *rwf = tab.rwf + base * tab.nr;
        const double *d2rwf = tab.d2rwf + base * tab.nr;
        const double ref0 = tab.gref[2 * jx], ref1 = tab.gref[2 * jx + 1];
        T delh = (T)(ref1 - ref0);
        T y0 = (T)rwf[jx - 1], y1 = (T)rwf[jx], s0 = (T)d2rwf[jx - 1], s1 = (T)d2rwf[jx];
        const T sixth = T(1) / T(6);
        T delhi = T(1) / delh;
        T bb = (r - (T)ref0) * delhi;
        T aa = ((T)ref1 - r) * delhi;
        T cc = aa * (aa * aa - T(1)) * delh * delh * sixth;
        T dd = bb * (bb * bb - T(1)) * delh * delh * sixth;
        wfv[0] = aa * y0 + bb * y1 + cc * s0 + dd * s1;
        if (ider >= 1) {
            wfv[1] = (y1 - y0) * delhi + (-(T(3) * aa * aa - T(1)) * s0 + (T(3) * bb * bb - T(1)) * s1) * delh * sixth;
            wfv[2] = aa * s0 + bb * s1;
        }
    }
}

template <typename T>
__device__ __forceinline__ void splfit(const SplineTab &tab, int it, int irb, T r, T wfv[3], int ider)
{
    splfit_jx<T>(tab, it, irb, r, grid_jx(tab, (double)r), wfv, ider);
}

// slm.f90:7-200 (ider < 3): monomial y, gradient dy, Laplacian, case by
// case.  The kernels use the branch-free slm_decode / slm_all below, which
// tests/cpu_emu/slm_check.cpp checks against this form for every l.
template <typename T>
__device__ __forceinline__ void slm(int l, T x, T yy, T z, T &y, T dy[3], T &lap)
{
    const T cs = (T)0.28209479177387814, cp = (T)0.48860251190291992;
    const T cd1 = (T)0.630783130505040, cd2 = (T)1.092548430592079;
    const T cf = (T)0.74635266518023078, cf2 = (T)1.66889529453113635, cf3 = (T)2.89061144264055405;
    dy[0] = dy[1] = dy[2] = 0;
    lap = 0;
    switch (l) {
    case 1: y = cs; break;
    case 2: y = cp * x; dy[0] = cp; break;
    case 3: y = cp * yy; dy[1] = cp; break;
    case 4: y = cp * z; dy[2] = cp; break;
    case 5: y = cd1 * x * x; dy[0] = T(2) * cd1 * x; lap = T(2) * cd1; break;
    case 6: y = cd2 * x * yy; dy[0] = cd2 * yy; dy[1] = cd2 * x; break;
    case 7: y = cd2 * x * z; dy[0] = cd2 * z; dy[2] = cd2 * x; break;
    case 8: y = cd1 * yy * yy; dy[1] = T(2) * cd1 * yy; lap = T(2) * cd1; break;
    case 9: y = cd2 * yy * z; dy[1] = cd2 * z; dy[2] = cd2 * yy; break;
    case 10: y = cd1 * z * z; dy[2] = T(2) * cd1 * z; lap = T(2) * cd1; break;
    case 11: y = cf * x * x * x; dy[0] = cf * T(3) * x * x; lap = cf * T(6) * x; break;
    case 12: y = cf2 * x * x * yy; dy[0] = cf2 * T(2) * x * yy; dy[1] = cf2 * x * x; lap = cf2 * T(2) * yy; break;
    case 13: y = cf2 * x * x * z; dy[0] = cf2 * T(2) * x * z; dy[2] = cf2 * x * x; lap = cf2 * T(2) * z; break;
    case 14: y = cf2 * yy * yy * x; dy[1] = cf2 * T(2) * yy * x; dy[0] = cf2 * yy * yy; lap = cf2 * T(2) * x; break;
    case 15: y = cf3 * x * yy * z; dy[0] = cf3 * yy * z; dy[1] = cf3 * x * z; dy[2] = cf3 * x * yy; break;
    case 16: y = cf2 * z * z * x; dy[0] = cf2 * z * z; dy[2] = cf2 * T(2) * z * x; lap = cf2 * T(2) * x; break;
    case 17: y = cf * yy * yy * yy; dy[1] = cf * T(3) * yy * yy; lap = cf * T(6) * yy; break;
    case 18: y = cf2 * yy * yy * z; dy[1] = cf2 * T(2) * yy * z; dy[2] = cf2 * yy * yy; lap = cf2 * T(2) * z; break;
    case 19: y = cf2 * z * z * yy; dy[1] = cf2 * z * z; dy[2] = cf2 * T(2) * z * yy; lap = cf2 * T(2) * yy; break;
    case 20: y = cf * z * z * z; dy[2] = cf * T(3) * z * z; lap = cf * T(6) * z; break;
    default: y = 0; break;
    }
}

// The monomial of slm as x^ex y^ey z^ez times a constant, decoded from l
// without branches (exponents packed 2 bits per l, the constant's class 3
// bits per l, in slm's order), so that lanes holding different l stay in step.
struct SlmMono {
    int ex, ey, ez;
    double c;
};

__device__ __forceinline__ SlmMono slm_decode(int l)
{
    const bool ok = l >= 1 && l <= 20;
    const int k = ok ? l - 1 : 0;
    SlmMono m;
    m.ex = (int)((0x56b01604ull >> (2 * k)) & 3);
    m.ey = (int)((0x1b18418410ull >> (2 * k)) & 3);
    m.ez = (int)((0xe491091040ull >> (2 * k)) & 3);
    const int cls = (int)((0x96cbadb134da248ull >> (3 * k)) & 7);
    const double cs = 0.28209479177387814, cp = 0.48860251190291992;
    const double cd1 = 0.630783130505040, cd2 = 1.092548430592079;
    const double cf = 0.74635266518023078, cf2 = 1.66889529453113635, cf3 = 2.89061144264055405;
    double c = cls == 0 ? cs : cls == 1 ? cp : cls == 2 ? cd1 : cls == 3 ? cd2 : cls == 4 ? cf : cls == 5 ? cf2 : cf3;
    m.c = ok ? c : 0.0;
    return m;
}

template <typename T>
__device__ __forceinline__ T ipow3(T t, int e)
{
    const T t2 = t * t;
    return e == 0 ? T(1) : e == 1 ? t : e == 2 ? t2 : t2 * t;
}

// value of slm's monomial at (x, y, z)
template <typename T>
__device__ __forceinline__ T slm_value(const SlmMono &m, T x, T y, T z)
{
    return (T)m.c * ipow3(x, m.ex) * ipow3(y, m.ey) * ipow3(z, m.ez);
}

// t^e, e t^(e-1) and e (e-1) t^(e-2) for e = 0..3 (zero where the power is negative)
template <typename T>
__device__ __forceinline__ void ipow3_d(T t, int e, T &p0, T &p1, T &p2)
{
    const T t2 = t * t;
    p0 = e == 0 ? T(1) : e == 1 ? t : e == 2 ? t2 : t2 * t;
    p1 = e == 0 ? T(0) : e == 1 ? T(1) : e == 2 ? T(2) * t : T(3) * t2;
    p2 = e < 2 ? T(0) : e == 2 ? T(2) : T(6) * t;
}

// slm's monomial y, gradient dy and Laplacian at (x, yy, z), branch-free
template <typename T>
__device__ __forceinline__ void slm_all(const SlmMono &m, T x, T yy, T z, T &y, T dy[3], T &lap)
{
    T x0, x1, x2, y0, y1, y2, z0, z1, z2;
    ipow3_d(x, m.ex, x0, x1, x2);
    ipow3_d(yy, m.ey, y0, y1, y2);
    ipow3_d(z, m.ez, z0, z1, z2);
    const T c = (T)m.c;
    y = c * x0 * y0 * z0;
    dy[0] = c * x1 * y0 * z0;
    dy[1] = c * x0 * y1 * z0;
    dy[2] = c * x0 * y0 * z1;
    lap = c * (x2 * y0 * z0 + x0 * y2 * z0 + x0 * y0 * z2);
}

// One AO from its radial function R, R', R'' (wfv) and its monomial l at the
// point (rx, ry, rz) = x - centre, |.| = r: phi_combine (basis_fns.f90:217-296).
// Returns value, gradient and (ider == 2) Laplacian.
template <typename T>
__device__ __forceinline__ void ao_combine(int l, T rx, T ry, T rz, T r, const T wfv[3], int ider, T out[5])
{
    T y, dy[3], lap;
    slm_all<T>(slm_decode(l), rx, ry, rz, y, dy, lap);  // lanes of one wave hold different l
    out[0] = y * wfv[0];
    if (ider >= 1) {
        T ri = T(1) / r;
        T xcri[3] = {rx * ri, ry * ri, rz * ri};
        T dum = 0;
        for (int jj = 0; jj < 3; jj++) {
            out[1 + jj] = y * xcri[jj] * wfv[1] + dy[jj] * wfv[0];
            dum = dum + dy[jj] * xcri[jj];
        }
        if (ider >= 2) {
            T d2 = y * wfv[2] + y * T(2) * ri * wfv[1] + lap * wfv[0];
            out[4] = d2 + T(2) * dum * wfv[1];
        }
    }
}

// One AO at one point: basis_fns.f90:72-110 + phi_combine :217-296.
template <typename T>
__device__ __forceinline__ void ao_eval(const SplineTab &tab, int it, int irad, int l, T rx, T ry, T rz, T r,
                                        double rmax, int ider, T out[5])
{
    T wfv[3] = {0, 0, 0};
    if ((double)r <= rmax) splfit<T>(tab, it, irad, r, wfv, ider);
    ao_combine<T>(l, rx, ry, rz, r, wfv, ider, out);
}

// ----------------------------------------------------------------- Jastrow
#define CHAMP_MAXCT 4
#define CHAMP_MAXT 32
#define CHAMP_NORDP 8

// ijas=4, isc=2 parameters (m_jastrow.f90:145-147) + asymptotes (set_scale_dist).
struct JasParams {
    int norda, nordb, nordc, nterms, nspin2b, nocuspb, nctype, pad;
    double kap, asymp_r;
    double a4[CHAMP_MAXCT][CHAMP_NORDP];
    double b[2][CHAMP_NORDP];
    double c[CHAMP_MAXCT][CHAMP_MAXT];
    double ja[CHAMP_MAXCT];
    double jb[2];
    signed char tk[CHAMP_MAXT], tl[CHAMP_MAXT], tm[CHAMP_MAXT];  // (k,l,m) of each e-e-n term
};

// scale_dist2 / switch_scale (scale_dist.f90:159-331), isc = 2
template <typename T>
__device__ __forceinline__ T scale_r(T kap, T r) { return (T(1) - exp(-kap * r)) / kap; }

// Electron-centre distance of electron e of a walker (xw [3][NE]): the one
// expression every kernel that forms r_en uses, so recomputed values equal
// the stored ones exactly.
__device__ __forceinline__ double en_distance(const double *xw, int NE, int e, const double *cent, int c)
{
    const double rv0 = xw[e] - cent[3 * c], rv1 = xw[NE + e] - cent[3 * c + 1], rv2 = xw[2 * NE + e] - cent[3 * c + 2];
    double r2 = 0;
    r2 = r2 + rv0 * rv0; r2 = r2 + rv1 * rv1; r2 = r2 + rv2 * rv2;
    return sqrt(r2);
}

// The partner side of an (electron, centre) pair as the FP32 move kernel uses
// it: the switched scaled distance x = 1 - kap rr and the factor
// -kap x / r of its gradient; (0, 0) at the asymptote rr = 1/kap, so x == 0
// marks it (jastrow4e skips a pair whose two distances are both there).
__device__ __forceinline__ float2 enp_pack(double kap, double asymp_r, double r, double rr)
{
    if (rr == asymp_r) return make_float2(0.0f, 0.0f);
    const double sw = 1.0 - kap * rr;
    return make_float2((float)sw, (float)(-kap * sw / r));
}

// e-e-n sums for one pair and one centre: jastrow4.f90:203-262 (fc, fu, fi,
// fj and, with want2, the second derivatives).  uu/ri/rj are the switched
// scaled distances.  Power tables have zeros at negative indices.
template <typename T, bool WANT2>
struct Een {
    T fc, fu, fi, fj, fuu, fii, fjj, fui, fuj;
};

template <typename T, bool WANT2>
__device__ __forceinline__ Een<T, WANT2> een_eval(const JasParams &jp, const double *c, const T *uu, const T *rri,
                                                  const T *rrj, const T *ss, const T *tt)
{
    Een<T, WANT2> e;
    e.fc = e.fu = e.fi = e.fj = 0;
    e.fuu = e.fii = e.fjj = e.fui = e.fuj = 0;
    for (int ll = 0; ll < jp.nterms; ll++) {
        const int k = jp.tk[ll], l = jp.tl[ll], m = jp.tm[ll];
        const T cl = (T)c[ll];
        const T stv = ss[l] * tt[m];
        e.fc = e.fc + cl * uu[k] * stv;
        e.fu = e.fu + cl * T(k) * uu[k - 1] * stv;
        const T gi = T(l + m) * rri[l + m - 1] * rrj[m] + T(m) * rri[m - 1] * rrj[l + m];
        const T gj = T(l + m) * rrj[l + m - 1] * rri[m] + T(m) * rrj[m - 1] * rri[l + m];
        e.fi = e.fi + cl * uu[k] * gi;
        e.fj = e.fj + cl * uu[k] * gj;
        if (WANT2) {
            e.fuu = e.fuu + cl * T(k * (k - 1)) * uu[k - 2] * stv;
            e.fii = e.fii + cl * uu[k] * (T((l + m) * (l + m - 1)) * rri[l + m - 2] * rrj[m] + T(m * (m - 1)) * rri[m - 2] * rrj[l + m]);
            e.fjj = e.fjj + cl * uu[k] * (T((l + m) * (l + m - 1)) * rrj[l + m - 2] * rri[m] + T(m * (m - 1)) * rrj[m - 2] * rri[l + m]);
            e.fui = e.fui + cl * T(k) * uu[k - 1] * gi;
            e.fuj = e.fuj + cl * T(k) * uu[k - 1] * gj;
        }
    }
    return e;
}

}  // namespace champ
