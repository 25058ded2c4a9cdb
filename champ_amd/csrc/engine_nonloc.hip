// engine_nonloc.hip -- ECP quadrature of hpsi over the active (electron,
// centre) pairs of the whole population: nonloc_pot.f90:3-108 and
// nonloc.f90:5-315 with nonlocd_quad (:462-506) and nonlocj_quad4
// (:656-899), single determinant.
//
// One 256-thread block per pair, quadrature points in blocks of QB.  Per
// point q of a pair (i, ic) the kernel needs
//   psi ratio   = sum_m phi_m(x_q) W[i][m],  W = slmi_i . coef (one GEMM per hpsi)
//   log J ratio = sum_{j!=i} [b(u_qj) + sum_c P(u_qj, R_qc, R_jc)]
//                 + sum_c a(R_qc) - sum_j F[i][j]
// with P the e-e-n polynomial.  For a fixed (q, c) the polynomial is written
// as P = sum_k u^k sum_n A_kn(R_qc) R_jc^n (een_poly.hpp), so a lane that owns
// one centre c keeps A in registers for QB points and streams over electrons
// j: about a dozen FMAs per (q, c, j).
//   phase A  lanes over (j, q): u_qj into LDS, e-e sums b(u_qj)
//   phase B  lanes over (c, j-group): e-e-n and e-n sums
//   phase C  lanes over AOs: the psi ratio
// Reductions go lane -> wave (shuffles) -> per-wave slots in LDS -> thread 0,
// in a fixed order, so results do not depend on timing.
#include "een_poly.hpp"
#include "hpsi_args.hpp"

#include <algorithm>
#include <cmath>

namespace champ {

// psianl (nonlpsi.f90:88-105): e-n term at scaled distance rr; x = 1 - kap rr
// (exp(-kap r)) for the float form, which subtracts the asymptote exactly
template <typename T>
__device__ __forceinline__ T en_term(const JasParams &jp, int it, T rr, T x)
{
    if constexpr (sizeof(T) == sizeof(float))
        return pade_minus_asym<T>(jp.a4[it], jp.norda, T(1), rr, asym_offset<T>(jp.asymp_r, x), (T)jp.asymp_r);
    if (rr == (T)jp.asymp_r) return T(0);
    T psi = (T)jp.a4[it][0] * rr / (T(1) + (T)jp.a4[it][1] * rr) - (T)jp.ja[it];
    T rrip = rr;
    for (int o = 2; o <= jp.norda; o++) { rrip = rrip * rr; psi = psi + (T)jp.a4[it][o] * rrip; }
    return psi;
}

// psibnl (nonlpsi.f90:107-133): e-e term of electrons (i, j) at scaled u;
// x = 1 - kap u (exp(-kap r)) for the float form
template <typename T>
__device__ __forceinline__ T ee_term(const JasParams &jp, int nup, int i, int j, T u, T x)
{
    if (sizeof(T) != sizeof(float) && u == (T)jp.asymp_r) return T(0);
    const int I = j > i ? j : i, J = j > i ? i : j;
    T sspinn = T(1);
    int isb = 0, ipar = 0;
    if (I < nup || J >= nup) {
        if (jp.nspin2b == 2) isb = 1;
        else if (jp.nocuspb == 0) sspinn = T(0.5);
        ipar = 1;
    }
    const double *b = jp.b[isb];
    if constexpr (sizeof(T) == sizeof(float))
        return pade_minus_asym<T>(b, jp.nordb, sspinn, u, asym_offset<T>(jp.asymp_r, x), (T)jp.asymp_r);
    const T fee = (T)b[0] * u / (T(1) + (T)b[1] * u);
    T psi = sspinn * fee - (T)jp.jb[ipar];
    T up = u;
    for (int o = 2; o <= jp.nordb; o++) { up = u * up; psi = psi + (T)b[o] * up; }
    return psi;
}

// lanes of a wave with equal (lane % QB) summed; the result is in lanes 0..QB-1
template <int QB>
__device__ __forceinline__ double strided_sum(double v)
{
#pragma unroll
    for (int off = WAVE / 2; off >= QB; off >>= 1) v += __shfl_xor(v, off, WAVE);
    return v;
}

// Tuning of k_nonloc: quadrature points per block at e-e-n order <= 4 and the
// minimum waves per SIMD asked of the register allocator.
#ifndef CHAMP_NL_QB
#define CHAMP_NL_QB 4
#endif
#ifndef CHAMP_NL_WPE
#define CHAMP_NL_WPE 1
#endif

constexpr int NL_THREADS = 256;
constexpr int NL_WAVES = NL_THREADS / WAVE;

// T: arithmetic type of the Jastrow part (double on the FP64 path, float on
// the mixed path).  The determinant ratio at the points stays double on both
// paths: sum_m phi_m W[i][m] cancels strongly when the inverse has large
// entries, so FP32 AO values would lose its leading digits.  Sums across
// lanes, the per-point assembly and every table in memory are double.
template <int NORDC, int QB, typename T>
__global__ void __launch_bounds__(NL_THREADS) CHAMP_WAVES_PER_EU(CHAMP_NL_WPE) k_nonloc(HArgs a)
{
    HIP_DYNAMIC_SHARED(double, lds)
    const int tid = threadIdx.x, lane = tid % WAVE, wv = tid / WAVE;
    const int NE = a.NE, NC = a.NC, NQ = a.NQ, NB = a.NB, NRI = a.NRI;
    const JasParams &jp = *a.jas;
    const T kap = (T)jp.kap, asymp = (T)jp.asymp_r;
    const double kapd = jp.kap, asympd = jp.asymp_r;
    double *S = lds;                        // [3][QB][NL_WAVES] per-wave sums: e-e, e-e-n + e-n, psi ratio
    double *XQ = S + 3 * QB * NL_WAVES;     // [QB][3] quadrature points
    double *FI = XQ + 3 * QB;               // [NL_WAVES] partial sums of F[i][.]
    double *GEO = FI + NL_WAVES;            // [QB][NC][5] point - centre: x, y, z, r, grid interval
    double *RAD = GEO + (size_t)QB * NC * 5;  // [QB][NRI] radial functions at the points
    T *U = (T *)(RAD + (size_t)QB * NRI);   // [NE][QB] switched scaled e-e distances u(q, j)

    // XCD-aware assignment: workgroups are dealt round-robin to the 8 XCDs,
    // so the blocks b = x, x + 8, ... run on XCD x, which takes one contiguous
    // eighth of the pair list; inside it the blocks stride by their count, so
    // the blocks resident at one time work on neighbouring pairs (the same
    // one or two walkers, whose tables stay in that XCD's L2).
    const int G = gridDim.x, b = blockIdx.x;
    const int nx = (G % 8 == 0) ? 8 : 1, gx = G / nx, xcd = b % nx, j0 = b / nx;
    const int total = a.pair_total[0];
    const int per_x = (total + nx - 1) / nx;
    const int x_lo = xcd * per_x, x_hi = min(total, x_lo + per_x);

    for (int t = x_lo + j0; t < x_hi; t += gx) {
        int lo = 0, hi = a.W - 1;
        while (lo < hi) {  // walker: largest w with off[w] <= t
            const int mid = (lo + hi + 1) / 2;
            if (a.pair_off[mid] <= t) lo = mid; else hi = mid - 1;
        }
        const int w = lo;
        const size_t slot = (size_t)w * a.cap + (t - a.pair_off[w]);
        const int i = a.pair_e[slot], ic = a.pair_c[slot];
        const int lp = a.ecp->lpot[a.iwctype[ic]];
        const double *xw = a.xT + (size_t)w * 3 * NE;
        const double *F = a.fso + ((size_t)w * NE + i) * NE;
        const double *Wm = a.wmat + ((size_t)w * NE + i) * NB;
        const double *rrt = a.rrT + (size_t)w * NE * NC;
        const double cx = a.cent[3 * ic], cy = a.cent[3 * ic + 1], cz = a.cent[3 * ic + 2];
        const double rv0 = xw[i] - cx, rv1 = xw[NE + i] - cy, rv2 = xw[2 * NE + i] - cz;
        const double ren = a.r_en[((size_t)w * NC + ic) * NE + i], ri = 1.0 / ren;
        const double *qx = a.xq + (size_t)w * NQ, *qy = a.yq + (size_t)w * NQ, *qz = a.zq + (size_t)w * NQ;

        // sum_j F[i][j]: the old e-e and e-n values of electron i (F symmetric)
        double fs = 0.0;
        for (int j = tid; j < NE; j += NL_THREADS) fs += F[j];
        fs = wave_sum(fs);
        __syncthreads();  // LDS of the previous pair is free
        if (lane == 0) FI[wv] = fs;

        double pair_sum = 0.0;  // meaningful on thread 0
        for (int q0 = 0; q0 < NQ; q0 += QB) {
            if (tid < 3 * QB) {
                const int qq = tid / 3, k = tid % 3, q = min(q0 + qq, NQ - 1);
                const double *qv = k == 0 ? qx : (k == 1 ? qy : qz);
                XQ[tid] = ren * qv[q] + a.cent[3 * ic + k];
            }
            __syncthreads();

            // ---- phase A: u(q, j) and the e-e sums (lane keeps q = tid % QB)
            {
                const int qq = tid % QB;
                const double px = XQ[3 * qq], py = XQ[3 * qq + 1], pz = XQ[3 * qq + 2];
                T see = T(0);
                for (int j = tid / QB; j < NE; j += NL_THREADS / QB) {
                    const T d0 = (T)(xw[j] - px), d1 = (T)(xw[NE + j] - py), d2 = (T)(xw[2 * NE + j] - pz);
                    T r2 = 0;
                    r2 = r2 + d0 * d0; r2 = r2 + d1 * d1; r2 = r2 + d2 * d2;
                    const T rjq = sqrt(r2);
                    const T u = scale_r<T>(kap, rjq);
                    const T xs = switched_from<T>(kapd, rjq, u);
                    U[(size_t)j * QB + qq] = xs;
                    if (j != i) see += ee_term<T>(jp, a.NUP, i, j, u, xs);
                }
                const double sees = strided_sum<QB>((double)see);
                if (lane < QB) S[(0 * QB + lane) * NL_WAVES + wv] = sees;
            }
            __syncthreads();

            // ---- phase B: e-e-n (lanes own a centre, stream over electrons) and e-n
            T acc[QB];
#pragma unroll
            for (int qq = 0; qq < QB; qq++) acc[qq] = T(0);
            {
                const int CT = NC < NL_THREADS ? NC : NL_THREADS, JG = NL_THREADS / CT;
                for (int c0 = 0; c0 < NC; c0 += CT) {
                    const int c = c0 + tid % CT, jg = tid / CT;
                    if (jg < JG && c < NC) {
                        const int it = a.iwctype[c];
                        const double ecx = a.cent[3 * c], ecy = a.cent[3 * c + 1], ecz = a.cent[3 * c + 2];
                        T A[QB][een_naz<NORDC>()];
                        bool any_asym = false;
                        bool asym[QB];
#pragma unroll
                        for (int qq = 0; qq < QB; qq++) {
                            const double e0 = XQ[3 * qq] - ecx, e1 = XQ[3 * qq + 1] - ecy, e2 = XQ[3 * qq + 2] - ecz;
                            if (jg == 0) {  // geometry of the point seen from centre c, for the AOs (double)
                                double r2d = 0;
                                r2d = r2d + e0 * e0; r2d = r2d + e1 * e1; r2d = r2d + e2 * e2;
                                const double rqd = c == ic ? ren : sqrt(r2d);
                                double *g = GEO + ((size_t)qq * NC + c) * 5;
                                g[0] = e0; g[1] = e1; g[2] = e2; g[3] = rqd;
                                g[4] = (double)grid_jx(a.tab, rqd);
                            }
                            const T d0 = (T)e0, d1 = (T)e1, d2 = (T)e2;
                            T r2 = 0;
                            r2 = r2 + d0 * d0; r2 = r2 + d1 * d1; r2 = r2 + d2 * d2;
                            const T rq = c == ic ? (T)ren : sqrt(r2);
                            const T rr = scale_r<T>(kap, rq);
                            const T xq = switched_from<T>(kapd, rq, rr);
                            asym[qq] = rr == asymp;
                            any_asym = any_asym || asym[qq];
                            if (jg == 0 && q0 + qq < NQ) acc[qq] += en_term<T>(jp, it, rr, xq);
                            if constexpr (NORDC > 1) een_tables<NORDC, 0, T>(jp.c[it], xq, A[qq]);
                        }
                        if constexpr (NORDC > 1) {
                            // one partner of lookahead on the scaled distance
                            double rr_next = jg < NE ? rrt[(size_t)jg * NC + c] : 0.0;
                            for (int j = jg; j < NE; j += JG) {
                                const double rrj = rr_next;
                                if (j + JG < NE) rr_next = rrt[(size_t)(j + JG) * NC + c];
                                if (j == i) continue;
                                const T rj = (T)(1.0 - kapd * rrj);
                                const T *uj = U + (size_t)j * QB;
                                T P[QB];
#pragma unroll
                                for (int qq = 0; qq < QB; qq++) P[qq] = een_eval_tab<NORDC, 0, T>(A[qq], uj[qq], rj).f;
                                if (any_asym && rrj == asympd) {  // psinl: both distances asymptotic -> 0
#pragma unroll
                                    for (int qq = 0; qq < QB; qq++)
                                        if (asym[qq]) P[qq] = T(0);
                                }
#pragma unroll
                                for (int qq = 0; qq < QB; qq++) acc[qq] += P[qq];
                            }
                        }
                    }
                }
            }
#pragma unroll
            for (int qq = 0; qq < QB; qq++) {
                const double v = wave_sum((double)acc[qq]);
                if (lane == 0) S[(1 * QB + qq) * NL_WAVES + wv] = v;
            }

            // ---- phase C: psi ratio sum_m phi_m(x_q) W[i][m]; radial functions once
            // per (point, centre, shell), the angular factor per AO
            __syncthreads();  // GEO complete
            for (int tt = tid; tt < QB * NRI; tt += NL_THREADS) {
                const int qq = tt / NRI, ri = tt % NRI;
                const double *g = GEO + ((size_t)qq * NC + a.ri_cent[ri]) * 5;
                double wfv[3] = {0.0, 0.0, 0.0};
                if (g[3] <= a.ri_rmax[ri]) splfit_jx<double>(a.tab, a.ri_type[ri], a.ri_rad[ri], g[3], (int)g[4], wfv, 0);
                RAD[(size_t)qq * NRI + ri] = wfv[0];
            }
            __syncthreads();
            double dr[QB];
#pragma unroll
            for (int qq = 0; qq < QB; qq++) dr[qq] = 0.0;
            for (int m = tid; m < NB; m += NL_THREADS) {
                const int c = a.ao_cent[m], ri = a.ao_ri[m];
                const SlmMono mono = slm_decode(a.ao_slm[m]);  // once per AO, branch-free
                const double wm = Wm[m];
#pragma unroll
                for (int qq = 0; qq < QB; qq++) {
                    const double *g = GEO + ((size_t)qq * NC + c) * 5;
                    const double y = slm_value<double>(mono, g[0], g[1], g[2]);
                    dr[qq] += (y * RAD[(size_t)qq * NRI + ri]) * wm;
                }
            }
#pragma unroll
            for (int qq = 0; qq < QB; qq++) {
                const double v = wave_sum(dr[qq]);
                if (lane == 0) S[(2 * QB + qq) * NL_WAVES + wv] = v;
            }
            __syncthreads();

            // ---- the vpsp_det terms of these points (nonloc.f90:248-274)
            if (tid == 0) {
                double fi = 0.0;
                for (int k = 0; k < NL_WAVES; k++) fi += FI[k];
                for (int qq = 0; qq < QB && q0 + qq < NQ; qq++) {
                    const int q = q0 + qq;
                    double see = 0.0, sen = 0.0, drq = 0.0;
                    for (int k = 0; k < NL_WAVES; k++) {
                        see += S[(0 * QB + qq) * NL_WAVES + k];
                        sen += S[(1 * QB + qq) * NL_WAVES + k];
                        drq += S[(2 * QB + qq) * NL_WAVES + k];
                    }
                    const double fsumn = see + sen - fi;
                    double costh = rv0 * qx[q] + rv1 * qy[q] + rv2 * qz[q];
                    costh = costh * ri;
                    double term_radial = 0.0;
                    for (int l = 1; l <= lp - 1; l++)
                        term_radial = term_radial + yl0(l, costh) * a.pair_vps[slot * CHAMP_MPS_L + l - 1];
                    const double trj = term_radial * a.ecp->wq[q] * exp(fsumn);
                    pair_sum = pair_sum + trj * drq;
                    // t_vpsp of the point (nonloc.f90:297) for DMC's T-moves
                    if (a.tstore) a.pair_psi[slot * NQ + q] = trj * drq;
                }
            }
            __syncthreads();  // S, XQ, U, GEO and RAD are rewritten by the next point block
        }
        if (tid == 0) a.pair_val[slot] = pair_sum;
    }
}

// ------------------------------------------------------------------------
// Matrix-core form of the quadrature's e-e-n sum.  By the x <-> y symmetry of
// the polynomial, P(u, x, y) = sum_k u^k sum_n A_kn(y) x^n with the tables
// A_kn of the partner's distance y (een_poly.hpp).  For the points q of a
// group of pairs and the partners j of one walker,
//   M_k[q][j] = sum_{(b, n)} x_qb^n A_kn(y_jb)          (K = centres x powers)
// is a product of a (points x K) matrix of powers by a (K x partners) matrix
// of partner tables, and the e-e-n sum of point q is
//   sum_{j != i} sum_k u_qj^k M_k[q][j].
// A block takes MMA_GP active pairs of one walker (MMA_GP * NQ <= 48 points:
// three 16-row tiles).  Items (16 partners) are dealt to the four waves.  An
// item walks the centres four at a time: the four K-slots of an MFMA are four
// centres at one table entry (k, n), so a lane builds all table entries of
// its (centre, partner) once and feeds, for each row tile, one
// v_mfma_f64_16x16x4f64 per entry into the accumulator of row k (lane l
// supplies A[l % 16][l / 16] and B[l / 16][l % 16]; element r of the
// accumulator holds D[4 r + l / 16][l % 16]).  The partner distances are read
// from rr_en, partners innermost, so the 16 lanes of a slot load one line.
// The epilogue recomputes u_qj for its 4 x 1 (point, partner) entries and adds
// sum_k u^k M_k to per-lane point sums.  The e-e and e-n sums and the psi
// ratio are computed as in k_nonloc.
constexpr int MMA_GP = NL_GROUP, MMA_MT = 3, MMA_MP = 16 * MMA_MT, MMA_THREADS = 256, MMA_WAVES = MMA_THREADS / WAVE;
constexpr int MMA_TG = MMA_THREADS / MMA_MP;  // thread groups over (point, partner) in the point phase

// accumulator type and instruction per arithmetic type; row of accumulator
// element r in lane slot sl = l / 16 (measured on gfx950 with one-hot
// operands: f64 D[4 r + l / 16][l % 16], f32 D[4 (l / 16) + r][l % 16])
template <typename T> struct Mma;
template <> struct Mma<double> {
    typedef dbl4 acc;
    static __device__ __forceinline__ acc zero() { return dbl4{0.0, 0.0, 0.0, 0.0}; }
    static __device__ __forceinline__ acc op(double a, double b, acc c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
    static __device__ __forceinline__ int row(int r, int sl) { return 4 * r + sl; }
};
template <> struct Mma<float> {
    typedef flt4 acc;
    static __device__ __forceinline__ acc zero() { return flt4{0.0f, 0.0f, 0.0f, 0.0f}; }
    static __device__ __forceinline__ acc op(float a, float b, acc c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
    static __device__ __forceinline__ int row(int r, int sl) { return 4 * sl + r; }
};

// with_ee: the item also adds the e-e terms b(u_qj) of its (point, partner)
// entries (the first of the two halves of a split tile, or a whole tile)
template <int N, typename T>
__device__ __forceinline__ void mma_item(const HArgs &a, int w, int j0, int c_lo, int c_hi, int npts,
                                         const double *XP, const T *XQ, const int *IPT, double *swv, bool with_ee,
                                         const JasParams &jp, const int *sT)
{
    constexpr EenTable<N> ET{};
    constexpr int NA = een_naz<N>(), KR = ET.kmax + 1;
    typedef typename Mma<T>::acc Acc;
    const int lane = threadIdx.x % WAVE, sl = lane / 16, ln = lane % 16;
    const int NC = a.NC, NE = a.NE;
    const double kap = jp.kap;
    const int j = j0 + ln;
    const bool jv = j < NE;
    const double *rrw = a.rr_en + (size_t)w * NC * NE;
    Acc acc[MMA_MT][KR];
#pragma unroll
    for (int mt = 0; mt < MMA_MT; mt++)
#pragma unroll
        for (int k = 0; k < KR; k++) acc[mt][k] = Mma<T>::zero();
    // partner distances two centre steps ahead
    double ynext = (jv && c_lo + sl < c_hi) ? rrw[(size_t)(c_lo + sl) * NE + j] : 0.0;
    double ynext2 = (jv && c_lo + sl + 4 < c_hi) ? rrw[(size_t)(c_lo + sl + 4) * NE + j] : 0.0;
    for (int b0 = c_lo; b0 < c_hi; b0 += 4) {
        const int b = b0 + sl;
        const bool bv = b < c_hi;
        const double rrb = ynext;
        ynext = ynext2;
        if (b + 8 < c_hi && jv) ynext2 = rrw[(size_t)(b + 8) * NE + j];
        // the partner's table entries at centre b: built once, used by every row tile
        T tab[NA];
        een_tables<N, 0, T>(jp.c[bv ? sT[b] : 0], (T)(1.0 - kap * rrb), tab);
        if (!(bv && jv)) {
#pragma unroll
            for (int e = 0; e < NA; e++) tab[e] = T(0);
        }
        T xp[MMA_MT][N + 1];
#pragma unroll
        for (int mt = 0; mt < MMA_MT; mt++) {
            const int pta = mt * 16 + ln;
            const bool live_a = bv && pta < npts;
            const T x = live_a ? XQ[(size_t)b * MMA_MP + pta] : T(0);
            xp[mt][0] = live_a ? T(1) : T(0);
#pragma unroll
            for (int p = 1; p <= N; p++) xp[mt][p] = xp[mt][p - 1] * x;
        }
        // consecutive MFMAs go to different accumulators (row tile, power of
        // u), so none waits for the previous one's result
#pragma unroll
        for (int n = 0; n <= N; n++)
#pragma unroll
            for (int k = 0; k < KR; k++) {
                const int e = ET.idx[k][n];
                if (e < 0) continue;
#pragma unroll
                for (int mt = 0; mt < MMA_MT; mt++) acc[mt][k] = Mma<T>::op(xp[mt][n], tab[e], acc[mt][k]);
            }
    }
    const double *xw = a.xT + (size_t)w * 3 * NE;
    const double xj = jv ? xw[j] : 0.0, yj = jv ? xw[NE + j] : 0.0, zj = jv ? xw[2 * NE + j] : 0.0;
#pragma unroll
    for (int mt = 0; mt < MMA_MT; mt++)
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int pt = mt * 16 + Mma<T>::row(r, sl);
            double v = 0.0;
            if (jv && pt < npts && j != IPT[pt]) {
                // the switched e-e distance as the point phase forms it
                const T d0 = (T)(xj - XP[3 * pt]), d1 = (T)(yj - XP[3 * pt + 1]), d2 = (T)(zj - XP[3 * pt + 2]);
                T r2 = 0;
                r2 = r2 + d0 * d0; r2 = r2 + d1 * d1; r2 = r2 + d2 * d2;
                const T rjq = sqrt(r2);
                const T rr = scale_r<T>((T)kap, rjq);
                const T u = switched_from<T>(kap, rjq, rr);
                T pv2 = acc[mt][KR - 1][r];
#pragma unroll
                for (int k = KR - 2; k >= 0; k--) pv2 = pv2 * u + acc[mt][k][r];
                v = (double)pv2;
                if (with_ee) v += (double)ee_term<T>(jp, a.NUP, IPT[pt], j, rr, u);
            }
            // sum over the 16 partners (lanes of equal l / 16) into the wave's point slot
            v += __shfl_xor(v, 1, WAVE); v += __shfl_xor(v, 2, WAVE);
            v += __shfl_xor(v, 4, WAVE); v += __shfl_xor(v, 8, WAVE);
            if (ln == 0) swv[pt] += v;
        }
}

// psi-ratio phase: lanes over (point, radial-instance slot), MMA_PS slots
constexpr int MMA_PS = MMA_THREADS / MMA_MP;
static_assert(MMA_PS * MMA_MP <= 2 * MMA_TG * MMA_MP, "psi-ratio partials fit the point-phase partials");

// LDS of k_nonloc_mma in doubles: XP, SP, TG partials (reused for the
// psi-ratio partials), SW, FI, IPT/PI (as doubles), XQ (tsize bytes per
// entry), the point-centre distances (double) and grid indices (int), then
// the radial instances in slot order (8 ints and the cut radius each) and the
// AOs in slot order (slm code and AO index)
__host__ __device__ __forceinline__ size_t mma_lds_doubles(int NC, int NB, int NRI, int tsize)
{
    return 3 * MMA_MP + 3 * MMA_MP + 2 * MMA_TG * MMA_MP + MMA_WAVES * MMA_MP + MMA_GP + MMA_MP + 4 * MMA_GP +
           ((size_t)NC * MMA_MP * tsize + 7) / 8 + (size_t)NC * MMA_MP + ((size_t)NC * MMA_MP * 4 + 7) / 8 +
           (size_t)NRI * 4 + (size_t)NRI + (size_t)NB;
}

#ifndef CHAMP_MMA_WPE
#define CHAMP_MMA_WPE 2
#endif
// T: arithmetic type of the Jastrow terms and of the matrix cores (double on
// the FP64 path, float on the mixed path); sums and the psi ratio are double.
template <int NORDC, typename T>
__global__ void __launch_bounds__(MMA_THREADS) CHAMP_WAVES_PER_EU(CHAMP_MMA_WPE) k_nonloc_mma(HArgs a)
{
    HIP_DYNAMIC_SHARED(double, lds)
    constexpr EenTable<NORDC> ET{};
    static_assert(MMA_GP * 12 <= MMA_MP, "group size");
    const int tid = threadIdx.x, lane = tid % WAVE, wv = tid / WAVE;
    const int NE = a.NE, NC = a.NC, NQ = a.NQ, NB = a.NB, NRI = a.NRI;
    // the Jastrow parameters and the centres' types in LDS (read before use
    // behind the group loop's first barrier): every centre step of every item
    // reads a type's e-e-n coefficients, every epilogue entry the e-e ones
    __shared__ JasParams sJ;
    __shared__ int sT[128];
    static_assert(sizeof(JasParams) % 8 == 0, "JasParams copied as 8-byte words");
    for (int t = tid; t < (int)(sizeof(JasParams) / 8); t += MMA_THREADS)
        reinterpret_cast<unsigned long long *>(&sJ)[t] = reinterpret_cast<const unsigned long long *>(a.jas)[t];
    for (int t = tid; t < NC; t += MMA_THREADS) sT[t] = a.iwctype[t];
    const JasParams &jp = sJ;
    const double kapd = a.jas->kap;  // read directly: sJ is complete only after a barrier
    const T kap = (T)kapd;
    double *XP = lds;                                          // [MP][3]
    double *SP = XP + 3 * MMA_MP;                              // [3][MP]: e-e + e-n, e-e-n, psi ratio
    double *TP = SP + 3 * MMA_MP;                              // [2][TG][MP] partial sums
    double *SW = TP + 2 * MMA_TG * MMA_MP;                     // [WAVES][MP]
    double *FI = SW + MMA_WAVES * MMA_MP;                      // [GP]
    int *IPT = (int *)(FI + MMA_GP);                           // [MP] electron of each point
    int *PI = IPT + 2 * MMA_MP;                                // [GP][4]: i, ic
    T *XQ = (T *)(PI + 8 * MMA_GP);                            // [NC][MP] switched point-centre distances
    double *GR = (double *)PI + 4 * MMA_GP + ((size_t)NC * MMA_MP * sizeof(T) + 7) / 8;  // [MP][NC] point-centre r
    int *GI = (int *)(GR + (size_t)NC * MMA_MP);               // [MP][NC] grid interval of r
    double *PS = TP;                                           // [PS][MP] psi-ratio partials
    // the radial instances in slot order (centre, type, radial function, AO range)
    // and their AOs, staged once per block: the psi-ratio lanes read them per
    // (point, instance), so from global memory their latency dominated that phase
    int *SMI = GI + (size_t)NC * MMA_MP;                       // [NRI][8]
    double *SRM = (double *)(SMI + (size_t)NRI * 8);           // [NRI] cut radius
    int *SAO = (int *)(SRM + NRI);                             // [NB][2] slm code, AO index
    const bool psi_ext = a.psi_ext != 0;  // psi ratios from k_nonloc_psi
    if (!psi_ext) {
        for (int sidx = tid; sidx < NRI; sidx += MMA_THREADS) {
            const int ri = a.ri_perm[sidx];
            SMI[8 * sidx] = a.ri_cent[ri]; SMI[8 * sidx + 1] = a.ri_type[ri]; SMI[8 * sidx + 2] = a.ri_rad[ri];
            SMI[8 * sidx + 3] = a.ri_aoff[ri]; SMI[8 * sidx + 4] = a.ri_aoff[ri + 1];
            SRM[sidx] = a.ri_rmax[ri];
        }
        for (int k = tid; k < NB; k += MMA_THREADS) {
            const int m = a.ri_aol[k];
            SAO[2 * k] = a.ao_slm[m];
            SAO[2 * k + 1] = m;
        }
    }

    const int G = gridDim.x, bk = blockIdx.x;
    const int nx = (G % 8 == 0) ? 8 : 1, gx = G / nx, xcd = bk % nx, g0 = bk / nx;
    const int total = a.pair_total[1];
    const int per_x = (total + nx - 1) / nx;
    const int x_lo = xcd * per_x, x_hi = min(total, x_lo + per_x);

    for (int gi = x_lo + g0; gi < x_hi; gi += gx) {
        int lo = 0, hi = a.W - 1;
        while (lo < hi) {  // walker: largest w with grp_off[w] <= gi
            const int mid = (lo + hi + 1) / 2;
            if (a.grp_off[mid] <= gi) lo = mid; else hi = mid - 1;
        }
        const int w = lo;
        const int p0 = MMA_GP * (gi - a.grp_off[w]);
        const int np = min(MMA_GP, a.pair_count[w] - p0), npts = np * NQ;
        const double *xw = a.xT + (size_t)w * 3 * NE;
        __syncthreads();  // LDS of the previous group is free
        if (tid < np) {
            const size_t slot = (size_t)w * a.cap + p0 + tid;
            PI[4 * tid] = a.pair_e[slot];
            PI[4 * tid + 1] = a.pair_c[slot];
        }
        __syncthreads();
        for (int t = tid; t < MMA_MP; t += MMA_THREADS) IPT[t] = t < npts ? PI[4 * (t / NQ)] : -1;
        for (int t = tid; t < 3 * npts; t += MMA_THREADS) {
            const int pt = t / 3, k = t % 3, p = pt / NQ, q = pt % NQ;
            const int i = PI[4 * p], ic = PI[4 * p + 1];
            const double ren = a.r_en[((size_t)w * NC + ic) * NE + i];
            const double *qv = (k == 0 ? a.xq : (k == 1 ? a.yq : a.zq)) + (size_t)w * NQ;
            XP[t] = ren * qv[q] + a.cent[3 * ic + k];
        }
        if (wv < np) {  // sum_j F[i][j] of each pair
            const double *F = a.fso + ((size_t)w * NE + PI[4 * wv]) * NE;
            double fs = 0.0;
            for (int j = lane; j < NE; j += WAVE) fs += F[j];
            fs = wave_sum(fs);
            if (lane == 0) FI[wv] = fs;
        }
        __syncthreads();

        // ---- points: x_qb and the e-n sums, the e-e sums
        {
            const int pt = tid % MMA_MP, tg = tid / MMA_MP;
            double sen = 0.0, see = 0.0;
            if (tg < MMA_TG) {
                const bool live = pt < npts;
                const int p = live ? pt / NQ : 0;
                const int i = live ? PI[4 * p] : 0, ic = live ? PI[4 * p + 1] : 0;
                const double px = live ? XP[3 * pt] : 0.0, py = live ? XP[3 * pt + 1] : 0.0;
                const double pz = live ? XP[3 * pt + 2] : 0.0;
                for (int b = tg; b < NC; b += MMA_TG) {
                    T xq = T(0);
                    if (live) {
                        const T e0 = (T)(px - a.cent[3 * b]), e1 = (T)(py - a.cent[3 * b + 1]), e2 = (T)(pz - a.cent[3 * b + 2]);
                        T r2 = 0;
                        r2 = r2 + e0 * e0; r2 = r2 + e1 * e1; r2 = r2 + e2 * e2;
                        const T rq = b == ic ? (T)a.r_en[((size_t)w * NC + ic) * NE + i] : sqrt(r2);
                        const T rr = scale_r<T>(kap, rq);
                        xq = switched_from<T>(kapd, rq, rr);
                        sen += (double)en_term<T>(jp, a.iwctype[b], rr, xq);
                    }
                    XQ[(size_t)b * MMA_MP + pt] = xq;
                }
                // with the e-e-n terms on the matrix cores (NORDC > 1) the e-e
                // terms are added in their epilogue, which forms u_qj anyway
                if (NORDC <= 1 && live)
                    for (int j = tg; j < NE; j += MMA_TG) {
                        if (j == i) continue;
                        const T d0 = (T)(xw[j] - px), d1 = (T)(xw[NE + j] - py), d2 = (T)(xw[2 * NE + j] - pz);
                        T r2 = 0;
                        r2 = r2 + d0 * d0; r2 = r2 + d1 * d1; r2 = r2 + d2 * d2;
                        const T rjq = sqrt(r2);
                        const T u = scale_r<T>(kap, rjq);
                        see += (double)ee_term<T>(jp, a.NUP, i, j, u, switched_from<T>(kapd, rjq, u));
                    }
                TP[tg * MMA_MP + pt] = see;
                TP[(MMA_TG + tg) * MMA_MP + pt] = sen;
            }
        }
        __syncthreads();
        for (int pt = tid; pt < MMA_MP; pt += MMA_THREADS) {
            double v = 0.0;
            for (int tg = 0; tg < MMA_TG; tg++) v += TP[tg * MMA_MP + pt];
            for (int tg = 0; tg < MMA_TG; tg++) v += TP[(MMA_TG + tg) * MMA_MP + pt];
            SP[pt] = v;
        }

        // ---- e-e-n on the matrix cores: items (16 partners, all row tiles) over
        // the waves; each wave sums its items into its own row of SW
        {
            double *swv = SW + wv * MMA_MP;
            for (int pt = lane; pt < MMA_MP; pt += WAVE) swv[pt] = 0.0;
            __syncwarp();
            if constexpr (NORDC > 1) {
                // half-tiles (16 partners x half the centres) dealt in contiguous
                // runs of equal length to the waves; two halves of one tile in a
                // run make one item over all centres
                const int NT = (NE + 15) / 16, nh = 2 * NT, per = (nh + MMA_WAVES - 1) / MMA_WAVES;
                const int ncm = min(NC, ((NC / 2 + 3) / 4) * 4);  // centres of the first half (a multiple of 4)
                const int h_hi = min(nh, (wv + 1) * per);
                for (int h = wv * per; h < h_hi;) {
                    const int tile = h / 2;
                    if (h % 2 == 0 && h + 1 < h_hi) {
                        mma_item<NORDC, T>(a, w, tile * 16, 0, NC, npts, XP, XQ, IPT, swv, true, sJ, sT);
                        h += 2;
                    } else {
                        mma_item<NORDC, T>(a, w, tile * 16, h % 2 ? ncm : 0, h % 2 ? NC : ncm, npts, XP, XQ, IPT, swv,
                                           h % 2 == 0, sJ, sT);
                        h += 1;
                    }
                    __syncwarp();
                }
            }
        }
        __syncthreads();
        for (int pt = tid; pt < MMA_MP; pt += MMA_THREADS) {
            double v = 0.0;
            for (int k = 0; k < MMA_WAVES; k++) v += SW[k * MMA_MP + pt];
            SP[MMA_MP + pt] = v;
        }

        // ---- psi ratio of every point: sum_m phi_m(x_q) W[i][m].  The point-centre
        // distances and grid intervals first, then lanes over (point, radial
        // instance): a lane forms the instance's radial value at its point and
        // dots the instance's angular factors with the W row.  Instances are
        // taken in order of their AO count (ri_perm), so the lanes of a wave
        // mostly run the same number of AOs.
        __syncthreads();  // SW and the point-phase partials are free
        if (!psi_ext) {
        for (int t = tid; t < npts * NC; t += MMA_THREADS) {
            const int q = t / NC, c = t % NC, p = q / NQ;
            const double *xp = XP + 3 * q;
            const double e0 = xp[0] - a.cent[3 * c], e1 = xp[1] - a.cent[3 * c + 1], e2 = xp[2] - a.cent[3 * c + 2];
            double r2 = 0;
            r2 = r2 + e0 * e0; r2 = r2 + e1 * e1; r2 = r2 + e2 * e2;
            const double rqd = c == PI[4 * p + 1] ? a.r_en[((size_t)w * NC + c) * NE + PI[4 * p]] : sqrt(r2);
            GR[t] = rqd;
            GI[t] = grid_jx(a.tab, rqd);
        }
        __syncthreads();
        {
            const int q = tid % MMA_MP, slot = tid / MMA_MP;
            double part = 0.0;
            if (slot < MMA_PS && q < npts) {
                const double px = XP[3 * q], py = XP[3 * q + 1], pz = XP[3 * q + 2];
                const double *Wm = a.wmat + ((size_t)w * NE + IPT[q]) * NB;
                const double *gr = GR + (size_t)q * NC;
                const int *gi = GI + (size_t)q * NC;
                for (int sidx = slot; sidx < NRI; sidx += MMA_PS) {
                    const int *mi = SMI + 8 * sidx;
                    const int cen = mi[0];
                    const double r = gr[cen];
                    if (!(r <= SRM[sidx])) continue;  // beyond the shell's cut radius: its AOs are exactly zero
                    double wfv[3] = {0.0, 0.0, 0.0};
                    splfit_jx<double>(a.tab, mi[1], mi[2], r, gi[cen], wfv, 0);
                    const double d0 = px - a.cent[3 * cen], d1 = py - a.cent[3 * cen + 1], d2 = pz - a.cent[3 * cen + 2];
                    double sa = 0.0;
                    for (int k = mi[3]; k < mi[4]; k++)
                        sa += slm_value<double>(slm_decode(SAO[2 * k]), d0, d1, d2) * Wm[SAO[2 * k + 1]];
                    part += sa * wfv[0];
                }
            }
            PS[tid] = part;
        }
        __syncthreads();
        for (int q = tid; q < npts; q += MMA_THREADS) {
            double v = 0.0;
            for (int k = 0; k < MMA_PS; k++) v += PS[k * MMA_MP + q];
            SP[2 * MMA_MP + q] = v;
        }
        }  // !psi_ext
        __syncthreads();

        // ---- the vpsp_det terms (nonloc.f90:248-274): a thread per point
        // forms its term, then a thread per pair sums its points in order
        double *TRJ = TP;  // [MP], free once the point sums are in SP
        if (tid < npts) {
            const int pt = tid, p = pt / NQ, q = pt % NQ, i = PI[4 * p], ic = PI[4 * p + 1];
            const size_t slot = (size_t)w * a.cap + p0 + p;
            const int lp = a.ecp->lpot[sT[ic]];
            const double cx = a.cent[3 * ic], cy = a.cent[3 * ic + 1], cz = a.cent[3 * ic + 2];
            const double rv0 = xw[i] - cx, rv1 = xw[NE + i] - cy, rv2 = xw[2 * NE + i] - cz;
            const double ri = 1.0 / a.r_en[((size_t)w * NC + ic) * NE + i];
            const double *qx = a.xq + (size_t)w * NQ, *qy = a.yq + (size_t)w * NQ, *qz = a.zq + (size_t)w * NQ;
            const double fsumn = SP[pt] + SP[MMA_MP + pt] - FI[p];
            double costh = rv0 * qx[q] + rv1 * qy[q] + rv2 * qz[q];
            costh = costh * ri;
            double term_radial = 0.0;
            for (int l = 1; l <= lp - 1; l++)
                term_radial = term_radial + yl0(l, costh) * a.pair_vps[slot * CHAMP_MPS_L + l - 1];
            const double trj = term_radial * a.ecp->wq[q] * exp(fsumn);
            TRJ[pt] = trj * (psi_ext ? a.pair_psi[slot * NQ + q] : SP[2 * MMA_MP + pt]);
            // t_vpsp of the point (nonloc.f90:297) for DMC's T-moves
            if (a.tstore) a.pair_psi[slot * NQ + q] = TRJ[pt];
        }
        __syncthreads();
        if (tid < np) {
            const int p = tid;
            double pair_sum = 0.0;
            for (int q = 0; q < NQ; q++) pair_sum = pair_sum + TRJ[p * NQ + q];
            a.pair_val[(size_t)w * a.cap + p0 + p] = pair_sum;
        }
    }
}

// ------------------------------------------------------------------------
// The psi ratio of every quadrature point on its own: sum_m phi_m(x_q) W[i][m]
// (the orbital part of nonloc.f90:224-247, contracted with the inverse row
// once per electron: wmat).  The same groups of MMA_GP pairs and the same lane
// form as the psi phase of k_nonloc_mma, in a kernel of few registers and
// small LDS, so that more waves hide its table and W loads.  Results go to
// pair_psi[slot][q] for k_nonloc_mma (psi_ext).
constexpr int PSI_MP = MMA_MP, PSI_THREADS = MMA_THREADS;

__host__ __device__ __forceinline__ size_t psi_lds_bytes(int NC, int NB, int NRI)
{
    const size_t dbl = 3 * PSI_MP + (size_t)PSI_MP * NC + NRI + MMA_PS * PSI_MP;
    const size_t ints = PSI_MP + 2 * MMA_GP + (size_t)PSI_MP * NC + 5 * (size_t)NRI + 2 * (size_t)NB;
    return dbl * 8 + ((ints * 4 + 7) / 8) * 8;
}

__global__ void __launch_bounds__(PSI_THREADS) k_nonloc_psi(HArgs a)
{
    HIP_DYNAMIC_SHARED(double, lds)
    const int tid = threadIdx.x;
    const int NE = a.NE, NC = a.NC, NQ = a.NQ, NB = a.NB, NRI = a.NRI;
    double *XP = lds;                                  // [MP][3]
    double *GR = XP + 3 * PSI_MP;                      // [MP][NC] point-centre r
    double *SRM = GR + (size_t)PSI_MP * NC;            // [NRI] cut radius in slot order
    double *PS = SRM + NRI;                            // [PS][MP] partials
    int *IPT = (int *)(PS + MMA_PS * PSI_MP);          // [MP] electron of each point
    int *PI = IPT + PSI_MP;                            // [GP][2]: i, ic
    int *GI = PI + 2 * MMA_GP;                         // [MP][NC] grid interval of r
    int *SMI = GI + (size_t)PSI_MP * NC;               // [NRI][5] centre, type, radial fn, AO range
    int *SAO = SMI + 5 * (size_t)NRI;                  // [NB][2] slm code, AO index
    __shared__ double sCent[3 * 128];                  // nuclear positions (ncent <= 128)
    for (int t = tid; t < 3 * NC; t += PSI_THREADS) sCent[t] = a.cent[t];
    for (int sidx = tid; sidx < NRI; sidx += PSI_THREADS) {
        const int ri = a.ri_perm[sidx];
        SMI[5 * sidx] = a.ri_cent[ri]; SMI[5 * sidx + 1] = a.ri_type[ri]; SMI[5 * sidx + 2] = a.ri_rad[ri];
        SMI[5 * sidx + 3] = a.ri_aoff[ri]; SMI[5 * sidx + 4] = a.ri_aoff[ri + 1];
        SRM[sidx] = a.ri_rmax[ri];
    }
    for (int k = tid; k < NB; k += PSI_THREADS) {
        const int m = a.ri_aol[k];
        SAO[2 * k] = a.ao_slm[m];
        SAO[2 * k + 1] = m;
    }
    const int G = gridDim.x, bk = blockIdx.x;
    const int nx = (G % 8 == 0) ? 8 : 1, gx = G / nx, xcd = bk % nx, g0 = bk / nx;
    const int total = a.pair_total[1];
    const int per_x = (total + nx - 1) / nx;
    const int x_lo = xcd * per_x, x_hi = min(total, x_lo + per_x);
    for (int gi = x_lo + g0; gi < x_hi; gi += gx) {
        int lo = 0, hi = a.W - 1;
        while (lo < hi) {  // walker: largest w with grp_off[w] <= gi
            const int mid = (lo + hi + 1) / 2;
            if (a.grp_off[mid] <= gi) lo = mid; else hi = mid - 1;
        }
        const int w = lo;
        const int p0 = MMA_GP * (gi - a.grp_off[w]);
        const int np = min(MMA_GP, a.pair_count[w] - p0), npts = np * NQ;
        __syncthreads();  // LDS of the previous group is free
        if (tid < np) {
            const size_t slot = (size_t)w * a.cap + p0 + tid;
            PI[2 * tid] = a.pair_e[slot];
            PI[2 * tid + 1] = a.pair_c[slot];
        }
        __syncthreads();
        for (int t = tid; t < PSI_MP; t += PSI_THREADS) IPT[t] = t < npts ? PI[2 * (t / NQ)] : -1;
        for (int t = tid; t < 3 * npts; t += PSI_THREADS) {
            const int pt = t / 3, k = t % 3, p = pt / NQ, q = pt % NQ;
            const int i = PI[2 * p], ic = PI[2 * p + 1];
            const double ren = a.r_en[((size_t)w * NC + ic) * NE + i];
            const double *qv = (k == 0 ? a.xq : (k == 1 ? a.yq : a.zq)) + (size_t)w * NQ;
            XP[t] = ren * qv[q] + sCent[3 * ic + k];
        }
        __syncthreads();
        for (int t = tid; t < npts * NC; t += PSI_THREADS) {
            const int q = t / NC, c = t % NC, p = q / NQ;
            const double *xp = XP + 3 * q;
            const double e0 = xp[0] - sCent[3 * c], e1 = xp[1] - sCent[3 * c + 1], e2 = xp[2] - sCent[3 * c + 2];
            double r2 = 0;
            r2 = r2 + e0 * e0; r2 = r2 + e1 * e1; r2 = r2 + e2 * e2;
            const double rqd = c == PI[2 * p + 1] ? a.r_en[((size_t)w * NC + c) * NE + PI[2 * p]] : sqrt(r2);
            GR[t] = rqd;
            GI[t] = grid_jx(a.tab, rqd);
        }
        __syncthreads();
        {
            const int q = tid % PSI_MP, slot = tid / PSI_MP;
            if (slot < MMA_PS) {
                double part = 0.0;
                if (q < npts) {
                    const double px = XP[3 * q], py = XP[3 * q + 1], pz = XP[3 * q + 2];
                    const double *Wm = a.wmat + ((size_t)w * NE + IPT[q]) * NB;
                    const double *gr = GR + (size_t)q * NC;
                    const int *gi = GI + (size_t)q * NC;
                    for (int sidx = slot; sidx < NRI; sidx += MMA_PS) {
                        const int *mi = SMI + 5 * sidx;
                        const int cen = mi[0];
                        const double r = gr[cen];
                        if (!(r <= SRM[sidx])) continue;  // beyond the shell's cut radius: its AOs are exactly zero
                        double wfv[3] = {0.0, 0.0, 0.0};
                        splfit_jx<double>(a.tab, mi[1], mi[2], r, gi[cen], wfv, 0);
                        const double d0 = px - sCent[3 * cen], d1 = py - sCent[3 * cen + 1], d2 = pz - sCent[3 * cen + 2];
                        double sa = 0.0;
                        for (int k = mi[3]; k < mi[4]; k++)
                            sa += slm_value<double>(slm_decode(SAO[2 * k]), d0, d1, d2) * Wm[SAO[2 * k + 1]];
                        part += sa * wfv[0];
                    }
                }
                PS[tid] = part;
            }
        }
        __syncthreads();
        for (int q = tid; q < npts; q += PSI_THREADS) {
            double v = 0.0;
            for (int k = 0; k < MMA_PS; k++) v += PS[k * PSI_MP + q];
            a.pair_psi[((size_t)w * a.cap + p0 + q / NQ) * NQ + q % NQ] = v;
        }
    }
}

#ifndef CHAMP_NL_SPLIT
#define CHAMP_NL_SPLIT 1
#endif

static int launch_psi(champ_ctx *c, const HArgs &a)
{
    const size_t lds = psi_lds_bytes(c->NC, c->NB, c->NRI);
    if (lds > 64 * 1024)
        HIPCHECK(hipFuncSetAttribute((const void *)k_nonloc_psi, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    int dev = 0, ncu = 256, per_cu = 1;
    HIPCHECK(hipGetDevice(&dev));
    HIPCHECK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    HIPCHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_nonloc_psi, PSI_THREADS, lds));
    per_cu = std::max(1, per_cu);
    hipLaunchKernelGGL(k_nonloc_psi, dim3(ncu * per_cu), dim3(PSI_THREADS), lds, c->stream, a);
    HIPCHECK(hipGetLastError());
    return CHAMP_OK;
}

template <int NORDC, typename T>
static int launch_mma(champ_ctx *c, const HArgs &a0)
{
    HArgs a = a0;
    const bool split = CHAMP_NL_SPLIT && psi_lds_bytes(c->NC, c->NB, c->NRI) <= 160 * 1024;
    int slot;
    if (split) {
        tbegin(c, K_HPSI_PSI, &slot);
        const int rc = launch_psi(c, a);
        tend(c, slot);
        if (rc) return rc;
        a.psi_ext = 1;
    }
    // without the psi phase the tables after XQ (point-centre distances,
    // instances, AOs) are not used
    const size_t lds = (mma_lds_doubles(c->NC, c->NB, c->NRI, (int)sizeof(T)) -
                        (split ? (size_t)c->NC * MMA_MP + ((size_t)c->NC * MMA_MP * 4 + 7) / 8 + (size_t)c->NRI * 4 +
                                     (size_t)c->NRI + (size_t)c->NB
                               : 0)) *
                       sizeof(double);
    if (lds > 160 * 1024) {
        set_error("nonloc: the quadrature kernel's LDS tables exceed 160 KB for this system");
        return CHAMP_ERR_UNSUPPORTED;
    }
    if (lds > 64 * 1024)
        HIPCHECK(hipFuncSetAttribute((const void *)k_nonloc_mma<NORDC, T>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    int dev = 0, ncu = 256, per_cu = 1;
    HIPCHECK(hipGetDevice(&dev));
    HIPCHECK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    HIPCHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_nonloc_mma<NORDC, T>, MMA_THREADS, lds));
    per_cu = std::max(1, per_cu);
    tbegin(c, K_HPSI_NONLOC, &slot);
    hipLaunchKernelGGL((k_nonloc_mma<NORDC, T>), dim3(ncu * per_cu), dim3(MMA_THREADS), lds, c->stream, a);
    tend(c, slot);
    HIPCHECK(hipGetLastError());
    return CHAMP_OK;
}

template <int NORDC, int QB, typename T>
static int launch_k(champ_ctx *c, const HArgs &a)
{
    const size_t lds = (3 * QB * NL_WAVES + 3 * QB + NL_WAVES + (size_t)QB * c->NC * 5 + (size_t)QB * c->NRI) * sizeof(double) +
                       (size_t)c->NE * QB * sizeof(T);
    if (lds > 160 * 1024) {
        set_error("nonloc: the quadrature kernel's LDS tables exceed 160 KB for this system");
        return CHAMP_ERR_UNSUPPORTED;
    }
    if (lds > 64 * 1024)
        HIPCHECK(hipFuncSetAttribute((const void *)k_nonloc<NORDC, QB, T>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    // one wave of blocks: as many as can be resident at once
    int dev = 0, ncu = 256, per_cu = 1;
    HIPCHECK(hipGetDevice(&dev));
    HIPCHECK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    HIPCHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_nonloc<NORDC, QB, T>, NL_THREADS, lds));
    per_cu = std::max(1, per_cu);
    int slot;
    tbegin(c, K_HPSI_NONLOC, &slot);
    hipLaunchKernelGGL((k_nonloc<NORDC, QB, T>), dim3(ncu * per_cu), dim3(NL_THREADS), lds, c->stream, a);
    tend(c, slot);
    HIPCHECK(hipGetLastError());
    return CHAMP_OK;
}

template <typename T>
static int launch_t(champ_ctx *c, const HArgs &a)
{
    switch (c->hjas.nordc) {
    case 0: return launch_k<0, CHAMP_NL_QB, T>(c, a);
    case 1: return launch_k<1, CHAMP_NL_QB, T>(c, a);
    case 2: return launch_k<2, CHAMP_NL_QB, T>(c, a);
    case 3: return launch_k<3, CHAMP_NL_QB, T>(c, a);
    case 4: return launch_k<4, CHAMP_NL_QB, T>(c, a);
    case 5: return launch_k<5, 2, T>(c, a);
    default:
        set_error("nonloc: e-e-n order above 5 is not supported");
        return CHAMP_ERR_UNSUPPORTED;
    }
}

#ifndef CHAMP_NL_MMA
#define CHAMP_NL_MMA 1
#endif

int launch_nonloc(champ_ctx *c, const HArgs &a)
{
    // the matrix-core form: up to 12 points per pair (48 points per group) and
    // its LDS within 160 KB
    const bool mixed = c->prec == CHAMP_PREC_MIXED;
    const bool mma = CHAMP_NL_MMA && c->tune_nonloc != 1 && c->NQ <= 12 && c->hjas.nordc >= 2 && c->hjas.nordc <= 5 &&
                     mma_lds_doubles(c->NC, c->NB, c->NRI, mixed ? 4 : 8) * 8 <= 160 * 1024;
    if (mma && mixed) {
        switch (c->hjas.nordc) {
        case 2: return launch_mma<2, float>(c, a);
        case 3: return launch_mma<3, float>(c, a);
        case 4: return launch_mma<4, float>(c, a);
        default: return launch_mma<5, float>(c, a);
        }
    }
    if (mma) {
        switch (c->hjas.nordc) {
        case 2: return launch_mma<2, double>(c, a);
        case 3: return launch_mma<3, double>(c, a);
        case 4: return launch_mma<4, double>(c, a);
        default: return launch_mma<5, double>(c, a);
        }
    }
    return c->prec == CHAMP_PREC_MIXED ? launch_t<float>(c, a) : launch_t<double>(c, a);
}

}  // namespace champ
