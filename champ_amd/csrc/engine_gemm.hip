// engine_gemm.hip -- population-wide FP64 contractions of the engine:
//   AO -> MO (orbitals / orbitalse, orbitals_no_qmckl.f90:73-128, 245-258):
//       orb[r][o] = sum_m ao[r][m] * coef[o][m]        (B transposed, "NT")
//   inverse rows contracted with the occupied MO coefficients for the
//   quadrature ratios (nonlocd_quad, nonloc.f90:462-506, reassociated):
//       wmat[b][i][m] = sum_j slmi[b][i][j] * cocc[j][m] ("NN", batched)
// Matrix cores: v_mfma_f64_16x16x4f64, one 16x16 output tile per wavefront.
#include "engine_internal.hpp"

namespace champ {

// 16 x 16 output tiles along N per wavefront of the tile form.  Three tiles
// share each A fragment: the per-move contraction (16384 x 80 x 600) measured
// 0.127 ms with one tile, 0.109 with two, 0.094 with three, 0.115 with four
// and 0.133 with five per wavefront (identical results).
#ifndef CHAMP_GEMM_TILE_WN
#define CHAMP_GEMM_TILE_WN 3
#endif

// Which GEMM kernel launch_gemm uses: 1 = matrix cores (k_gemm_mfma),
// 0 = the LDS-tiled VALU kernel (k_gemm_valu).
#ifndef CHAMP_GEMM_MFMA
#define CHAMP_GEMM_MFMA 1
#endif

// LDS-tiled VALU GEMM: 64x64 output tile per 256-thread block, 4x4 per
// thread, K staged through LDS in slabs of 16.
#define GB_M 64
#define GB_N 64
#define GB_K 16

template <bool BT>
__global__ void __launch_bounds__(256) k_gemm_valu(int M, int N, int K, const double *__restrict__ A, int lda,
                                              long long sA, const double *__restrict__ B, int ldb, long long sB,
                                              double *__restrict__ Cm, int ldc, long long sC)
{
    __shared__ double As[GB_K][GB_M + 1];
    __shared__ double Bs[GB_K][GB_N + 1];
    const int tid = threadIdx.x, tx = tid % 16, ty = tid / 16;
    const int m0 = blockIdx.x * GB_M, n0 = blockIdx.y * GB_N;
    A += (long long)blockIdx.z * sA;
    B += (long long)blockIdx.z * sB;
    Cm += (long long)blockIdx.z * sC;
    double acc[4][4];
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) acc[i][j] = 0.0;
    for (int k0 = 0; k0 < K; k0 += GB_K) {
        // A slab: 64 rows x 16 k, 4 elements per thread, k fastest in memory
        for (int t = tid; t < GB_M * GB_K; t += 256) {
            const int r = t / GB_K, kk = t % GB_K;
            const int gm = m0 + r, gk = k0 + kk;
            As[kk][r] = (gm < M && gk < K) ? A[(long long)gm * lda + gk] : 0.0;
        }
        for (int t = tid; t < GB_N * GB_K; t += 256) {
            int nn, kk;
            if (BT) { nn = t / GB_K; kk = t % GB_K; }
            else { kk = t / GB_N; nn = t % GB_N; }
            const int gn = n0 + nn, gk = k0 + kk;
            double v = 0.0;
            if (gn < N && gk < K) v = BT ? B[(long long)gn * ldb + gk] : B[(long long)gk * ldb + gn];
            Bs[kk][nn] = v;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < GB_K; kk++) {
            double a[4], b[4];
            for (int i = 0; i < 4; i++) a[i] = As[kk][ty + 16 * i];
            for (int j = 0; j < 4; j++) b[j] = Bs[kk][tx + 16 * j];
            for (int i = 0; i < 4; i++)
                for (int j = 0; j < 4; j++) acc[i][j] += a[i] * b[j];
        }
        __syncthreads();
    }
    for (int i = 0; i < 4; i++) {
        const int gm = m0 + ty + 16 * i;
        if (gm >= M) continue;
        for (int j = 0; j < 4; j++) {
            const int gn = n0 + tx + 16 * j;
            if (gn < N) Cm[(long long)gm * ldc + gn] = acc[i][j];
        }
    }
}


// FP64 MFMA GEMM.  The instruction computes D(16x16) += A(16x4) B(4x16) with
// lane l holding A[l % 16][l / 16], B[l / 16][l % 16] and D[4 r + l / 16][l % 16]
// in accumulator element r = 0..3 (layout measured on gfx950).  A K-chunk of 16 is fed as four instructions in which lane l's
// k-slot (l / 16) carries k = k0 + 4 (l / 16) + s at step s: the same mapping
// for A and B, so the chunk's 16 products are all summed, and each lane reads
// 4 consecutive k of its row (a contiguous 32 bytes of A, and of B^T).
// One wavefront owns one 16 x (16 WN) tile of C; no LDS, no block barrier.
template <bool BT, int WN>
__global__ void __launch_bounds__(256) k_gemm_mfma(int M, int N, int K, const double *__restrict__ A, int lda,
                                                   long long sA, const double *__restrict__ B, int ldb, long long sB,
                                                   double *__restrict__ Cm, int ldc, long long sC)
{
    const int lane = threadIdx.x % WAVE;
    const int nrt = (M + 15) / 16, ncg = (N + 16 * WN - 1) / (16 * WN);
    const long long g = (long long)blockIdx.x * (blockDim.x / WAVE) + threadIdx.x / WAVE;
    if (g >= (long long)nrt * ncg) return;  // whole wavefront
    const int cg = (int)(g % ncg), rt = (int)(g / ncg);
    A += (long long)blockIdx.y * sA;
    B += (long long)blockIdx.y * sB;
    Cm += (long long)blockIdx.y * sC;
    const int row = rt * 16 + lane % 16, kq = 4 * (lane / 16);
    const bool rok = row < M;
    const double *Ar = A + (long long)(rok ? row : 0) * lda;
    int col[WN];
    bool cok[WN];
#pragma unroll
    for (int t = 0; t < WN; t++) {
        col[t] = cg * 16 * WN + t * 16 + lane % 16;
        cok[t] = col[t] < N;
    }
    dbl4 acc[WN];
#pragma unroll
    for (int t = 0; t < WN; t++) acc[t] = dbl4{0.0, 0.0, 0.0, 0.0};
    // the operands of chunk k0 + 16 are loaded while chunk k0 runs on the
    // matrix cores (one chunk of software pipelining)
    auto load = [&](int k0, double *a, double (*b)[4]) {
        const int kb = k0 + kq;
#pragma unroll
        for (int s = 0; s < 4; s++) a[s] = (rok && kb + s < K) ? Ar[kb + s] : 0.0;
#pragma unroll
        for (int t = 0; t < WN; t++)
#pragma unroll
            for (int s = 0; s < 4; s++) {
                const bool ok = cok[t] && kb + s < K;
                const long long off = BT ? (long long)(cok[t] ? col[t] : 0) * ldb + (kb + s < K ? kb + s : 0)
                                         : (long long)(kb + s < K ? kb + s : 0) * ldb + (cok[t] ? col[t] : 0);
                b[t][s] = ok ? B[off] : 0.0;
            }
    };
    double an[4], bn[WN][4];
    load(0, an, bn);
    for (int k0 = 0; k0 < K; k0 += 16) {
        double a[4], b[WN][4];
#pragma unroll
        for (int s = 0; s < 4; s++) a[s] = an[s];
#pragma unroll
        for (int t = 0; t < WN; t++)
#pragma unroll
            for (int s = 0; s < 4; s++) b[t][s] = bn[t][s];
        if (k0 + 16 < K) load(k0 + 16, an, bn);
#pragma unroll
        for (int t = 0; t < WN; t++)
#pragma unroll
            for (int s = 0; s < 4; s++) acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[s], b[t][s], acc[t], 0, 0, 0);
    }
#pragma unroll
    for (int t = 0; t < WN; t++) {
        if (!cok[t]) continue;
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int ro = rt * 16 + 4 * r + lane / 16;
            if (ro < M) Cm[(long long)ro * ldc + col[t]] = acc[t][r];
        }
    }
}

// FP32 operands, FP32 accumulation, FP64 result: the per-move AO->MO
// contraction of the mixed path on v_mfma_f32_16x16x4f32 (A[l % 16][l / 16],
// B[l / 16][l % 16] as above; accumulator element r of lane l holds
// D[4 (l / 16) + r][l % 16], measured on gfx950).  Same k-slot pairing as
// k_gemm_mfma: each lane reads 4 consecutive k of its row (16 bytes).
template <bool BT, int WN>
__global__ void __launch_bounds__(256) k_gemm_mfma_f32(int M, int N, int K, const float *__restrict__ A, int lda,
                                                       const float *__restrict__ B, int ldb, double *__restrict__ Cm,
                                                       int ldc)
{
    const int lane = threadIdx.x % WAVE;
    const int nrt = (M + 15) / 16, ncg = (N + 16 * WN - 1) / (16 * WN);
    const long long g = (long long)blockIdx.x * (blockDim.x / WAVE) + threadIdx.x / WAVE;
    if (g >= (long long)nrt * ncg) return;  // whole wavefront
    const int cg = (int)(g % ncg), rt = (int)(g / ncg);
    const int row = rt * 16 + lane % 16, kq = 4 * (lane / 16);
    const bool rok = row < M;
    const float *Ar = A + (long long)(rok ? row : 0) * lda;
    int col[WN];
    bool cok[WN];
#pragma unroll
    for (int t = 0; t < WN; t++) {
        col[t] = cg * 16 * WN + t * 16 + lane % 16;
        cok[t] = col[t] < N;
    }
    flt4 acc[WN];
#pragma unroll
    for (int t = 0; t < WN; t++) acc[t] = flt4{0.0f, 0.0f, 0.0f, 0.0f};
    for (int k0 = 0; k0 < K; k0 += 16) {
        const int kb = k0 + kq;
        float a[4];
#pragma unroll
        for (int s = 0; s < 4; s++) a[s] = (rok && kb + s < K) ? Ar[kb + s] : 0.0f;
#pragma unroll
        for (int t = 0; t < WN; t++) {
            float b[4];
#pragma unroll
            for (int s = 0; s < 4; s++) {
                const bool ok = cok[t] && kb + s < K;
                const long long off = BT ? (long long)(cok[t] ? col[t] : 0) * ldb + (kb + s < K ? kb + s : 0)
                                         : (long long)(kb + s < K ? kb + s : 0) * ldb + (cok[t] ? col[t] : 0);
                b[s] = ok ? B[off] : 0.0f;
            }
#pragma unroll
            for (int s = 0; s < 4; s++) acc[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[s], b[s], acc[t], 0, 0, 0);
        }
    }
#pragma unroll
    for (int t = 0; t < WN; t++) {
        if (!cok[t]) continue;
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int ro = rt * 16 + 4 * (lane / 16) + r;
            if (ro < M) Cm[(long long)ro * ldc + col[t]] = (double)acc[t][r];
        }
    }
}

int launch_gemm_f32(champ_ctx *c, int M, int N, int K, const float *A, int lda, const float *B, int ldb, double *Cm,
                    int ldc)
{
    if (M <= 0 || N <= 0) return CHAMP_OK;
    constexpr int TWN = CHAMP_GEMM_TILE_WN;  // tiles along N per wavefront, as the FP64 tile form
    const long long waves = (long long)((M + 15) / 16) * ((N + 16 * TWN - 1) / (16 * TWN));
    const long long blocks = (waves + 3) / 4;
    if (blocks > 0x7fffffffLL) {
        set_error("gemm: problem too large for one launch");
        return CHAMP_ERR_UNSUPPORTED;
    }
    hipLaunchKernelGGL((k_gemm_mfma_f32<true, TWN>), dim3((unsigned)blocks), dim3(256), 0, c->stream, M, N, K, A, lda,
                       B, ldb, Cm, ldc);
    HIPCHECK(hipGetLastError());
    return CHAMP_OK;
}

// Strip variant: each wavefront owns 16 rows x (16 WN) columns, so every A
// value it loads feeds WN instructions; the block's four wavefronts share the
// B chunk of 16 k x (16 WN) columns through LDS (double-buffered, one barrier
// per chunk).  Same k re-pairing and accumulator layout as k_gemm_mfma.
// LOWER: blocks entirely above the diagonal (first column beyond the last
// row) return at once, for symmetric products of which the lower triangle
// is kept; blocks across the diagonal write their whole tile.
template <bool BT, int WN, bool LOWER = false>
__global__ void __launch_bounds__(256) k_gemm_mfma_strip(int M, int N, int K, const double *__restrict__ A, int lda,
                                                         long long sA, const double *__restrict__ B, int ldb,
                                                         long long sB, double *__restrict__ Cm, int ldc, long long sC)
{
    constexpr int BN = 16 * WN;
    __shared__ double Bs[2][16][BN + 1];
    if (LOWER && (long long)blockIdx.y * BN > (long long)blockIdx.x * 64 + 63) return;  // the whole block
    const int tid = threadIdx.x, lane = tid % WAVE, wv = tid / WAVE;
    const int m0 = blockIdx.x * 64 + wv * 16, n0 = blockIdx.y * BN;
    A += (long long)blockIdx.z * sA;
    B += (long long)blockIdx.z * sB;
    Cm += (long long)blockIdx.z * sC;
    const int row = m0 + lane % 16, kq = 4 * (lane / 16);
    const bool rok = row < M;
    const double *Ar = A + (long long)(rok ? row : 0) * lda;
    dbl4 acc[WN];
#pragma unroll
    for (int t = 0; t < WN; t++) acc[t] = dbl4{0.0, 0.0, 0.0, 0.0};
    auto load_b = [&](int buf, int k0) {
        for (int e = tid; e < 16 * BN; e += 256) {
            int kk, nn;
            if (BT) { nn = e / 16; kk = e % 16; }
            else { kk = e / BN; nn = e % BN; }
            const int gk = k0 + kk, gn = n0 + nn;
            double v = 0.0;
            if (gk < K && gn < N) v = BT ? B[(long long)gn * ldb + gk] : B[(long long)gk * ldb + gn];
            Bs[buf][kk][nn] = v;
        }
    };
    const int nk = (K + 15) / 16;
    load_b(0, 0);
    __syncthreads();
    for (int kc = 0; kc < nk; kc++) {
        const int buf = kc & 1, k0 = kc * 16;
        if (kc + 1 < nk) load_b(buf ^ 1, k0 + 16);  // the other buffer was last read before the previous barrier
        const int kb = k0 + kq;
        double a[4];
#pragma unroll
        for (int s = 0; s < 4; s++) a[s] = (rok && kb + s < K) ? Ar[kb + s] : 0.0;
#pragma unroll
        for (int t = 0; t < WN; t++)
#pragma unroll
            for (int s = 0; s < 4; s++)
                acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[s], Bs[buf][kq + s][t * 16 + lane % 16], acc[t], 0, 0, 0);
        __syncthreads();
    }
#pragma unroll
    for (int t = 0; t < WN; t++) {
        const int col = n0 + t * 16 + lane % 16;
        if (col >= N) continue;
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int ro = m0 + 4 * r + lane / 16;
            if (ro < M) Cm[(long long)ro * ldc + col] = acc[t][r];
        }
    }
}

// Cm[i][j] = t[i][j] + t[M + i][j]: the two split-K partials (t ld N)
__global__ void k_gemm_add_halves(int M, int N, const double *__restrict__ t, double *__restrict__ Cm, int ldc)
{
    const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (long long)M * N) return;
    const int i = (int)(e / N), j = (int)(e % N);
    Cm[(long long)i * ldc + j] = t[e] + t[(long long)M * N + e];
}

// Scratch for a split-K product of M x N (reserved before any launch that may
// be captured in a graph; launch_gemm falls back to the tile form without it).
int gemm_reserve_splitk(champ_ctx *c, int M, int N)
{
    const size_t need = (size_t)2 * M * N * sizeof(double);
    if (need <= c->gemm_tmp_cap) return CHAMP_OK;
    if (c->gemm_tmp) HIPCHECK(hipFree(c->gemm_tmp));
    c->gemm_tmp = nullptr;
    c->gemm_tmp_cap = 0;
    HIPCHECK(hipMalloc(&c->gemm_tmp, need));
    c->gemm_tmp_cap = need;
    return CHAMP_OK;
}

// The lower triangle (and the tiles across the diagonal) of C = A A^T, C
// (n x n, ldc), A (n x K, lda): the strip kernel without its upper blocks.
int launch_syrk_lower(champ_ctx *c, int n, int K, const double *A, int lda, double *Cm, int ldc)
{
    if (n <= 0) return CHAMP_OK;
    constexpr int WN = 5, BN = 16 * WN;
    const long long gx = (n + 63) / 64, gy = (n + BN - 1) / BN;
    if (gx > 0x7fffffffLL || gy > 65535) {
        set_error("gemm: problem too large for one launch");
        return CHAMP_ERR_UNSUPPORTED;
    }
    hipLaunchKernelGGL((k_gemm_mfma_strip<true, WN, true>), dim3((unsigned)gx, (unsigned)gy, 1), dim3(256), 0, c->stream,
                       n, n, K, A, lda, 0LL, A, lda, 0LL, Cm, ldc, 0LL);
    HIPCHECK(hipGetLastError());
    return CHAMP_OK;
}

// C[b] (M x N, row-major, ldc) = A[b] (M x K, lda) * op(B[b]); op = B^T when
// bt (B stored N x K) else B (K x N).
int launch_gemm(champ_ctx *c, bool bt, int M, int N, int K, const double *A, int lda, long long sA, const double *B,
                int ldb, long long sB, double *Cm, int ldc, long long sC, int batch, int kind)
{
    if (M <= 0 || N <= 0 || batch <= 0) return CHAMP_OK;
    if (kind == GEMM_SPLITK) {
        // the tile form on k in [0, K/2) and [K/2, K) as a batch of two, so twice
        // as many wavefronts share the work; the partials are then added in a
        // fixed order (reproducible, no atomics)
        constexpr int TWN = CHAMP_GEMM_TILE_WN;
        const long long waves = (long long)((M + 15) / 16) * ((N + 16 * TWN - 1) / (16 * TWN));
        const long long blocks = (waves + 3) / 4;
        const bool ok = CHAMP_GEMM_MFMA && batch == 1 && K >= 32 && K % 2 == 0 && blocks <= 0x7fffffffLL &&
                        c->gemm_tmp && c->gemm_tmp_cap >= (size_t)2 * M * N * sizeof(double);
        if (ok) {
            const int Kh = K / 2;
            const long long sBh = bt ? (long long)Kh : (long long)Kh * ldb;
            dim3 grid((unsigned)blocks, 2);
            if (bt)
                hipLaunchKernelGGL((k_gemm_mfma<true, TWN>), grid, dim3(256), 0, c->stream, M, N, Kh, A, lda, (long long)Kh,
                                   B, ldb, sBh, c->gemm_tmp, N, (long long)M * N);
            else
                hipLaunchKernelGGL((k_gemm_mfma<false, TWN>), grid, dim3(256), 0, c->stream, M, N, Kh, A, lda,
                                   (long long)Kh, B, ldb, sBh, c->gemm_tmp, N, (long long)M * N);
            HIPCHECK(hipGetLastError());
            const long long ne = (long long)M * N;
            hipLaunchKernelGGL(k_gemm_add_halves, dim3((unsigned)((ne + 255) / 256)), dim3(256), 0, c->stream, M, N,
                               c->gemm_tmp, Cm, ldc);
            HIPCHECK(hipGetLastError());
            return CHAMP_OK;
        }
        kind = GEMM_TILE;
    }
    if (!CHAMP_GEMM_MFMA || kind == GEMM_VALU) {
        if (batch > 65535) {
            set_error("gemm: batch too large for one launch");
            return CHAMP_ERR_UNSUPPORTED;
        }
        dim3 grid((M + GB_M - 1) / GB_M, (N + GB_N - 1) / GB_N, batch);
        if (bt)
            hipLaunchKernelGGL(k_gemm_valu<true>, grid, dim3(256), 0, c->stream, M, N, K, A, lda, sA, B, ldb, sB, Cm, ldc, sC);
        else
            hipLaunchKernelGGL(k_gemm_valu<false>, grid, dim3(256), 0, c->stream, M, N, K, A, lda, sA, B, ldb, sB, Cm, ldc, sC);
        HIPCHECK(hipGetLastError());
        return CHAMP_OK;
    }
    // strips of 16 x 80 columns per wavefront, 64 rows per block, when that
    // still gives at least two blocks per CU (tall problems such as the
    // orbitals of all electrons); otherwise one 16x16 tile per wavefront
    {
        constexpr int WN = 5, BN = 16 * WN;
        const long long gx = (M + 63) / 64, gy = (N + BN - 1) / BN;
        const bool fits = gx <= 0x7fffffffLL && gy <= 65535 && batch <= 65535;
        if (fits && (kind == GEMM_STRIP || (kind == GEMM_AUTO && gx * gy * batch >= 512))) {
            dim3 grid((unsigned)gx, (unsigned)gy, batch);
            if (bt)
                hipLaunchKernelGGL((k_gemm_mfma_strip<true, WN>), grid, dim3(256), 0, c->stream, M, N, K, A, lda, sA, B, ldb,
                                   sB, Cm, ldc, sC);
            else
                hipLaunchKernelGGL((k_gemm_mfma_strip<false, WN>), grid, dim3(256), 0, c->stream, M, N, K, A, lda, sA, B,
                                   ldb, sB, Cm, ldc, sC);
            HIPCHECK(hipGetLastError());
            return CHAMP_OK;
        }
    }
    // TWN 16 x 16 tiles along N per wavefront (A read once for all of them)
    constexpr int TWN = CHAMP_GEMM_TILE_WN;
    const long long waves = (long long)((M + 15) / 16) * ((N + 16 * TWN - 1) / (16 * TWN));
    const long long blocks = (waves + 3) / 4;
    if (blocks > 0x7fffffffLL || batch > 65535) {
        set_error("gemm: problem too large for one launch");
        return CHAMP_ERR_UNSUPPORTED;
    }
    dim3 grid((unsigned)blocks, batch);
    if (bt)
        hipLaunchKernelGGL((k_gemm_mfma<true, TWN>), grid, dim3(256), 0, c->stream, M, N, K, A, lda, sA, B, ldb, sB, Cm, ldc, sC);
    else
        hipLaunchKernelGGL((k_gemm_mfma<false, TWN>), grid, dim3(256), 0, c->stream, M, N, K, A, lda, sA, B, ldb, sB, Cm, ldc, sC);
    HIPCHECK(hipGetLastError());
    return CHAMP_OK;
}

}  // namespace champ
