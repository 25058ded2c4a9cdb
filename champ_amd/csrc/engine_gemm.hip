// engine_gemm.hip -- population-wide FP64 contractions of the engine:
//   AO -> MO (orbitals / orbitalse, orbitals_no_qmckl.f90:73-128, 245-258):
//       orb[r][o] = sum_m ao[r][m] * coef[o][m]        (B transposed, "NT")
//   inverse rows contracted with the occupied MO coefficients for the
//   quadrature ratios (nonlocd_quad, nonloc.f90:462-506, reassociated):
//       wmat[b][i][m] = sum_j slmi[b][i][j] * cocc[j][m] ("NN", batched)
// One LDS-tiled kernel: 64x64 output tile per 256-thread block, 4x4 per
// thread, K staged through LDS in slabs of 16.
#include "engine_internal.hpp"

namespace champ {

#define GB_M 64
#define GB_N 64
#define GB_K 16

template <bool BT>
__global__ void __launch_bounds__(256) k_gemm(int M, int N, int K, const double *__restrict__ A, int lda,
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

// C[b] (M x N, row-major, ldc) = A[b] (M x K, lda) * op(B[b]); op = B^T when
// bt (B stored N x K) else B (K x N).
int launch_gemm(champ_ctx *c, bool bt, int M, int N, int K, const double *A, int lda, long long sA, const double *B,
                int ldb, long long sB, double *Cm, int ldc, long long sC, int batch)
{
    if (M <= 0 || N <= 0 || batch <= 0) return CHAMP_OK;
    dim3 grid((M + GB_M - 1) / GB_M, (N + GB_N - 1) / GB_N, batch);
    if (bt)
        hipLaunchKernelGGL(k_gemm<true>, grid, dim3(256), 0, c->stream, M, N, K, A, lda, sA, B, ldb, sB, Cm, ldc, sC);
    else
        hipLaunchKernelGGL(k_gemm<false>, grid, dim3(256), 0, c->stream, M, N, K, A, lda, sA, B, ldb, sB, Cm, ldc, sC);
    HIPCHECK(hipGetLastError());
    return CHAMP_OK;
}

}  // namespace champ
