// CPU emulation of the two rocBLAS calls the engine makes (column-major
// dgemm and strided-batched dgemm).  Test infrastructure only.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

typedef struct emu_rocblas *rocblas_handle;
typedef int32_t rocblas_int;
typedef int64_t rocblas_stride;
typedef enum rocblas_operation { rocblas_operation_none = 111, rocblas_operation_transpose = 112 } rocblas_operation;
typedef enum rocblas_status { rocblas_status_success = 0, rocblas_status_invalid_size = 3 } rocblas_status;

rocblas_status rocblas_create_handle(rocblas_handle *h);
rocblas_status rocblas_destroy_handle(rocblas_handle h);
rocblas_status rocblas_set_stream(rocblas_handle h, hipStream_t s);
rocblas_status rocblas_dgemm(rocblas_handle h, rocblas_operation ta, rocblas_operation tb, rocblas_int m, rocblas_int n,
                             rocblas_int k, const double *alpha, const double *A, rocblas_int lda, const double *B,
                             rocblas_int ldb, const double *beta, double *C, rocblas_int ldc);
rocblas_status rocblas_dgemm_strided_batched(rocblas_handle h, rocblas_operation ta, rocblas_operation tb,
                                             rocblas_int m, rocblas_int n, rocblas_int k, const double *alpha,
                                             const double *A, rocblas_int lda, rocblas_stride sa, const double *B,
                                             rocblas_int ldb, rocblas_stride sb, const double *beta, double *C,
                                             rocblas_int ldc, rocblas_stride sc, rocblas_int batch);
