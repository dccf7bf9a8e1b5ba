// Host-only stand-in for the rocBLAS API used by libqxhip's host code, for the
// sanitizer build (sim_device.cpp checks every operand's reach instead of running
// the product).  Only the names and signatures the host code uses.
#pragma once
#include <cstdint>

#include <hip/hip_runtime.h>

struct SimBlas;
typedef SimBlas *rocblas_handle;
typedef int32_t rocblas_int;
typedef int64_t rocblas_stride;
enum rocblas_status { rocblas_status_success = 0, rocblas_status_invalid_size = 3 };
enum rocblas_operation { rocblas_operation_none = 111, rocblas_operation_transpose = 112 };
enum rocblas_datatype { rocblas_datatype_f32_r = 151, rocblas_datatype_bf16_r = 168 };
enum rocblas_gemm_algo { rocblas_gemm_algo_standard = 0 };
struct rocblas_double_complex { double re, im; };
struct rocblas_float_complex { float re, im; };

rocblas_status rocblas_create_handle(rocblas_handle *h);
rocblas_status rocblas_destroy_handle(rocblas_handle h);
rocblas_status rocblas_set_workspace(rocblas_handle h, void *ws, size_t bytes);
rocblas_status rocblas_set_stream(rocblas_handle h, hipStream_t s);
const char *rocblas_status_to_string(rocblas_status s);
rocblas_status rocblas_zgemm_strided_batched(rocblas_handle h, rocblas_operation ta, rocblas_operation tb,
                                             rocblas_int m, rocblas_int n, rocblas_int k,
                                             const rocblas_double_complex *alpha, const rocblas_double_complex *A,
                                             rocblas_int lda, rocblas_stride sa, const rocblas_double_complex *B,
                                             rocblas_int ldb, rocblas_stride sb, const rocblas_double_complex *beta,
                                             rocblas_double_complex *C, rocblas_int ldc, rocblas_stride sc,
                                             rocblas_int batch);
rocblas_status rocblas_cgemm_strided_batched(rocblas_handle h, rocblas_operation ta, rocblas_operation tb,
                                             rocblas_int m, rocblas_int n, rocblas_int k,
                                             const rocblas_float_complex *alpha, const rocblas_float_complex *A,
                                             rocblas_int lda, rocblas_stride sa, const rocblas_float_complex *B,
                                             rocblas_int ldb, rocblas_stride sb, const rocblas_float_complex *beta,
                                             rocblas_float_complex *C, rocblas_int ldc, rocblas_stride sc,
                                             rocblas_int batch);
rocblas_status rocblas_gemm_strided_batched_ex(rocblas_handle h, rocblas_operation ta, rocblas_operation tb,
                                               rocblas_int m, rocblas_int n, rocblas_int k, const void *alpha,
                                               const void *A, rocblas_datatype at, rocblas_int lda, rocblas_stride sa,
                                               const void *B, rocblas_datatype bt, rocblas_int ldb, rocblas_stride sb,
                                               const void *beta, const void *C, rocblas_datatype ct, rocblas_int ldc,
                                               rocblas_stride sc, void *D, rocblas_datatype dt, rocblas_int ldd,
                                               rocblas_stride sd, rocblas_int batch, rocblas_datatype compute,
                                               rocblas_gemm_algo algo, int32_t sol, uint32_t flags);
