// TEST INFRASTRUCTURE: serial CPU stand-in for the few rocBLAS entry points
// libhfx's SCF driver calls (column-major dgemm), for the CPU emulation build
// of tests/cpuemu.  Loaded by tests only.
#pragma once
#include <cstdint>
#include <vector>

#include "../hip/hip_runtime.h"

typedef int rocblas_int;
typedef int64_t rocblas_stride;
typedef enum { rocblas_status_success = 0, rocblas_status_invalid_value = 3 } rocblas_status;
typedef enum { rocblas_operation_none = 111, rocblas_operation_transpose = 112 } rocblas_operation;
typedef enum { rocblas_fill_upper = 121, rocblas_fill_lower = 122 } rocblas_fill;
typedef enum { rocblas_evect_original = 211, rocblas_evect_none = 213 } rocblas_evect;
typedef struct hfx_emu_blas* rocblas_handle;
struct hfx_emu_blas {
  int dummy;
};

inline rocblas_status rocblas_create_handle(rocblas_handle* h) {
  *h = new hfx_emu_blas{0};
  return rocblas_status_success;
}
inline rocblas_status rocblas_destroy_handle(rocblas_handle h) {
  delete h;
  return rocblas_status_success;
}
inline rocblas_status rocblas_set_stream(rocblas_handle, void*) { return rocblas_status_success; }

// C = alpha op(A) op(B) + beta C, column-major
inline rocblas_status hfx_emu_dgemm(rocblas_handle, rocblas_operation ta, rocblas_operation tb, rocblas_int m,
                                    rocblas_int n, rocblas_int k, const double* alpha, const double* A,
                                    rocblas_int lda, const double* B, rocblas_int ldb, const double* beta, double* C,
                                    rocblas_int ldc) {
  std::vector<double> out((size_t)m * n);
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < m; ++i) {
      double s = 0.0;
      for (int l = 0; l < k; ++l) {
        const double a = ta == rocblas_operation_none ? A[(size_t)l * lda + i] : A[(size_t)i * lda + l];
        const double b = tb == rocblas_operation_none ? B[(size_t)j * ldb + l] : B[(size_t)l * ldb + j];
        s += a * b;
      }
      out[(size_t)j * m + i] = *alpha * s + (*beta == 0.0 ? 0.0 : *beta * C[(size_t)j * ldc + i]);
    }
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < m; ++i) C[(size_t)j * ldc + i] = out[(size_t)j * m + i];
  return rocblas_status_success;
}

// recorded like a kernel while a stream is captured (see hip/hip_runtime.h)
inline rocblas_status rocblas_dgemm(rocblas_handle h, rocblas_operation ta, rocblas_operation tb, rocblas_int m,
                                    rocblas_int n, rocblas_int k, const double* alpha, const double* A,
                                    rocblas_int lda, const double* B, rocblas_int ldb, const double* beta, double* C,
                                    rocblas_int ldc) {
  const double al = *alpha, be = *beta;
  hfx_emu_launch(1, [=]() { hfx_emu_dgemm(h, ta, tb, m, n, k, &al, A, lda, B, ldb, &be, C, ldc); });
  return rocblas_status_success;
}

inline rocblas_status rocblas_dgemm_strided_batched(rocblas_handle h, rocblas_operation ta, rocblas_operation tb,
                                                    rocblas_int m, rocblas_int n, rocblas_int k, const double* alpha,
                                                    const double* A, rocblas_int lda, rocblas_stride sA,
                                                    const double* B, rocblas_int ldb, rocblas_stride sB,
                                                    const double* beta, double* C, rocblas_int ldc, rocblas_stride sC,
                                                    rocblas_int batch) {
  for (rocblas_int b = 0; b < batch; ++b)
    rocblas_dgemm(h, ta, tb, m, n, k, alpha, A + b * sA, lda, B + b * sB, ldb, beta, C + b * sC, ldc);
  return rocblas_status_success;
}
