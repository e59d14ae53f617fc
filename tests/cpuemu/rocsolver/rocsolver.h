// TEST INFRASTRUCTURE: serial CPU stand-in for rocsolver_dsyevd (symmetric
// eigensolve, eigenvalues ascending, eigenvectors in the columns), for the
// CPU emulation build of tests/cpuemu: cyclic Jacobi rotations on the full
// symmetric matrix built from the requested triangle.  Loaded by tests only.
#pragma once
#include <algorithm>
#include <cmath>
#include <numeric>
#include <vector>

#include "../rocblas/rocblas.h"

inline rocblas_status hfx_emu_dsyevd(rocblas_handle, rocblas_evect evect, rocblas_fill uplo, rocblas_int n,
                                       double* A, rocblas_int lda, double* D, double* E, rocblas_int* info) {
  (void)E;
  std::vector<double> a((size_t)n * n), v((size_t)n * n, 0.0);
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i) {
      const bool upper = i <= j;   // element (i, j), column-major
      const double x = (uplo == rocblas_fill_upper) == upper ? A[(size_t)j * lda + i] : A[(size_t)i * lda + j];
      a[(size_t)j * n + i] = x;
    }
  for (int i = 0; i < n; ++i) v[(size_t)i * n + i] = 1.0;
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0.0;
    for (int j = 0; j < n; ++j)
      for (int i = 0; i < j; ++i) off += a[(size_t)j * n + i] * a[(size_t)j * n + i];
    if (off < 1e-30) break;
    for (int p = 0; p < n; ++p)
      for (int q = p + 1; q < n; ++q) {
        const double apq = a[(size_t)q * n + p];
        if (std::fabs(apq) < 1e-300) continue;
        const double app = a[(size_t)p * n + p], aqq = a[(size_t)q * n + q];
        const double theta = (aqq - app) / (2.0 * apq);
        const double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
        const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
        for (int k = 0; k < n; ++k) {   // columns p, q
          const double akp = a[(size_t)p * n + k], akq = a[(size_t)q * n + k];
          a[(size_t)p * n + k] = c * akp - s * akq;
          a[(size_t)q * n + k] = s * akp + c * akq;
        }
        for (int k = 0; k < n; ++k) {   // rows p, q
          const double apk = a[(size_t)k * n + p], aqk = a[(size_t)k * n + q];
          a[(size_t)k * n + p] = c * apk - s * aqk;
          a[(size_t)k * n + q] = s * apk + c * aqk;
        }
        for (int k = 0; k < n; ++k) {
          const double vkp = v[(size_t)p * n + k], vkq = v[(size_t)q * n + k];
          v[(size_t)p * n + k] = c * vkp - s * vkq;
          v[(size_t)q * n + k] = s * vkp + c * vkq;
        }
      }
  }
  std::vector<int> idx(n);
  std::iota(idx.begin(), idx.end(), 0);
  std::sort(idx.begin(), idx.end(), [&](int x, int y) { return a[(size_t)x * n + x] < a[(size_t)y * n + y]; });
  for (int j = 0; j < n; ++j) {
    D[j] = a[(size_t)idx[j] * n + idx[j]];
    if (evect == rocblas_evect_original)
      for (int i = 0; i < n; ++i) A[(size_t)j * lda + i] = v[(size_t)idx[j] * n + i];
  }
  *info = 0;
  return rocblas_status_success;
}

inline rocblas_status rocsolver_dsyevd(rocblas_handle h, rocblas_evect evect, rocblas_fill uplo, rocblas_int n,
                                       double* A, rocblas_int lda, double* D, double* E, rocblas_int* info) {
  hfx_emu_launch(1, [=]() { hfx_emu_dsyevd(h, evect, uplo, n, A, lda, D, E, info); });
  return rocblas_status_success;
}

inline rocblas_status rocsolver_dsyevd_strided_batched(rocblas_handle h, rocblas_evect evect, rocblas_fill uplo,
                                                       rocblas_int n, double* A, rocblas_int lda, rocblas_stride sA,
                                                       double* D, rocblas_stride sD, double* E, rocblas_stride sE,
                                                       rocblas_int* info, rocblas_int batch) {
  for (rocblas_int b = 0; b < batch; ++b)
    rocsolver_dsyevd(h, evect, uplo, n, A + b * sA, lda, D + b * sD, E + b * sE, info + b);
  return rocblas_status_success;
}
