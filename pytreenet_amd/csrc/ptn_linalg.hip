// Dense factorizations for the node splits: reduced QR and thin SVD in FP64.
//
// Both run on the caller's stream through rocSOLVER (AMD's LAPACK on ROCm)
// with one lazily created handle per device.  R and the packed singular
// vectors are produced by small kernels here.
#include "ptn_common.h"

#include <mutex>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

namespace {

std::mutex g_mu;
rocblas_handle g_handles[64] = {nullptr};

int get_handle(hipStream_t s, rocblas_handle *out) {
  int dev = 0;
  PTN_HIP_CHECK(hipGetDevice(&dev));
  if (dev < 0 || dev >= 64) {
    ptn::set_error("device index out of range");
    return PTN_ERANGE;
  }
  if (!g_handles[dev]) {
    if (rocblas_create_handle(&g_handles[dev]) != rocblas_status_success) {
      ptn::set_error("rocblas_create_handle failed");
      return PTN_ESOLVER;
    }
  }
  if (rocblas_set_stream(g_handles[dev], s) != rocblas_status_success) {
    ptn::set_error("rocblas_set_stream failed");
    return PTN_ESOLVER;
  }
  *out = g_handles[dev];
  return PTN_OK;
}

// R (k x n col-major) = upper triangle of A (m x n col-major, lda = m), k = min(m, n)
__global__ void extract_r(int64_t m, int64_t n, int64_t k, const double2 *A, double2 *R) {
  const int64_t total = k * n;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += stride) {
    const int64_t i = t % k, j = t / k;
    R[t] = (i <= j) ? A[j * m + i] : make_double2(0, 0);
  }
}

unsigned grid_for(int64_t total) {
  int64_t g = (total + 255) / 256;
  if (g < 1) g = 1;
  if (g > 65536) g = 65536;
  return (unsigned)g;
}

}  // namespace

extern "C" int ptn_qr(void *stream, int64_t m, int64_t n, void *A, void *R, void *tau) {
  if (m <= 0 || n <= 0 || !A || !R || !tau || m > 0x7fffffff || n > 0x7fffffff) {
    ptn::set_error("ptn_qr: bad arguments (m=%lld n=%lld)", (long long)m, (long long)n);
    return PTN_EINVAL;
  }
  hipStream_t s = static_cast<hipStream_t>(stream);
  std::lock_guard<std::mutex> lk(g_mu);
  rocblas_handle h;
  int rc = get_handle(s, &h);
  if (rc) return rc;
  const int64_t k = m < n ? m : n;
  auto *a = static_cast<rocblas_double_complex *>(A);
  auto *t = static_cast<rocblas_double_complex *>(tau);
  if (rocsolver_zgeqrf(h, (rocblas_int)m, (rocblas_int)n, a, (rocblas_int)m, t) !=
      rocblas_status_success) {
    ptn::set_error("rocsolver_zgeqrf failed");
    return PTN_ESOLVER;
  }
  hipLaunchKernelGGL(extract_r, dim3(grid_for(k * n)), dim3(256), 0, s, m, n, k,
                     static_cast<const double2 *>(A), static_cast<double2 *>(R));
  PTN_LAUNCH_CHECK();
  if (rocsolver_zungqr(h, (rocblas_int)m, (rocblas_int)k, (rocblas_int)k, a, (rocblas_int)m,
                       t) != rocblas_status_success) {
    ptn::set_error("rocsolver_zungqr failed");
    return PTN_ESOLVER;
  }
  return PTN_OK;
}

extern "C" int ptn_svd(void *stream, int64_t m, int64_t n, void *A, void *S, void *U, void *Vh,
                       void *E) {
  if (m <= 0 || n <= 0 || !A || !S || !U || !Vh || !E || m > 0x7fffffff || n > 0x7fffffff) {
    ptn::set_error("ptn_svd: bad arguments (m=%lld n=%lld)", (long long)m, (long long)n);
    return PTN_EINVAL;
  }
  hipStream_t s = static_cast<hipStream_t>(stream);
  std::lock_guard<std::mutex> lk(g_mu);
  rocblas_handle h;
  int rc = get_handle(s, &h);
  if (rc) return rc;
  const int64_t k = m < n ? m : n;
  // E holds k doubles of superdiagonal workspace followed by the info word
  double *e = static_cast<double *>(E);
  rocblas_int *info = reinterpret_cast<rocblas_int *>(e + k);
  if (rocsolver_zgesvd(h, rocblas_svect_singular, rocblas_svect_singular, (rocblas_int)m,
                       (rocblas_int)n, static_cast<rocblas_double_complex *>(A),
                       (rocblas_int)m, static_cast<double *>(S),
                       static_cast<rocblas_double_complex *>(U), (rocblas_int)m,
                       static_cast<rocblas_double_complex *>(Vh), (rocblas_int)k, e,
                       rocblas_outofplace, info) != rocblas_status_success) {
    ptn::set_error("rocsolver_zgesvd failed");
    return PTN_ESOLVER;
  }
  return PTN_OK;
}
