// Dense factorizations for the node splits, FP64.
//
// * ptn_potrf_lower / ptn_trtri_lower: the small n x n Cholesky factor and its
//   inverse used by the CholeskyQR3 node split (Gram matrices and triangular
//   updates of that split run as gather-GEMMs).  One workgroup factors the
//   matrix in place in global memory (L2 resident for n <= 1024); the inverse
//   runs one workgroup per column.
// * ptn_qr / ptn_svd: Householder QR and thin SVD through rocSOLVER (AMD's
//   LAPACK on ROCm) with one lazily created handle per device -- the QR is the
//   fallback for rank-deficient panels, the SVD serves the truncated splits.
#include "ptn_common.h"

#include <mutex>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

namespace {

std::mutex g_mu;
rocblas_handle g_handles[64] = {nullptr};

// callers hold g_mu: it guards the handle table and each handle's stream binding
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

// ---- strided complex GEMM through rocBLAS ------------------------------------
struct OpLd {
  rocblas_operation op;
  rocblas_int ld;
};

// operand X, logically R x Cc with element strides (sr, sc): as a column-major
// matrix (op none, ld = sc) or as the transpose of one (op T / C, ld = sr)
bool pick_op(int64_t R, int64_t Cc, int64_t sr, int64_t sc, bool conj, OpLd *o) {
  const int64_t r1 = R > 1 ? R : 1, c1 = Cc > 1 ? Cc : 1;
  if (!conj && (sr == 1 || R == 1)) {
    const int64_t ld = Cc > 1 ? sc : r1;
    if (ld >= r1 && ld <= INT32_MAX) {
      o->op = rocblas_operation_none;
      o->ld = (rocblas_int)ld;
      return true;
    }
  }
  if (sc == 1 || Cc == 1) {
    const int64_t ld = R > 1 ? sr : c1;
    if (ld >= c1 && ld <= INT32_MAX) {
      o->op = conj ? rocblas_operation_conjugate_transpose : rocblas_operation_transpose;
      o->ld = (rocblas_int)ld;
      return true;
    }
  }
  return false;
}

struct StridedPlan {
  bool swap;  // false: C = op(A) op(B) column-major; true: C^T = op(B^T) op(A^T)
  OpLd a, b;
  rocblas_int ldc;
};

bool plan_strided(const ptn_gemm_strided_desc *d, StridedPlan *p) {
  if (!d || d->M < 1 || d->N < 1 || d->K < 1 || d->M > INT32_MAX || d->N > INT32_MAX ||
      d->K > INT32_MAX || !d->A || !d->B || !d->C ||
      (d->dtype != PTN_C128 && d->dtype != PTN_C64))
    return false;
  const int64_t M = d->M, N = d->N, K = d->K;
  if (d->c_sm == 1 || M == 1) {
    const int64_t ldc = N > 1 ? d->c_sn : M;
    if (ldc >= M && ldc <= INT32_MAX && pick_op(M, K, d->a_sm, d->a_sk, d->a_conj, &p->a) &&
        pick_op(K, N, d->b_sk, d->b_sn, d->b_conj, &p->b)) {
      p->swap = false;
      p->ldc = (rocblas_int)ldc;
      return true;
    }
  }
  if (d->c_sn == 1 || N == 1) {
    const int64_t ldc = M > 1 ? d->c_sm : N;
    if (ldc >= N && ldc <= INT32_MAX && pick_op(N, K, d->b_sn, d->b_sk, d->b_conj, &p->b) &&
        pick_op(K, M, d->a_sk, d->a_sm, d->a_conj, &p->a)) {
      p->swap = true;
      p->ldc = (rocblas_int)ldc;
      return true;
    }
  }
  return false;
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

// ---- Cholesky (lower, in place), optional diagonal shift = shift_factor * trace
constexpr int CT = 1024;

__global__ __launch_bounds__(CT) void potrf_lower_kernel(int n, const double2 *G, int64_t ldg,
                                                         double2 *L, double shift_factor,
                                                         double min_pivot, int *status) {
  __shared__ double2 colj[1024];
  __shared__ double red[CT / 64];
  __shared__ double sh_shift, sh_inv;
  __shared__ int sh_fail;
  const int tid = threadIdx.x;
  // trace for the shift
  double tr = 0;
  for (int i = tid; i < n; i += CT) tr += G[(int64_t)i * ldg + i].x;
  for (int off = 32; off > 0; off >>= 1) tr += __shfl_down(tr, off, 64);
  if ((tid & 63) == 0) red[tid >> 6] = tr;
  __syncthreads();
  if (tid == 0) {
    double t = 0;
    for (int w = 0; w < CT / 64; ++w) t += red[w];
    sh_shift = shift_factor * t;
    sh_fail = 0;
  }
  __syncthreads();
  const int64_t nn = (int64_t)n * n;
  for (int64_t t = tid; t < nn; t += CT) {
    const int i = (int)(t / n), j = (int)(t - (int64_t)i * n);
    double2 v = make_double2(0, 0);
    if (j <= i) {
      v = G[(int64_t)i * ldg + j];
      if (i == j) v = make_double2(v.x + sh_shift, 0.0);
    }
    L[t] = v;
  }
  __syncthreads();
  for (int j = 0; j < n; ++j) {
    if (tid == 0) {
      double d = L[(int64_t)j * n + j].x;
      if (!(d > min_pivot)) {  // also catches NaN
        sh_fail = 1;
        d = 1.0;
      }
      const double r = sqrt(d);
      L[(int64_t)j * n + j] = make_double2(r, 0.0);
      sh_inv = 1.0 / r;
    }
    __syncthreads();
    const double inv = sh_inv;
    for (int i = j + 1 + tid; i < n; i += CT) {
      double2 v = L[(int64_t)i * n + j];
      v.x *= inv;
      v.y *= inv;
      L[(int64_t)i * n + j] = v;
      colj[i] = v;
    }
    __syncthreads();
    const int rem = n - j - 1;
    const int64_t cnt = (int64_t)rem * rem;
    for (int64_t t = tid; t < cnt; t += CT) {
      const int i = j + 1 + (int)(t / rem), k = j + 1 + (int)(t % rem);
      if (k > i) continue;
      const double2 a = colj[i], b = colj[k];
      double2 v = L[(int64_t)i * n + k];
      // v -= a * conj(b)
      v.x -= a.x * b.x + a.y * b.y;
      v.y -= a.y * b.x - a.x * b.y;
      L[(int64_t)i * n + k] = v;
    }
    __syncthreads();
  }
  if (tid == 0) *status = sh_fail;
}

// X = L^{-1} (lower); block j solves L x = e_j by forward substitution
constexpr int TT = 256;
__global__ __launch_bounds__(TT) void trtri_lower_kernel(int n, const double2 *L, double2 *X) {
  __shared__ double2 x[1024];
  __shared__ double2 red[TT / 64];
  const int j = blockIdx.x, tid = threadIdx.x;
  for (int i = tid; i < n; i += TT) x[i] = make_double2(0, 0);
  __syncthreads();
  if (tid == 0) {
    const double2 d = L[(int64_t)j * n + j];
    const double den = d.x * d.x + d.y * d.y;
    x[j] = make_double2(d.x / den, -d.y / den);
  }
  __syncthreads();
  for (int i = j + 1; i < n; ++i) {
    double sr = 0, si = 0;
    for (int k = j + tid; k < i; k += TT) {
      const double2 l = L[(int64_t)i * n + k], xk = x[k];
      sr += l.x * xk.x - l.y * xk.y;
      si += l.x * xk.y + l.y * xk.x;
    }
    for (int off = 32; off > 0; off >>= 1) {
      sr += __shfl_down(sr, off, 64);
      si += __shfl_down(si, off, 64);
    }
    if ((tid & 63) == 0) red[tid >> 6] = make_double2(sr, si);
    __syncthreads();
    if (tid == 0) {
      double tr = 0, ti = 0;
      for (int w = 0; w < TT / 64; ++w) {
        tr += red[w].x;
        ti += red[w].y;
      }
      const double2 d = L[(int64_t)i * n + i];
      const double den = d.x * d.x + d.y * d.y;
      // x_i = -(sum) / d
      x[i] = make_double2(-(tr * d.x + ti * d.y) / den, -(ti * d.x - tr * d.y) / den);
    }
    __syncthreads();
  }
  for (int i = tid; i < n; i += TT) X[(int64_t)i * n + j] = x[i];
}

}  // namespace

extern "C" int ptn_potrf_lower(void *stream, int32_t n, const void *G, int64_t ldg, void *L,
                               double shift_factor, double min_pivot, void *status) {
  if (n <= 0 || n > 1024 || ldg < n || !G || !L || !status) {
    ptn::set_error("ptn_potrf_lower: bad arguments (n=%d, max 1024)", n);
    return n > 1024 ? PTN_ERANGE : PTN_EINVAL;
  }
  hipStream_t s = static_cast<hipStream_t>(stream);
  hipLaunchKernelGGL(potrf_lower_kernel, dim3(1), dim3(CT), 0, s, n,
                     static_cast<const double2 *>(G), ldg, static_cast<double2 *>(L),
                     shift_factor, min_pivot, static_cast<int *>(status));
  PTN_LAUNCH_CHECK();
  return PTN_OK;
}

extern "C" int ptn_trtri_lower(void *stream, int32_t n, const void *L, void *X) {
  if (n <= 0 || n > 1024 || !L || !X) {
    ptn::set_error("ptn_trtri_lower: bad arguments (n=%d, max 1024)", n);
    return n > 1024 ? PTN_ERANGE : PTN_EINVAL;
  }
  hipStream_t s = static_cast<hipStream_t>(stream);
  hipLaunchKernelGGL(trtri_lower_kernel, dim3(n), dim3(TT), 0, s, n,
                     static_cast<const double2 *>(L), static_cast<double2 *>(X));
  PTN_LAUNCH_CHECK();
  return PTN_OK;
}

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

extern "C" int ptn_gemm_strided_ok(const ptn_gemm_strided_desc *d) {
  StridedPlan p;
  return plan_strided(d, &p) ? 1 : 0;
}

extern "C" int ptn_gemm_strided(void *stream, const ptn_gemm_strided_desc *d) {
  StridedPlan p;
  if (!plan_strided(d, &p)) {
    ptn::set_error("ptn_gemm_strided: layout not expressible as one strided GEMM");
    return PTN_EINVAL;
  }
  hipStream_t s = static_cast<hipStream_t>(stream);
  std::lock_guard<std::mutex> lk(g_mu);
  rocblas_handle h;
  int rc = get_handle(s, &h);
  if (rc) return rc;
  const rocblas_int M = (rocblas_int)d->M, N = (rocblas_int)d->N, K = (rocblas_int)d->K;
  rocblas_status st;
  if (d->dtype == PTN_C128) {
    const rocblas_double_complex alpha(d->alpha_re, d->alpha_im);
    const rocblas_double_complex beta(d->beta_one ? 1.0 : 0.0, 0.0);
    auto *A = static_cast<const rocblas_double_complex *>(d->A);
    auto *B = static_cast<const rocblas_double_complex *>(d->B);
    auto *C = static_cast<rocblas_double_complex *>(d->C);
    st = p.swap ? rocblas_zgemm(h, p.b.op, p.a.op, N, M, K, &alpha, B, p.b.ld, A, p.a.ld, &beta,
                                C, p.ldc)
                : rocblas_zgemm(h, p.a.op, p.b.op, M, N, K, &alpha, A, p.a.ld, B, p.b.ld, &beta,
                                C, p.ldc);
  } else {
    const rocblas_float_complex alpha((float)d->alpha_re, (float)d->alpha_im);
    const rocblas_float_complex beta(d->beta_one ? 1.0f : 0.0f, 0.0f);
    auto *A = static_cast<const rocblas_float_complex *>(d->A);
    auto *B = static_cast<const rocblas_float_complex *>(d->B);
    auto *C = static_cast<rocblas_float_complex *>(d->C);
    st = p.swap ? rocblas_cgemm(h, p.b.op, p.a.op, N, M, K, &alpha, B, p.b.ld, A, p.a.ld, &beta,
                                C, p.ldc)
                : rocblas_cgemm(h, p.a.op, p.b.op, M, N, K, &alpha, A, p.a.ld, B, p.b.ld, &beta,
                                C, p.ldc);
  }
  if (st != rocblas_status_success) {
    ptn::set_error("ptn_gemm_strided: rocblas gemm failed (%d)", (int)st);
    return PTN_ESOLVER;
  }
  return PTN_OK;
}
