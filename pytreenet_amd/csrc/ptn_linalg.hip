// Dense factorizations for the node splits, FP64.
//
// * ptn_potrf_lower / ptn_trtri_lower: the small n x n Cholesky factor and its
//   inverse used by the CholeskyQR3 node split (Gram matrices and triangular
//   updates of that split run as gather-GEMMs).  n <= 16: one workgroup factors
//   the matrix in global memory (L2 resident); above, one launch per block column
//   of 16 with one workgroup per block row (left-looking).  The inverse runs one
//   workgroup per 16 columns (n <= 512) or per column.
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

// Blocked right-looking Cholesky in one workgroup (the working matrix L lives in
// global memory, L2 resident for n <= 1024): for each block column of PB
// columns, (a) the PB x PB diagonal block is factored in LDS, (b) the panel below
// it is solved against that block, one thread per row (x L_JJ^H = a), and (c) the
// trailing lower triangle is updated, L22 -= L21 L21^H, in 4x4 register tiles
// spread over all threads.  Same shift, pivot test and status as the unblocked
// kernel above.
constexpr int PB = 16;
constexpr int PT = 512;   // threads of the blocked kernel (256 VGPRs per lane)
constexpr int PSTAGE = 4096;   // panel elements staged in LDS (rows <= 256)

__global__ __launch_bounds__(PT) void potrf_blocked_kernel(int n, const double2 *G, int64_t ldg,
                                                           double2 *L, double shift_factor,
                                                           double min_pivot, int *status) {
  __shared__ double2 D[PB][PB + 1];
  __shared__ double2 Pn[PSTAGE];
  __shared__ double red[PT / 64];
  __shared__ double sh_shift;
  __shared__ int sh_fail;
  const int tid = threadIdx.x;
  double tr = 0;
  for (int i = tid; i < n; i += PT) tr += G[(int64_t)i * ldg + i].x;
  for (int off = 32; off > 0; off >>= 1) tr += __shfl_down(tr, off, 64);
  if ((tid & 63) == 0) red[tid >> 6] = tr;
  __syncthreads();
  if (tid == 0) {
    double t = 0;
    for (int w = 0; w < PT / 64; ++w) t += red[w];
    sh_shift = shift_factor * t;
    sh_fail = 0;
  }
  __syncthreads();
  const int64_t nn = (int64_t)n * n;
  for (int64_t t = tid; t < nn; t += PT) {
    const int i = (int)(t / n), j = (int)(t - (int64_t)i * n);
    double2 v = make_double2(0, 0);
    if (j <= i) {
      v = G[(int64_t)i * ldg + j];
      if (i == j) v = make_double2(v.x + sh_shift, 0.0);
    }
    L[t] = v;
  }
  __syncthreads();
  for (int J = 0; J < n; J += PB) {
    const int jb = (n - J) < PB ? (n - J) : PB;
    // (a) diagonal block, factorised by the first 16 lanes of wave 0 with no
    //     block barrier: lane i holds row i in registers, the pivot and the
    //     entries of column j are broadcast by 16-wide shuffles
    if (tid < PB) {
      const int i = tid;
      double2 row[PB];
#pragma unroll
      for (int k = 0; k < PB; ++k)
        row[k] = (i < jb && k <= i) ? L[(int64_t)(J + i) * n + J + k] : make_double2(0, 0);
      int fail = 0;
#pragma unroll
      for (int j = 0; j < PB; ++j) {
        const bool act = j < jb;  // columns past a short last block do nothing
        double d = __shfl(row[j].x, j, PB);
        if (act && !(d > min_pivot)) {  // also catches NaN
          fail = 1;
          d = 1.0;
        }
        if (!act) d = 1.0;
        const double r = sqrt(d), inv = 1.0 / r;
        if (act && i == j) row[j] = make_double2(r, 0.0);
        if (act && i > j) {
          row[j].x *= inv;
          row[j].y *= inv;
        }
#pragma unroll
        for (int k = j + 1; k < PB; ++k) {
          // D[k][j] from lane k, then D[i][k] -= D[i][j] conj(D[k][j]) for k <= i
          const double br = __shfl(row[j].x, k, PB), bi = __shfl(row[j].y, k, PB);
          if (act && k <= i && i < jb) {
            row[k].x -= row[j].x * br + row[j].y * bi;
            row[k].y -= row[j].y * br - row[j].x * bi;
          }
        }
      }
#pragma unroll
      for (int k = 0; k < PB; ++k) {
        const double2 v = (i < jb && k <= i) ? row[k] : make_double2(0, 0);
        D[i][k] = v;
        if (i < jb && k <= i) L[(int64_t)(J + i) * n + J + k] = v;
      }
      if (fail) sh_fail = 1;
    }
    __syncthreads();
    // (b) panel rows below the block: x L_JJ^H = a, i.e. x_k = (a_k - sum_{l<k} x_l
    //     conj(L_kl)) / L_kk.  Sixteen lanes per row, lane k owns x_k; x_l is
    //     broadcast within the 16-lane group once it is final.
    //     Each 16-lane group solves PRB rows together (their loads in flight at
    //     once); lane lk keeps row lk of the block and its pivot inverse in
    //     registers.
    {
      const int lk = tid & (PB - 1), grp = tid / PB;
      constexpr int PRB = 8, GROUPS = PT / PB;
      const double dinv = lk < jb ? 1.0 / D[lk][lk].x : 0.0;
      double2 drow[PB];
#pragma unroll
      for (int l = 0; l < PB; ++l) drow[l] = D[lk][l];   // zero outside the block
      for (int i0 = J + jb; i0 < n; i0 += PRB * GROUPS) {
        double2 v[PRB];
#pragma unroll
        for (int rb = 0; rb < PRB; ++rb) {
          const int i = i0 + rb * GROUPS + grp;
          v[rb] = (i < n && lk < jb) ? L[(int64_t)i * n + J + lk] : make_double2(0, 0);
        }
#pragma unroll
        for (int l = 0; l < PB; ++l) {
          if (l >= jb) continue;  // uniform
#pragma unroll
          for (int rb = 0; rb < PRB; ++rb) {
            if (lk == l) {
              v[rb].x *= dinv;
              v[rb].y *= dinv;
            }
            const double xr = __shfl(v[rb].x, l, PB), xi = __shfl(v[rb].y, l, PB);
            if (lk > l) {
              v[rb].x -= xr * drow[l].x + xi * drow[l].y;
              v[rb].y -= xi * drow[l].x - xr * drow[l].y;
            }
          }
        }
#pragma unroll
        for (int rb = 0; rb < PRB; ++rb) {
          const int i = i0 + rb * GROUPS + grp;
          if (i < n && lk < jb) L[(int64_t)i * n + J + lk] = v[rb];
        }
      }
    }
    __syncthreads();
    // (c) trailing update of the lower triangle, 4x4 tiles.  The panel columns
    //     (rows below the block) are staged in LDS when they fit, so the tile
    //     loops read LDS instead of L2; columns past jb are zero
    const int r0 = J + jb, rem = n - r0;
    const bool staged = rem > 0 && rem * PB <= PSTAGE;
    if (staged) {
      for (int t = tid; t < rem * PB; t += PT) {
        const int r = t / PB, l = t % PB;
        Pn[t] = l < jb ? L[(int64_t)(r0 + r) * n + J + l] : make_double2(0, 0);
      }
      __syncthreads();
    }
    if (rem > 0) {
      const int nt = (rem + 3) / 4;
      const int ntiles = nt * (nt + 1) / 2;
      for (int t = tid; t < ntiles; t += PT) {
        int ti = (int)((sqrt(8.0 * (double)t + 1.0) - 1.0) * 0.5);
        while ((ti + 1) * (ti + 2) / 2 <= t) ++ti;
        while (ti * (ti + 1) / 2 > t) --ti;
        const int tk = t - ti * (ti + 1) / 2;
        double2 acc[4][4];
#pragma unroll
        for (int r = 0; r < 4; ++r)
#pragma unroll
          for (int c = 0; c < 4; ++c) acc[r][c] = make_double2(0, 0);
        int rows[4], cols[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          rows[r] = r0 + ti * 4 + r;
          cols[r] = r0 + tk * 4 + r;
        }

        auto madd = [&](const double2 (&a)[4], const double2 (&b)[4]) {
#pragma unroll
          for (int r = 0; r < 4; ++r)
#pragma unroll
            for (int c = 0; c < 4; ++c) {
              acc[r][c].x += a[r].x * b[c].x + a[r].y * b[c].y;
              acc[r][c].y += a[r].y * b[c].x - a[r].x * b[c].y;
            }
        };
        if (staged) {
#pragma unroll 2
          for (int l = 0; l < PB; ++l) {
            double2 a[4], b[4];
#pragma unroll
            for (int r = 0; r < 4; ++r) {
              a[r] = rows[r] < n ? Pn[(rows[r] - r0) * PB + l] : make_double2(0, 0);
              b[r] = cols[r] < n ? Pn[(cols[r] - r0) * PB + l] : make_double2(0, 0);
            }
            madd(a, b);
          }
        } else {
          for (int l = 0; l < jb; ++l) {
            double2 a[4], b[4];
#pragma unroll
            for (int r = 0; r < 4; ++r) {
              a[r] = rows[r] < n ? L[(int64_t)rows[r] * n + J + l] : make_double2(0, 0);
              b[r] = cols[r] < n ? L[(int64_t)cols[r] * n + J + l] : make_double2(0, 0);
            }
            madd(a, b);
          }
        }
        // all 16 current values are loaded before any store, so the loads fly
        // together instead of each waiting behind the previous store
        double2 cur[4][4];
#pragma unroll
        for (int r = 0; r < 4; ++r)
#pragma unroll
          for (int c = 0; c < 4; ++c)
            cur[r][c] = (rows[r] < n && cols[c] <= rows[r])
                            ? L[(int64_t)rows[r] * n + cols[c]] : make_double2(0, 0);
#pragma unroll
        for (int r = 0; r < 4; ++r)
#pragma unroll
          for (int c = 0; c < 4; ++c) {
            if (rows[r] < n && cols[c] <= rows[r])
              L[(int64_t)rows[r] * n + cols[c]] =
                  make_double2(cur[r][c].x - acc[r][c].x, cur[r][c].y - acc[r][c].y);
          }
      }
    }
    __syncthreads();
  }
  if (tid == 0) *status = sh_fail;
}

// Left-looking blocked Cholesky: one launch per block column J of CB columns, one
// workgroup per block row I >= J (workgroup 0: the diagonal block).  Each
// workgroup forms its blocks from G minus the products of the finished panels
// L[., 0..J) (staged through LDS in chunks of CKC columns), factors the CB x CB
// diagonal block in one wave (lane i holds row i; pivots and columns pass by
// shuffles; every workgroup does this, so none waits for another), and solves its
// own panel block X L_JJ^H = A_IJ, 32 lanes per row.  The matrix is spread over
// up to n / CB CUs per launch instead of one; same shift, pivot test and status as
// the kernels above.
constexpr int CKC = 32;
template <int CB>
__global__ __launch_bounds__(256) void potrf_left_kernel(int n, int J, const double2 *G,
                                                         int64_t ldg, double2 *L,
                                                         double shift_factor,
                                                         double min_pivot, int *status) {
  __shared__ double2 D[CB][CB + 1];
  __shared__ double2 P[CB][CB + 1];
  __shared__ double2 Lj[CB][CKC + 1];
  __shared__ double2 Li[CB][CKC + 1];
  __shared__ double red[4];
  __shared__ double sh_shift;
  __shared__ int sh_fail;
  const int tid = threadIdx.x;
  const int I = J + blockIdx.x * CB;
  const bool diag = blockIdx.x == 0;
  const int jb = (n - J) < CB ? (n - J) : CB;
  const int ib = (n - I) < CB ? (n - I) : CB;
  double tr = 0;
  if (shift_factor != 0.0)
    for (int i = tid; i < n; i += 256) tr += G[(int64_t)i * ldg + i].x;
  for (int off = 32; off > 0; off >>= 1) tr += __shfl_down(tr, off, 64);
  if ((tid & 63) == 0) red[tid >> 6] = tr;
  __syncthreads();
  if (tid == 0) {
    sh_shift = shift_factor * (red[0] + red[1] + red[2] + red[3]);
    sh_fail = 0;
  }
  __syncthreads();
  // thread tid owns elements (tid / CB + RS q, tid % CB), q < EPT, of D and of P
  constexpr int RS = 256 / CB, EPT = CB * CB / 256;
  double2 aD[EPT], aP[EPT];
#pragma unroll
  for (int q = 0; q < EPT; ++q) {
    const int r = tid / CB + RS * q, c = tid % CB;
    double2 v = make_double2(0, 0);
    if (r < jb && c <= r) {
      v = G[(int64_t)(J + r) * ldg + J + c];
      if (r == c) v = make_double2(v.x + sh_shift, 0.0);
    }
    aD[q] = v;
    aP[q] = (!diag && r < ib && c < jb) ? G[(int64_t)(I + r) * ldg + J + c] : make_double2(0, 0);
  }
  for (int k0 = 0; k0 < J; k0 += CKC) {
    for (int t = tid; t < CB * CKC; t += 256) {
      const int r = t / CKC, kk = t % CKC;
      const bool in = k0 + kk < J;   // only the finished columns
      Lj[r][kk] = (in && r < jb) ? L[(int64_t)(J + r) * n + k0 + kk] : make_double2(0, 0);
      Li[r][kk] = (in && !diag && r < ib) ? L[(int64_t)(I + r) * n + k0 + kk]
                                          : make_double2(0, 0);
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < EPT; ++q) {
      const int r = tid / CB + RS * q, c = tid % CB;
      double dr = 0, di = 0, pr = 0, pi = 0;
#pragma unroll 8
      for (int kk = 0; kk < CKC; ++kk) {
        const double2 b = Lj[c][kk], a = Lj[r][kk], x = Li[r][kk];
        // a conj(b), x conj(b)
        dr += a.x * b.x + a.y * b.y;
        di += a.y * b.x - a.x * b.y;
        pr += x.x * b.x + x.y * b.y;
        pi += x.y * b.x - x.x * b.y;
      }
      if (c <= r) aD[q] = make_double2(aD[q].x - dr, aD[q].y - di);
      aP[q] = make_double2(aP[q].x - pr, aP[q].y - pi);
    }
    __syncthreads();
  }
#pragma unroll
  for (int q = 0; q < EPT; ++q) {
    const int r = tid / CB + RS * q, c = tid % CB;
    D[r][c] = c <= r ? aD[q] : make_double2(0, 0);
    P[r][c] = aP[q];
  }
  __syncthreads();
  // the diagonal block, factored by wave 0: lane i holds row i
  if (tid < 64) {
    const int i = tid;
    double2 row[CB];
#pragma unroll
    for (int k = 0; k < CB; ++k)
      row[k] = (i < jb && k <= i) ? D[i][k] : make_double2(0, 0);
    int fail = 0;
#pragma unroll
    for (int j = 0; j < CB; ++j) {
      const bool act = j < jb;
      double d = __shfl(row[j].x, j, 64);
      if (act && !(d > min_pivot)) {  // also catches NaN
        fail = 1;
        d = 1.0;
      }
      if (!act) d = 1.0;
      const double rt = sqrt(d), inv = 1.0 / rt;
      if (act && i == j) row[j] = make_double2(rt, 0.0);
      if (act && i > j) {
        row[j].x *= inv;
        row[j].y *= inv;
      }
#pragma unroll
      for (int k = j + 1; k < CB; ++k) {
        const double br = __shfl(row[j].x, k, 64), bi = __shfl(row[j].y, k, 64);
        if (act && k <= i && i < jb) {
          row[k].x -= row[j].x * br + row[j].y * bi;
          row[k].y -= row[j].y * br - row[j].x * bi;
        }
      }
    }
    if (i < CB) {
#pragma unroll
      for (int k = 0; k < CB; ++k) D[i][k] = (i < jb && k <= i) ? row[k] : make_double2(0, 0);
    }
    if (i < jb && fail) sh_fail = 1;
  }
  __syncthreads();
  if (diag) {
    // rows J..J+jb of L from column J on: the factored block, zero to its right
    const int w = n - J;
    for (int t = tid; t < jb * w; t += 256) {
      const int r = t / w, cc = t % w;
      L[(int64_t)(J + r) * n + J + cc] = cc < CB ? D[r][cc] : make_double2(0, 0);
    }
    if (tid == 0) {
      if (J == 0)
        *status = sh_fail;
      else if (sh_fail)
        *status = 1;
    }
    return;
  }
  // panel block: row r solves x L_JJ^H = p, lane k (of 32) owns x_k
  const int k = tid & (CB - 1), base = (tid & 63) & ~(CB - 1);
  const double dkk = D[k][k].x;
  for (int r = tid / CB; r < CB; r += RS) {
    double2 v = P[r][k];
#pragma unroll
    for (int l = 0; l < CB; ++l) {
      if (k == l && l < jb) {
        v.x /= dkk;
        v.y /= dkk;
      }
      const double xr = __shfl(v.x, base | l, 64), xi = __shfl(v.y, base | l, 64);
      if (k > l && l < jb) {
        const double2 dl = D[k][l];
        v.x -= xr * dl.x + xi * dl.y;
        v.y -= xi * dl.x - xr * dl.y;
      }
    }
    if (r < ib && k < jb) L[(int64_t)(I + r) * n + J + k] = v;
  }
}

// X = L^{-1} (lower), n <= TRB_MAXN: workgroup c solves L X[:, 16c..16c+15] = E by
// forward substitution over 16-row blocks.  For a row block I, the rows
// L[I..I+15, J..I) are staged through LDS in coalesced chunks of TRB_KC columns
// and thread (r, c) subtracts L[I+r, J..I) X[J..I, c] (the earlier row blocks of
// its column, kept in LDS); then the 16 x 16 diagonal solve runs in registers:
// the 16 lanes of one column sit in one wave and pass each finished x_k along
// with a shuffle.  A few barriers per row block instead of one per row.
constexpr int TRB = 16;
constexpr int TRB_MAXN = 512;
constexpr int TRB_KC = 64;
__global__ __launch_bounds__(256) void trtri_blocked_kernel(int n, const double2 *L,
                                                            double2 *X) {
  __shared__ double2 xs[TRB_MAXN * TRB];        // rows J.. of this column block
  __shared__ double2 lp[TRB][TRB_KC + 1];       // a chunk of the row block's L rows
  const int J = blockIdx.x * TRB;
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int r = lane & (TRB - 1);               // row within the row block
  const int c = (tid >> 6) * 4 + (lane >> 4);   // column within the column block
  const int col = J + c;
  // rows above the column block are zero in X
  for (int t = tid; t < J * TRB; t += 256) {
    const int row = t / TRB, cc = J + t % TRB;
    if (cc < n) X[(int64_t)row * n + cc] = make_double2(0, 0);
  }
  for (int I = J; I < n; I += TRB) {
    const int row = I + r;
    const bool live = row < n;
    double br = (live && row == col) ? 1.0 : 0.0, bi = 0.0;
    for (int k0 = J; k0 < I; k0 += TRB_KC) {
      const int kc = (I - k0) < TRB_KC ? (I - k0) : TRB_KC;
      for (int t = tid; t < TRB * TRB_KC; t += 256) {
        const int rr = t / TRB_KC, kk = t - rr * TRB_KC;
        lp[rr][kk] = (I + rr < n && kk < kc) ? L[(int64_t)(I + rr) * n + k0 + kk]
                                             : make_double2(0, 0);
      }
      __syncthreads();
      for (int kk = 0; kk < kc; ++kk) {
        const double2 l = lp[r][kk], xk = xs[(k0 + kk - J) * TRB + c];
        br -= l.x * xk.x - l.y * xk.y;
        bi -= l.x * xk.y + l.y * xk.x;
      }
      __syncthreads();
    }
    // diagonal block: L[I+r][I+k], k < r, and the pivot 1 / L[I+r][I+r]
    double2 lrow[TRB];
#pragma unroll
    for (int k = 0; k < TRB; ++k)
      lrow[k] = (live && k < r && I + k < n) ? L[(int64_t)row * n + I + k] : make_double2(0, 0);
    const double2 d = live ? L[(int64_t)row * n + row] : make_double2(1, 0);
    const double den = d.x * d.x + d.y * d.y;
    const double2 dinv = make_double2(d.x / den, -d.y / den);
    double xr = 0, xi = 0;
#pragma unroll
    for (int k = 0; k < TRB; ++k) {
      if (r == k) {
        xr = br * dinv.x - bi * dinv.y;
        xi = br * dinv.y + bi * dinv.x;
      }
      // x_k from lane k of this 16-lane group
      const double kr = __shfl(xr, (lane & ~(TRB - 1)) | k, 64);
      const double ki = __shfl(xi, (lane & ~(TRB - 1)) | k, 64);
      if (r > k) {
        br -= lrow[k].x * kr - lrow[k].y * ki;
        bi -= lrow[k].x * ki + lrow[k].y * kr;
      }
    }
    if (live) {
      xs[(row - J) * TRB + c] = make_double2(xr, xi);
      if (col < n) X[(int64_t)row * n + col] = make_double2(xr, xi);
    }
    __syncthreads();
  }
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
  constexpr int cb = 16;
  if (n <= cb) {
    hipLaunchKernelGGL(potrf_blocked_kernel, dim3(1), dim3(PT), 0, s, n,
                       static_cast<const double2 *>(G), ldg, static_cast<double2 *>(L),
                       shift_factor, min_pivot, static_cast<int *>(status));
    PTN_LAUNCH_CHECK();
    return PTN_OK;
  }
  for (int J = 0; J < n; J += cb) {
    hipLaunchKernelGGL(potrf_left_kernel<cb>, dim3((n - J + cb - 1) / cb), dim3(256), 0, s, n, J,
                       static_cast<const double2 *>(G), ldg, static_cast<double2 *>(L),
                       shift_factor, min_pivot, static_cast<int *>(status));
    PTN_LAUNCH_CHECK();
  }
  return PTN_OK;
}

extern "C" int ptn_trtri_lower(void *stream, int32_t n, const void *L, void *X) {
  if (n <= 0 || n > 1024 || !L || !X) {
    ptn::set_error("ptn_trtri_lower: bad arguments (n=%d, max 1024)", n);
    return n > 1024 ? PTN_ERANGE : PTN_EINVAL;
  }
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (n <= TRB_MAXN) {
    hipLaunchKernelGGL(trtri_blocked_kernel, dim3((n + TRB - 1) / TRB), dim3(256), 0, s, n,
                       static_cast<const double2 *>(L), static_cast<double2 *>(X));
    PTN_LAUNCH_CHECK();
    return PTN_OK;
  }
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

// ---- QR split of a small tensor in one workgroup ------------------------------
// The m x n panel (rows: the tensor's other legs, row-major in leg order;
// columns: leg `leg`) is gathered into LDS, factorised by Householder QR
// (reflectors as LAPACK's zlarfg / zgeqr2 build them: beta real, H^H applied
// to the trailing columns), R (its upper triangle) is written, then Q is formed
// in place (zung2r order: H_{k-1} first).  Row r of R and column r of Q are
// negated where beta_r < 0, so R has a real non-negative diagonal: for a
// full-rank panel the unique QR of that gauge.  KEEP shape: Q has the input's
// dims (column c >= k = min(m, n) zero), R is n x n (rows >= k zero).
namespace {
constexpr int QST = 256;
constexpr int QS_MAXE = 3072;
constexpr int QS_MAXN = 256;
constexpr int QS_MAXR = 8;

struct QrSmallArgs {
  int64_t dims[QS_MAXR];
  int64_t istr[QS_MAXR];  // input strides (elements)
  int64_t ostr[QS_MAXR];  // Q strides: contiguous row-major over dims
  int rank, leg;
  int m, n, k;
};

__device__ __forceinline__ double2 qcmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// conj(a) * b
__device__ __forceinline__ double2 qcjmul(double2 a, double2 b) {
  return make_double2(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);
}

__device__ double qs_block_sum(double v, double *red) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double s = 0;
  for (int w = 0; w < QST / 64; ++w) s += red[w];
  __syncthreads();
  return s;
}

// input and Q offsets of panel row i (the other legs' flattened index)
__device__ __forceinline__ void qs_row_offsets(const QrSmallArgs &a, int64_t i, int64_t &io,
                                               int64_t &oo) {
  io = 0;
  oo = 0;
  for (int d = a.rank - 1; d >= 0; --d) {
    if (d == a.leg) continue;
    const int64_t idx = i % a.dims[d];
    i /= a.dims[d];
    io += idx * a.istr[d];
    oo += idx * a.ostr[d];
  }
}

__global__ __launch_bounds__(QST) void qr_split_small_kernel(QrSmallArgs a, const double2 *T,
                                                             double2 *Q, double2 *R) {
  __shared__ double2 P[QS_MAXE];
  __shared__ double2 sw[QS_MAXN], stau[QS_MAXN];
  __shared__ double ssgn[QS_MAXN];
  __shared__ double red[QST / 64];
  __shared__ double2 sscale;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int m = a.m, n = a.n, k = a.k;
  for (int e = tid; e < m * n; e += QST) {
    const int i = e % m, j = e / m;
    int64_t io, oo;
    qs_row_offsets(a, i, io, oo);
    P[e] = T[io + (int64_t)j * a.istr[a.leg]];
  }
  __syncthreads();
  for (int j = 0; j < k; ++j) {
    double loc = 0;
    for (int i = j + 1 + tid; i < m; i += QST) loc += P[i + j * m].x * P[i + j * m].x +
                                                      P[i + j * m].y * P[i + j * m].y;
    const double xn2 = qs_block_sum(loc, red);
    if (tid == 0) {
      const double2 al = P[j + j * m];
      if (xn2 == 0.0 && al.y == 0.0) {
        stau[j] = make_double2(0, 0);
        sscale = make_double2(0, 0);
      } else {
        const double nr = sqrt(al.x * al.x + al.y * al.y + xn2);
        const double beta = al.x >= 0 ? -nr : nr;
        stau[j] = make_double2((beta - al.x) / beta, -al.y / beta);
        // 1 / (alpha - beta)
        const double dr = al.x - beta, di = al.y, dd = dr * dr + di * di;
        sscale = make_double2(dr / dd, -di / dd);
        P[j + j * m] = make_double2(beta, 0.0);
      }
    }
    __syncthreads();
    const double2 tau = stau[j];
    if (tau.x == 0.0 && tau.y == 0.0) continue;  // H_j = I (uniform)
    const double2 sc = sscale;
    for (int i = j + 1 + tid; i < m; i += QST) P[i + j * m] = qcmul(sc, P[i + j * m]);
    __syncthreads();
    // w_c = v^H P[j:m, c], v_j = 1
    for (int c = j + 1 + wave; c < n; c += QST / 64) {
      double2 acc = make_double2(0, 0);
      for (int i = j + lane; i < m; i += 64) {
        const double2 v = i == j ? make_double2(1, 0) : P[i + j * m];
        const double2 p = qcjmul(v, P[i + c * m]);
        acc.x += p.x;
        acc.y += p.y;
      }
      for (int off = 32; off > 0; off >>= 1) {
        acc.x += __shfl_down(acc.x, off, 64);
        acc.y += __shfl_down(acc.y, off, 64);
      }
      if (lane == 0) sw[c] = acc;
    }
    __syncthreads();
    // P[j:m, c] -= conj(tau) v w_c
    const double2 ctau = make_double2(tau.x, -tau.y);
    const int rows = m - j, cols = n - j - 1;
    for (int e = tid; e < rows * cols; e += QST) {
      const int i = j + e % rows, c = j + 1 + e / rows;
      const double2 v = i == j ? make_double2(1, 0) : P[i + j * m];
      const double2 t = qcmul(ctau, qcmul(v, sw[c]));
      P[i + c * m].x -= t.x;
      P[i + c * m].y -= t.y;
    }
    __syncthreads();
  }
  // gauge: real non-negative diagonal of R
  for (int r = tid; r < k; r += QST) ssgn[r] = P[r + r * m].x < 0.0 ? -1.0 : 1.0;
  __syncthreads();
  for (int e = tid; e < n * n; e += QST) {
    const int r = e / n, c = e % n;
    double2 v = make_double2(0, 0);
    if (r < k && c >= r) {
      v = P[r + c * m];
      v.x *= ssgn[r];
      v.y *= ssgn[r];
    }
    R[e] = v;
  }
  __syncthreads();
  // Q = H_0 ... H_{k-1} [I; 0], formed in place from the last reflector back
  for (int j = k - 1; j >= 0; --j) {
    const double2 tau = stau[j];
    if (j < k - 1) {
      for (int c = j + 1 + wave; c < k; c += QST / 64) {
        double2 acc = make_double2(0, 0);
        for (int i = j + lane; i < m; i += 64) {
          const double2 v = i == j ? make_double2(1, 0) : P[i + j * m];
          const double2 p = qcjmul(v, P[i + c * m]);
          acc.x += p.x;
          acc.y += p.y;
        }
        for (int off = 32; off > 0; off >>= 1) {
          acc.x += __shfl_down(acc.x, off, 64);
          acc.y += __shfl_down(acc.y, off, 64);
        }
        if (lane == 0) sw[c] = acc;
      }
      __syncthreads();
      const int rows = m - j, cols = k - j - 1;
      for (int e = tid; e < rows * cols; e += QST) {
        const int i = j + e % rows, c = j + 1 + e / rows;
        const double2 v = i == j ? make_double2(1, 0) : P[i + j * m];
        const double2 t = qcmul(tau, qcmul(v, sw[c]));
        P[i + c * m].x -= t.x;
        P[i + c * m].y -= t.y;
      }
      __syncthreads();
    }
    for (int i = tid; i < m; i += QST) {
      double2 v;
      if (i < j)
        v = make_double2(0, 0);
      else if (i == j)
        v = make_double2(1.0 - tau.x, -tau.y);
      else {
        const double2 t = qcmul(tau, P[i + j * m]);
        v = make_double2(-t.x, -t.y);
      }
      P[i + j * m] = v;
    }
    __syncthreads();
  }
  for (int e = tid; e < m * n; e += QST) {
    const int i = e % m, c = e / m;
    int64_t io, oo;
    qs_row_offsets(a, i, io, oo);
    double2 v = make_double2(0, 0);
    if (c < k) {
      v = P[i + c * m];
      v.x *= ssgn[c];
      v.y *= ssgn[c];
    }
    Q[oo + (int64_t)c * a.ostr[a.leg]] = v;
  }
}
}  // namespace

extern "C" int ptn_qr_split_small(void *stream, int32_t rank, const int64_t *dims,
                                  const int64_t *strides, int32_t leg, const void *T, void *Q,
                                  void *R) {
  if (rank < 1 || rank > QS_MAXR || leg < 0 || leg >= rank || !dims || !strides || !T || !Q ||
      !R) {
    ptn::set_error("ptn_qr_split_small: bad arguments (rank=%d leg=%d)", rank, leg);
    return PTN_EINVAL;
  }
  QrSmallArgs a;
  int64_t total = 1;
  for (int d = 0; d < rank; ++d) {
    if (dims[d] < 1) {
      ptn::set_error("ptn_qr_split_small: dimension %d is %lld", d, (long long)dims[d]);
      return PTN_EINVAL;
    }
    a.dims[d] = dims[d];
    a.istr[d] = strides[d];
    total *= dims[d];
  }
  int64_t acc = 1;
  for (int d = rank - 1; d >= 0; --d) {
    a.ostr[d] = acc;
    acc *= dims[d];
  }
  const int64_t n = dims[leg], m = total / n;
  if (total > QS_MAXE || n > QS_MAXN) {
    ptn::set_error("ptn_qr_split_small: panel %lld x %lld beyond %d elements / %d columns",
                   (long long)m, (long long)n, QS_MAXE, QS_MAXN);
    return PTN_ERANGE;
  }
  a.rank = rank;
  a.leg = leg;
  a.m = (int)m;
  a.n = (int)n;
  a.k = (int)(m < n ? m : n);
  hipLaunchKernelGGL(qr_split_small_kernel, dim3(1), dim3(QST), 0,
                     static_cast<hipStream_t>(stream), a, static_cast<const double2 *>(T),
                     static_cast<double2 *>(Q), static_cast<double2 *>(R));
  PTN_LAUNCH_CHECK();
  return PTN_OK;
}
