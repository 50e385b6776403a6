// One-sided Jacobi SVD of small complex matrices, one workgroup per matrix.
//
// Hestenes' method: the columns of G (m_g x n_g, column-major, m_g >= n_g) are
// rotated pairwise by 2 x 2 unitary Jacobi rotations until every pair is
// orthogonal to working precision; then sigma_j = |g_j|, U = G diag(1/sigma)
// and V is the product of the rotations.  The rotation of a pair (a, b) with
// alpha = |a|^2, beta = |b|^2, gamma = a^H b = |gamma| e^{i phi} diagonalises
// the pair's Gram matrix:
//   zeta = (beta - alpha) / (2 |gamma|),  t = sign(zeta) / (|zeta| + sqrt(1 + zeta^2)),
//   c = 1 / sqrt(1 + t^2),  s = c t,
//   a' = c a - s e^{-i phi} b,   b' = s e^{i phi} a + c b   (same on the columns of V).
// A pair is rotated when |gamma| > sqrt(m_g) eps sqrt(alpha beta); a sweep that
// rotates nothing ends the iteration.  The method computes small singular values
// to high relative accuracy, so the truncation rules see them as LAPACK's
// bidiagonal SVD does (to O(eps |A|)).  Applied to R^H of a QR factor (on_rows
// of a triangle) it converges in ~10 sweeps also for graded spectra.
//
// Parallel order: the round-robin tournament (circle method) on the n_g columns
// padded to even N: N - 1 rounds of N/2 disjoint pairs per sweep, every pair once.
// A group of TPP lanes (one wave or a power-of-two part of it) owns one pair:
// the lanes split the rows, reduce alpha, beta, gamma by shuffles within the
// group, and update both columns of G and V.  Pairs of a round touch disjoint
// columns, so one workgroup barrier per round orders the rounds.
//
// Replaces np.linalg.svd / scipy gesvd in tensor_svd (util/tensor_splitting.py:122-177)
// for the panels of the truncated two-site splits.
#include "ptn_common.h"

#include <math.h>
#include <string.h>

#include <vector>

namespace {

constexpr int SVD_NT = 1024;        // threads per workgroup
constexpr int SVD_MAXN = 256;       // largest n_g (columns rotated)
constexpr int SVD_MAXM = 4096;      // largest m_g (rows per column)
constexpr int SVD_MAX_SWEEPS = 64;

struct SvdJob {
  int64_t m, n;          // A is m x n, column-major with leading dimension lda
  const double2 *A;
  int64_t lda;
  int32_t on_rows;       // 0: rotate the columns of A (m >= n); 1: of A^H (m <= n)
  int32_t _pad;
  double *S;             // k = min(m, n) singular values, descending
  double2 *U;            // m x k, column-major (ld m)
  double2 *Vh;           // k x n, column-major (ld k)
  double2 *ws;           // m n + k k complex128 workspace
  int32_t *info;         // 0: converged; 1: not within SVD_MAX_SWEEPS sweeps
  int32_t _pad2[2];
};

__device__ __forceinline__ double2 cmulc(double2 a, double2 b) {  // conj(a) b
  return make_double2(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);
}

// pair i of round r in the circle method on N players (N even): player 0 stays,
// players 1..N-1 rotate
__device__ __forceinline__ void rr_pair(int N, int r, int i, int &p, int &q) {
  const int L = N - 1;
  if (i == 0) {
    p = 0;
    q = 1 + (r % L);
  } else {
    p = 1 + ((r + i) % L);
    q = 1 + ((r - i + L) % L);
  }
  if (p > q) {
    const int t = p;
    p = q;
    q = t;
  }
}

template <int TPP>
__device__ __forceinline__ double group_sum(double v) {
#pragma unroll
  for (int off = TPP / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// one rotation sweep step of the pair (a, b) of G (rows mg) and (va, vb) of V
// (rows ng) by the TPP lanes of a group; returns true when the pair was rotated.
// Rows are handled UNR at a time with every load issued before its use, so a
// lane waits for one memory latency per UNR rows, not per row.
template <int TPP, int UNR>
__device__ __forceinline__ bool rotate_pair(double2 *a, double2 *b, int64_t mg, double2 *va,
                                            double2 *vb, int ng, int gl, double tol) {
  double al = 0, be = 0, gr = 0, gi = 0;
  for (int64_t i0 = gl; i0 < mg; i0 += (int64_t)TPP * UNR) {
    double2 x[UNR], y[UNR];
#pragma unroll
    for (int u = 0; u < UNR; ++u) {
      const int64_t i = i0 + (int64_t)u * TPP;
      x[u] = i < mg ? a[i] : make_double2(0, 0);
      y[u] = i < mg ? b[i] : make_double2(0, 0);
    }
#pragma unroll
    for (int u = 0; u < UNR; ++u) {
      al += x[u].x * x[u].x + x[u].y * x[u].y;
      be += y[u].x * y[u].x + y[u].y * y[u].y;
      const double2 g = cmulc(x[u], y[u]);
      gr += g.x;
      gi += g.y;
    }
  }
  al = group_sum<TPP>(al);
  be = group_sum<TPP>(be);
  gr = group_sum<TPP>(gr);
  gi = group_sum<TPP>(gi);
  const double ag = hypot(gr, gi);
  if (!(ag > tol * sqrt(al) * sqrt(be) && ag > 0.0)) return false;
  const double zeta = (be - al) / (2.0 * ag);
  const double t = (zeta >= 0.0 ? 1.0 : -1.0) / (fabs(zeta) + hypot(1.0, zeta));
  const double c = 1.0 / sqrt(1.0 + t * t), s = c * t;
  const double er = gr / ag, ei = gi / ag;  // e^{i phi}
  // a' = c a - s conj(e) b,  b' = s e a + c b
  auto rot = [&](double2 *pa, double2 *pb, int64_t len) {
    for (int64_t i0 = gl; i0 < len; i0 += (int64_t)TPP * UNR) {
      double2 x[UNR], y[UNR];
#pragma unroll
      for (int u = 0; u < UNR; ++u) {
        const int64_t i = i0 + (int64_t)u * TPP;
        x[u] = i < len ? pa[i] : make_double2(0, 0);
        y[u] = i < len ? pb[i] : make_double2(0, 0);
      }
#pragma unroll
      for (int u = 0; u < UNR; ++u) {
        const int64_t i = i0 + (int64_t)u * TPP;
        if (i < len) {
          pa[i] = make_double2(c * x[u].x - s * (er * y[u].x + ei * y[u].y),
                               c * x[u].y - s * (er * y[u].y - ei * y[u].x));
          pb[i] = make_double2(s * (er * x[u].x - ei * x[u].y) + c * y[u].x,
                               s * (er * x[u].y + ei * x[u].x) + c * y[u].y);
        }
      }
    }
  };
  rot(a, b, mg);
  rot(va, vb, ng);
  return true;
}

// LDS_ELEMS complex128 of LDS hold G and V when mg ng + ng ng fits (128 KB)
constexpr int64_t LDS_ELEMS = 8192;

// singular values, their descending order, U and Vh from the converged G and V
// (one workgroup of SVD_NT threads; groups of TPP lanes reduce the column norms)
template <int TPP>
__device__ void svd_finish(const SvdJob &job, const double2 *G, const double2 *V, int64_t mg,
                           int ng, bool converged) {
  __shared__ double sig[SVD_MAXN];
  __shared__ int rank_of[SVD_MAXN];
  const bool rows = job.on_rows != 0;
  const int tid = threadIdx.x;
  const int grp = tid / TPP, gl = tid % TPP;
  // singular values: group g reduces the norms of columns 2g and 2g + 1
  for (int cbase = 2 * grp; cbase < ng; cbase += 2 * (SVD_NT / TPP)) {
    for (int h = 0; h < 2; ++h) {
      const int j = cbase + h;
      if (j >= ng) break;
      const double2 *g = G + (int64_t)j * mg;
      double acc = 0;
      for (int64_t i = gl; i < mg; i += TPP) acc += g[i].x * g[i].x + g[i].y * g[i].y;
      acc = group_sum<TPP>(acc);
      if (gl == 0) sig[j] = sqrt(acc);
    }
  }
  __syncthreads();
  // descending order, ties in column order
  for (int j = tid; j < ng; j += SVD_NT) {
    const double sj = sig[j];
    int rk = 0;
    for (int k = 0; k < ng; ++k) rk += (sig[k] > sj || (sig[k] == sj && k < j)) ? 1 : 0;
    rank_of[j] = rk;
    job.S[rk] = sj;
  }
  __syncthreads();
  // A = U S Vh.  Columns: G = Ug S Vg^H, U = Ug, Vh = Vg^H.  Rows: A^H = G, so
  // A = Vg S Ug^H: U = Vg, Vh = Ug^H.  k = n_g in both cases.
  const int k = ng;
  for (int64_t e = tid; e < mg * ng; e += SVD_NT) {
    const int64_t i = e % mg;
    const int j = (int)(e / mg);
    const double inv = sig[j] > 0.0 ? 1.0 / sig[j] : 0.0;
    const double2 u = make_double2(G[e].x * inv, G[e].y * inv);
    const int rk = rank_of[j];
    if (!rows) job.U[i + (int64_t)rk * job.m] = u;            // U (m x k)
    else job.Vh[rk + i * (int64_t)k] = make_double2(u.x, -u.y);  // Vh (k x n), n = mg
  }
  for (int e = tid; e < ng * ng; e += SVD_NT) {
    const int i = e % ng, j = e / ng;
    const double2 v = V[e];
    const int rk = rank_of[j];
    if (!rows) job.Vh[rk + (int64_t)i * k] = make_double2(v.x, -v.y);  // Vh (k x n), n = ng
    else job.U[i + (int64_t)rk * job.m] = v;                            // U (m x k), m = ng
  }
  if (tid == 0) *job.info = converged ? 0 : 1;
}

template <int TPP, bool IN_LDS>
__device__ void svd_jacobi_body(const SvdJob &job, double2 *G, double2 *V) {
  __shared__ int rotated;
  constexpr int UNR = IN_LDS ? 2 : 4;
  const bool rows = job.on_rows != 0;
  const int64_t mg = rows ? job.n : job.m;  // rows of G
  const int ng = (int)(rows ? job.m : job.n);  // columns of G
  const int tid = threadIdx.x;
  // G = op(A), V = I
  for (int64_t e = tid; e < mg * ng; e += SVD_NT) {
    const int64_t i = e % mg, j = e / mg;
    if (!rows) {
      G[e] = job.A[i + j * job.lda];
    } else {
      const double2 a = job.A[j + i * job.lda];  // G(i, j) = conj(A(j, i))
      G[e] = make_double2(a.x, -a.y);
    }
  }
  for (int e = tid; e < ng * ng; e += SVD_NT)
    V[e] = make_double2((e % ng) == (e / ng) ? 1.0 : 0.0, 0.0);
  __syncthreads();

  const int N = ng + (ng & 1);
  const int npairs = N / 2;
  const int grp = tid / TPP, gl = tid % TPP;
  const double tol = sqrt((double)mg) * 2.220446049250313e-16;
  int sweep = 0;
  bool converged = ng < 2;
  for (; sweep < SVD_MAX_SWEEPS && !converged; ++sweep) {
    if (tid == 0) rotated = 0;
    __syncthreads();
    for (int r = 0; r < N - 1; ++r) {
      if (grp < npairs) {
        int p, q;
        rr_pair(N, r, grp, p, q);
        if (q < ng &&
            rotate_pair<TPP, UNR>(G + (int64_t)p * mg, G + (int64_t)q * mg, mg,
                                  V + (int64_t)p * ng, V + (int64_t)q * ng, ng, gl, tol) &&
            gl == 0)
          rotated = 1;
      }
      __syncthreads();
    }
    converged = rotated == 0;
    __syncthreads();  // every thread has read `rotated` before the next reset
  }
  svd_finish<TPP>(job, G, V, mg, ng, converged);
}

template <int TPP>
__device__ __forceinline__ void svd_jacobi_dispatch(const SvdJob &job) {
  __shared__ double2 lds[LDS_ELEMS];
  const int64_t mg = job.on_rows ? job.n : job.m, ng = job.on_rows ? job.m : job.n;
  if (mg * ng + ng * ng <= LDS_ELEMS) {
    svd_jacobi_body<TPP, true>(job, lds, lds + mg * ng);
  } else {
    svd_jacobi_body<TPP, false>(job, job.ws, job.ws + mg * ng);
  }
}

template <int TPP>
__global__ __launch_bounds__(SVD_NT) void svd_jacobi_one(SvdJob job) {
  svd_jacobi_dispatch<TPP>(job);
}

template <int TPP>
__global__ __launch_bounds__(SVD_NT) void svd_jacobi_many(const SvdJob *__restrict__ jobs) {
  const SvdJob job = jobs[blockIdx.x];
  svd_jacobi_dispatch<TPP>(job);
}

// ---- wide form: one matrix over many workgroups ------------------------------------
// For a single matrix with many columns the one-workgroup kernel is bound by the L2
// bandwidth of one CU (a 256-column triangle: ~17 ms per sweep).  The wide form runs
// the same rotations in the same round-robin order with one launch per round: one
// wave per pair, npairs workgroups, so a round takes a few microseconds on the
// whole chip.  G, V and one "rotated" flag per sweep live in the workspace; the
// rounds of sweep s return at once when sweep s - 1 rotated nothing.
constexpr int WIDE_MIN_COLS = 65;   // columns from which a single matrix runs wide
constexpr int WIDE_FIRST_SWEEPS = 8;
constexpr int WIDE_MORE_SWEEPS = 2;

__device__ __forceinline__ int *wide_flags(const SvdJob &job, int64_t mg, int64_t ng) {
  return reinterpret_cast<int *>(job.ws + mg * ng + ng * ng);
}

__global__ __launch_bounds__(256) void svd_wide_init(SvdJob job) {
  const bool rows = job.on_rows != 0;
  const int64_t mg = rows ? job.n : job.m, ng = rows ? job.m : job.n;
  double2 *G = job.ws, *V = job.ws + mg * ng;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  const int64_t t0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  for (int64_t e = t0; e < mg * ng; e += stride) {
    const int64_t i = e % mg, j = e / mg;
    if (!rows) {
      G[e] = job.A[i + j * job.lda];
    } else {
      const double2 a = job.A[j + i * job.lda];  // G(i, j) = conj(A(j, i))
      G[e] = make_double2(a.x, -a.y);
    }
  }
  for (int64_t e = t0; e < ng * ng; e += stride)
    V[e] = make_double2((e % ng) == (e / ng) ? 1.0 : 0.0, 0.0);
  if (t0 < SVD_MAX_SWEEPS) wide_flags(job, mg, ng)[t0] = 0;
}

__global__ __launch_bounds__(64) void svd_wide_round(SvdJob job, int r, int sweep) {
  const bool rows = job.on_rows != 0;
  const int64_t mg = rows ? job.n : job.m;
  const int ng = (int)(rows ? job.m : job.n);
  int *flags = wide_flags(job, mg, ng);
  if (sweep > 0 && flags[sweep - 1] == 0) return;   // converged before this sweep
  const int N = ng + (ng & 1);
  int p, q;
  rr_pair(N, r, (int)blockIdx.x, p, q);
  if (q >= ng) return;
  double2 *G = job.ws, *V = job.ws + mg * ng;
  const double tol = sqrt((double)mg) * 2.220446049250313e-16;
  const bool rot = rotate_pair<64, 4>(G + (int64_t)p * mg, G + (int64_t)q * mg, mg,
                                      V + (int64_t)p * ng, V + (int64_t)q * ng, ng,
                                      (int)threadIdx.x, tol);
  if (rot && threadIdx.x == 0) flags[sweep] = 1;
}

__global__ __launch_bounds__(SVD_NT) void svd_wide_finish(SvdJob job, int converged) {
  const bool rows = job.on_rows != 0;
  const int64_t mg = rows ? job.n : job.m;
  const int ng = (int)(rows ? job.m : job.n);
  svd_finish<8>(job, job.ws, job.ws + mg * ng, mg, ng, converged != 0);
}

// The wide solve: init, sweeps in chunks (the flags of a chunk read back once, the
// host waits on the stream), finish.
int svd_wide(hipStream_t s, const SvdJob &j) {
  const int64_t mg = j.on_rows ? j.n : j.m, ng = j.on_rows ? j.m : j.n;
  const int N = (int)(ng + (ng & 1));
  const int npairs = N / 2;
  hipLaunchKernelGGL(svd_wide_init, dim3(256), dim3(256), 0, s, j);
  PTN_LAUNCH_CHECK();
  int done = 0, converged = 0;
  int flags[SVD_MAX_SWEEPS];
  while (done < SVD_MAX_SWEEPS && !converged) {
    const int chunk = done == 0 ? WIDE_FIRST_SWEEPS : WIDE_MORE_SWEEPS;
    const int upto = done + chunk < SVD_MAX_SWEEPS ? done + chunk : SVD_MAX_SWEEPS;
    for (int sw = done; sw < upto; ++sw)
      for (int r = 0; r < N - 1; ++r)
        hipLaunchKernelGGL(svd_wide_round, dim3(npairs), dim3(64), 0, s, j, r, sw);
    PTN_LAUNCH_CHECK();
    const int *dflags = reinterpret_cast<const int *>(j.ws + mg * ng + ng * ng);
    if (hipMemcpyAsync(flags, dflags, sizeof(int) * upto, hipMemcpyDeviceToHost, s) !=
            hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess) {
      ptn::set_error("ptn_svd_jacobi: reading the sweep flags failed");
      return PTN_EHIP;
    }
    for (int sw = done; sw < upto; ++sw)
      if (flags[sw] == 0) converged = 1;
    done = upto;
  }
  hipLaunchKernelGGL(svd_wide_finish, dim3(1), dim3(SVD_NT), 0, s, j, converged);
  PTN_LAUNCH_CHECK();
  return PTN_OK;
}

int tpp_for(int ng) {
  const int npairs = (ng + 1) / 2;
  if (npairs <= 16) return 64;
  if (npairs <= 32) return 32;
  if (npairs <= 64) return 16;
  return 8;
}

int check_job(const SvdJob &j) {
  const int64_t mg = j.on_rows ? j.n : j.m, ng = j.on_rows ? j.m : j.n;
  if (j.m <= 0 || j.n <= 0 || !j.A || !j.S || !j.U || !j.Vh || !j.ws || !j.info ||
      j.lda < j.m) {
    ptn::set_error("ptn_svd_jacobi: bad arguments (m=%lld n=%lld lda=%lld)", (long long)j.m,
                   (long long)j.n, (long long)j.lda);
    return PTN_EINVAL;
  }
  if (mg < ng) {
    ptn::set_error("ptn_svd_jacobi: on_rows=%d needs %s (m=%lld n=%lld)", j.on_rows,
                   j.on_rows ? "m <= n" : "m >= n", (long long)j.m, (long long)j.n);
    return PTN_EINVAL;
  }
  if (ng > SVD_MAXN || mg > SVD_MAXM) {
    ptn::set_error("ptn_svd_jacobi: %lld x %lld beyond the one-workgroup limits "
                   "(short side <= %d, long side <= %d)", (long long)j.m, (long long)j.n,
                   SVD_MAXN, SVD_MAXM);
    return PTN_ERANGE;
  }
  return PTN_OK;
}

}  // namespace

extern "C" int64_t ptn_svd_jacobi_work_bytes(int64_t m, int64_t n) {
  const int64_t k = m < n ? m : n;
  // G and V, then the wide form's per-sweep flags (256 bytes)
  return (int64_t)sizeof(double2) * (m * n + k * k) + 256;
}

extern "C" int ptn_svd_jacobi(void *stream, int64_t m, int64_t n, const void *A, int64_t lda,
                              int32_t on_rows, void *S, void *U, void *Vh, void *work,
                              void *info) {
  SvdJob j;
  memset(&j, 0, sizeof(j));
  j.m = m;
  j.n = n;
  j.A = static_cast<const double2 *>(A);
  j.lda = lda;
  j.on_rows = on_rows;
  j.S = static_cast<double *>(S);
  j.U = static_cast<double2 *>(U);
  j.Vh = static_cast<double2 *>(Vh);
  j.ws = static_cast<double2 *>(work);
  j.info = static_cast<int32_t *>(info);
  const int rc = check_job(j);
  if (rc) return rc;
  hipStream_t s = static_cast<hipStream_t>(stream);
  if ((on_rows ? m : n) >= WIDE_MIN_COLS) return svd_wide(s, j);
  switch (tpp_for((int)(on_rows ? m : n))) {
    case 64: hipLaunchKernelGGL(svd_jacobi_one<64>, dim3(1), dim3(SVD_NT), 0, s, j); break;
    case 32: hipLaunchKernelGGL(svd_jacobi_one<32>, dim3(1), dim3(SVD_NT), 0, s, j); break;
    case 16: hipLaunchKernelGGL(svd_jacobi_one<16>, dim3(1), dim3(SVD_NT), 0, s, j); break;
    default: hipLaunchKernelGGL(svd_jacobi_one<8>, dim3(1), dim3(SVD_NT), 0, s, j); break;
  }
  PTN_LAUNCH_CHECK();
  return PTN_OK;
}

extern "C" int64_t ptn_svd_jacobi_batched_work_bytes(int32_t count) {
  return (int64_t)sizeof(SvdJob) * (count > 0 ? count : 0);
}

extern "C" int ptn_svd_jacobi_batched(void *stream, int32_t count, const int64_t *m,
                                      const int64_t *n, const void *const *A, const int64_t *lda,
                                      const int32_t *on_rows, void *const *S, void *const *U,
                                      void *const *Vh, void *const *ws, void *const *info,
                                      void *work) {
  if (count <= 0 || !work) {
    ptn::set_error("ptn_svd_jacobi_batched: bad arguments (count=%d)", count);
    return PTN_EINVAL;
  }
  std::vector<SvdJob> jobs((size_t)count);
  int tpp = 0;
  for (int b = 0; b < count; ++b) {
    SvdJob &j = jobs[(size_t)b];
    memset(&j, 0, sizeof(j));
    j.m = m[b];
    j.n = n[b];
    j.A = static_cast<const double2 *>(A[b]);
    j.lda = lda[b];
    j.on_rows = on_rows[b];
    j.S = static_cast<double *>(S[b]);
    j.U = static_cast<double2 *>(U[b]);
    j.Vh = static_cast<double2 *>(Vh[b]);
    j.ws = static_cast<double2 *>(ws[b]);
    j.info = static_cast<int32_t *>(info[b]);
    const int rc = check_job(j);
    if (rc) return rc;
    // one group size per launch: the smallest over the jobs (fewer lanes per pair
    // never drops a pair: 1024 / tpp groups cover the largest pair count)
    const int t = tpp_for((int)(j.on_rows ? j.m : j.n));
    tpp = (tpp == 0 || t < tpp) ? t : tpp;
  }
  hipStream_t s = static_cast<hipStream_t>(stream);
  const int rc = ptn::stage_to_device(s, jobs.data(), sizeof(SvdJob) * (size_t)count, work);
  if (rc) return rc;
  const SvdJob *dj = static_cast<const SvdJob *>(work);
  switch (tpp) {
    case 64: hipLaunchKernelGGL(svd_jacobi_many<64>, dim3(count), dim3(SVD_NT), 0, s, dj); break;
    case 32: hipLaunchKernelGGL(svd_jacobi_many<32>, dim3(count), dim3(SVD_NT), 0, s, dj); break;
    case 16: hipLaunchKernelGGL(svd_jacobi_many<16>, dim3(count), dim3(SVD_NT), 0, s, dj); break;
    default: hipLaunchKernelGGL(svd_jacobi_many<8>, dim3(count), dim3(SVD_NT), 0, s, dj); break;
  }
  PTN_LAUNCH_CHECK();
  return PTN_OK;
}
