// Level-1 kernels: Krylov re-orthogonalisation, norms, dots, linear
// combinations, strided permutation copies and fills.
//
// Reductions are deterministic: a fixed grid (a function of n only) writes
// per-block partials, a single block reduces them in a fixed order.  Vectors
// are streamed with 16-byte complex128 loads; each pass over the Krylov basis
// reads every q_i once per pass (groups of up to KV vectors per sweep over w).
#include "ptn_common.h"

#include <stdarg.h>
#include <string.h>

namespace ptn {
static thread_local char g_err[512];
void set_error(const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}
}  // namespace ptn

extern "C" const char *ptn_last_error(void) { return ptn::g_err; }
extern "C" int ptn_version(void) { return 100; }

namespace {

constexpr int RT = 256;       // threads per reduction block
constexpr int KV = 8;         // vectors per sweep over w
constexpr int MAXV = 64;      // max basis vectors per call
constexpr int MAXBLK = 1024;  // max partial blocks

inline int nblocks_for(int64_t n) {
  int64_t b = (n + RT * 4 - 1) / (RT * 4);
  if (b < 1) b = 1;
  if (b > MAXBLK) b = MAXBLK;
  return (int)b;
}

__device__ __forceinline__ double2 cmul_conj_a(double2 a, double2 b) {
  // conj(a) * b
  return make_double2(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);
}

// block-wide sum of `cnt` complex values per thread held in LDS rows
template <int CNT>
__device__ void block_reduce_store(double2 (&acc)[CNT], int cnt, double2 *out_row) {
  __shared__ double2 red[RT / 64][CNT];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int v = 0; v < cnt; ++v) {
    double x = acc[v].x, y = acc[v].y;
    for (int off = 32; off > 0; off >>= 1) {
      x += __shfl_down(x, off, 64);
      y += __shfl_down(y, off, 64);
    }
    if (lane == 0) red[wave][v] = make_double2(x, y);
  }
  __syncthreads();
  if (threadIdx.x < cnt) {
    double2 s = make_double2(0, 0);
    for (int w = 0; w < RT / 64; ++w) {
      s.x += red[w][threadIdx.x].x;
      s.y += red[w][threadIdx.x].y;
    }
    out_row[threadIdx.x] = s;
  }
  __syncthreads();
}

// partial[b][v] = sum over block b's range of conj(Q[v0+v]) * w, v < cnt;
// if self_slot >= 0 the slot `self_slot` gets sum |w|^2 (imag 0).
__global__ __launch_bounds__(RT) void dots_partial(int64_t n, const double2 *Q, int64_t ldq,
                                                   int v0, int cnt, const double2 *w,
                                                   double2 *partial, int pstride,
                                                   int self_slot) {
  double2 acc[KV + 1];
#pragma unroll
  for (int v = 0; v <= KV; ++v) acc[v] = make_double2(0, 0);
  const int64_t stride = (int64_t)gridDim.x * RT;
  for (int64_t i = (int64_t)blockIdx.x * RT + threadIdx.x; i < n; i += stride) {
    const double2 wi = w[i];
#pragma unroll
    for (int v = 0; v < KV; ++v) {
      if (v < cnt) {
        const double2 q = Q[(int64_t)(v0 + v) * ldq + i];
        const double2 t = cmul_conj_a(q, wi);
        acc[v].x += t.x;
        acc[v].y += t.y;
      }
    }
    if (self_slot >= 0) acc[KV].x += wi.x * wi.x + wi.y * wi.y;
  }
  double2 *row = partial + (int64_t)blockIdx.x * pstride;
  // move the self term next to the dots so one reduction covers both
  if (self_slot >= 0) acc[cnt] = acc[KV];
  block_reduce_store<KV + 1>(acc, cnt + (self_slot >= 0 ? 1 : 0), row + v0);
}

// out[v] = sum_b partial[b][v] for v < count (fixed order), optional add-into
__global__ __launch_bounds__(RT) void reduce_partials(const double2 *partial, int pstride,
                                                      int nblk, int count, double2 *out,
                                                      int accumulate) {
  for (int v = threadIdx.x; v < count; v += RT) {
    double2 s = make_double2(0, 0);
    for (int b = 0; b < nblk; ++b) {
      const double2 t = partial[(int64_t)b * pstride + v];
      s.x += t.x;
      s.y += t.y;
    }
    if (accumulate) {
      out[v].x += s.x;
      out[v].y += s.y;
    } else {
      out[v] = s;
    }
  }
}

// w -= sum_v h[v] Q[v];  optionally partial |w|^2 into partial[b][0]
__global__ __launch_bounds__(RT) void subtract_proj(int64_t n, const double2 *Q, int64_t ldq,
                                                    int nvec, const double2 *h, double2 *w,
                                                    double2 *partial, int pstride) {
  __shared__ double2 hs[MAXV];
  for (int v = threadIdx.x; v < nvec; v += RT) hs[v] = h[v];
  __syncthreads();
  double nrm = 0;
  const int64_t stride = (int64_t)gridDim.x * RT;
  for (int64_t i = (int64_t)blockIdx.x * RT + threadIdx.x; i < n; i += stride) {
    double2 wi = w[i];
    for (int v = 0; v < nvec; ++v) {
      const double2 q = Q[(int64_t)v * ldq + i];
      const double2 hv = hs[v];
      wi.x -= hv.x * q.x - hv.y * q.y;
      wi.y -= hv.x * q.y + hv.y * q.x;
    }
    w[i] = wi;
    nrm += wi.x * wi.x + wi.y * wi.y;
  }
  if (partial) {
    double2 acc[1] = {make_double2(nrm, 0)};
    block_reduce_store<1>(acc, 1, partial + (int64_t)blockIdx.x * pstride);
  }
}

// final packing: out[0..nvec) = h1 + h2, out[nvec] = (w0sq, beta_sq)
__global__ void pack_orth(int nvec, const double2 *h1, const double2 *h2, const double2 *w0sq,
                          const double2 *bsq, double2 *out) {
  for (int v = threadIdx.x; v < nvec; v += blockDim.x)
    out[v] = make_double2(h1[v].x + h2[v].x, h1[v].y + h2[v].y);
  if (threadIdx.x == 0) out[nvec] = make_double2(w0sq->x, bsq->x);
}

__global__ __launch_bounds__(RT) void combine_kernel(int64_t n, const double2 *Q, int64_t ldq,
                                                     int nvec, const double2 *c, double2 *out) {
  __shared__ double2 cs[MAXV];
  for (int v = threadIdx.x; v < nvec; v += RT) cs[v] = c[v];
  __syncthreads();
  const int64_t stride = (int64_t)gridDim.x * RT;
  for (int64_t i = (int64_t)blockIdx.x * RT + threadIdx.x; i < n; i += stride) {
    double2 s = make_double2(0, 0);
    for (int v = 0; v < nvec; ++v) {
      const double2 q = Q[(int64_t)v * ldq + i];
      s.x += cs[v].x * q.x - cs[v].y * q.y;
      s.y += cs[v].x * q.y + cs[v].y * q.x;
    }
    out[i] = s;
  }
}

// ---- strided permutation copy ------------------------------------------------
struct PermArgs {
  int64_t dims[8];
  int64_t ss[8];
  int64_t ds[8];
  int rank;
  int conj;
  int64_t total;
};

template <typename TS, typename TD>
__global__ __launch_bounds__(256) void permute_kernel(PermArgs a, const TS *src, TD *dst) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t lin = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; lin < a.total;
       lin += stride) {
    int64_t rem = lin, so = 0, dof = 0;
    for (int d = a.rank - 1; d >= 0; --d) {
      const int64_t idx = rem % a.dims[d];
      rem /= a.dims[d];
      so += idx * a.ss[d];
      dof += idx * a.ds[d];
    }
    double re, im;
    ptn::cload(src + so, re, im);
    if (a.conj) im = -im;
    ptn::cstore(dst + dof, re, im);
  }
}

template <typename TD>
__global__ void fill_kernel(TD *dst, int64_t n, double re, double im) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    ptn::cstore(dst + i, re, im);
}

inline unsigned grid_for(int64_t total) {
  int64_t g = (total + 255) / 256;
  if (g < 1) g = 1;
  if (g > 65536) g = 65536;
  return (unsigned)g;
}

}  // namespace

extern "C" int64_t ptn_krylov_work_bytes(int64_t n, int32_t nvec) {
  // partial rows: MAXBLK x (MAXV + 2) complex, plus 2*MAXV + 4 complex scratch
  (void)n;
  (void)nvec;
  return (int64_t)sizeof(double2) * ((int64_t)MAXBLK * (MAXV + 2) + 2 * MAXV + 8);
}

extern "C" int ptn_krylov_orthogonalize(void *stream, int64_t n, const void *Qv, int64_t ldq,
                                        int32_t nvec, void *wv, void *outv, void *work) {
  if (n <= 0 || nvec < 1 || nvec > MAXV || ldq < n || !Qv || !wv || !outv || !work) {
    ptn::set_error("ptn_krylov_orthogonalize: bad arguments (n=%lld nvec=%d)", (long long)n,
                   nvec);
    return nvec > MAXV ? PTN_ERANGE : PTN_EINVAL;
  }
  hipStream_t s = static_cast<hipStream_t>(stream);
  const double2 *Q = static_cast<const double2 *>(Qv);
  double2 *w = static_cast<double2 *>(wv);
  double2 *out = static_cast<double2 *>(outv);
  double2 *partial = static_cast<double2 *>(work);
  const int pstride = MAXV + 2;
  double2 *scratch = partial + (int64_t)MAXBLK * pstride;
  double2 *h1 = scratch, *h2 = scratch + MAXV, *w0 = scratch + 2 * MAXV,
          *bsq = scratch + 2 * MAXV + 1;
  const int nb = nblocks_for(n);
  for (int pass = 0; pass < 2; ++pass) {
    double2 *h = pass == 0 ? h1 : h2;
    for (int v0 = 0; v0 < nvec; v0 += KV) {
      const int cnt = (nvec - v0) < KV ? (nvec - v0) : KV;
      const bool last = v0 + KV >= nvec;
      // the self-norm of w (pass 0 only) lands in slot nvec
      const int self_slot = (pass == 0 && last) ? 1 : -1;
      hipLaunchKernelGGL(dots_partial, dim3(nb), dim3(RT), 0, s, n, Q, ldq, v0, cnt, w,
                         partial, pstride, self_slot);
      PTN_LAUNCH_CHECK();
    }
    const int count = nvec + (pass == 0 ? 1 : 0);
    hipLaunchKernelGGL(reduce_partials, dim3(1), dim3(RT), 0, s, partial, pstride, nb, count,
                       h, 0);
    PTN_LAUNCH_CHECK();
    if (pass == 0) {
      // copy w0sq out of h1[nvec] before h1 is used as coefficients
      hipLaunchKernelGGL(reduce_partials, dim3(1), dim3(RT), 0, s, partial + nvec, pstride,
                         nb, 1, w0, 0);
      PTN_LAUNCH_CHECK();
    }
    hipLaunchKernelGGL(subtract_proj, dim3(nb), dim3(RT), 0, s, n, Q, ldq, nvec, h, w,
                       pass == 1 ? partial : nullptr, pstride);
    PTN_LAUNCH_CHECK();
  }
  hipLaunchKernelGGL(reduce_partials, dim3(1), dim3(RT), 0, s, partial, pstride, nb, 1, bsq, 0);
  PTN_LAUNCH_CHECK();
  hipLaunchKernelGGL(pack_orth, dim3(1), dim3(64), 0, s, nvec, h1, h2, w0, bsq, out);
  PTN_LAUNCH_CHECK();
  return PTN_OK;
}

extern "C" int ptn_krylov_combine(void *stream, int64_t n, const void *Q, int64_t ldq,
                                  int32_t nvec, const void *coeffs, void *out) {
  if (n <= 0 || nvec < 1 || nvec > MAXV || ldq < n || !Q || !coeffs || !out) {
    ptn::set_error("ptn_krylov_combine: bad arguments");
    return nvec > MAXV ? PTN_ERANGE : PTN_EINVAL;
  }
  hipStream_t s = static_cast<hipStream_t>(stream);
  hipLaunchKernelGGL(combine_kernel, dim3(nblocks_for(n)), dim3(RT), 0, s, n,
                     static_cast<const double2 *>(Q), ldq, nvec,
                     static_cast<const double2 *>(coeffs), static_cast<double2 *>(out));
  PTN_LAUNCH_CHECK();
  return PTN_OK;
}

extern "C" int ptn_norm_sq(void *stream, int64_t n, const void *x, void *out, void *work) {
  if (n <= 0 || !x || !out || !work) {
    ptn::set_error("ptn_norm_sq: bad arguments");
    return PTN_EINVAL;
  }
  hipStream_t s = static_cast<hipStream_t>(stream);
  const int nb = nblocks_for(n);
  double2 *partial = static_cast<double2 *>(work);
  const int pstride = MAXV + 2;
  // cnt = 0 dots, self norm only -> lands in slot 0
  hipLaunchKernelGGL(dots_partial, dim3(nb), dim3(RT), 0, s, n,
                     static_cast<const double2 *>(x), n, 0, 0,
                     static_cast<const double2 *>(x), partial, pstride, 1);
  PTN_LAUNCH_CHECK();
  hipLaunchKernelGGL(reduce_partials, dim3(1), dim3(RT), 0, s, partial, pstride, nb, 1,
                     static_cast<double2 *>(out), 0);
  PTN_LAUNCH_CHECK();
  return PTN_OK;
}

extern "C" int ptn_dotc(void *stream, int64_t n, const void *x, const void *y, void *out,
                        void *work) {
  if (n <= 0 || !x || !y || !out || !work) {
    ptn::set_error("ptn_dotc: bad arguments");
    return PTN_EINVAL;
  }
  hipStream_t s = static_cast<hipStream_t>(stream);
  const int nb = nblocks_for(n);
  double2 *partial = static_cast<double2 *>(work);
  const int pstride = MAXV + 2;
  hipLaunchKernelGGL(dots_partial, dim3(nb), dim3(RT), 0, s, n,
                     static_cast<const double2 *>(x), n, 0, 1,
                     static_cast<const double2 *>(y), partial, pstride, -1);
  PTN_LAUNCH_CHECK();
  hipLaunchKernelGGL(reduce_partials, dim3(1), dim3(RT), 0, s, partial, pstride, nb, 1,
                     static_cast<double2 *>(out), 0);
  PTN_LAUNCH_CHECK();
  return PTN_OK;
}

extern "C" int ptn_permute(void *stream, int32_t rank, const int64_t *dims, const void *src,
                           const int64_t *src_strides, int32_t src_dtype, void *dst,
                           const int64_t *dst_strides, int32_t dst_dtype, int32_t conj) {
  if (rank < 0 || rank > 8 || !src || !dst || (rank > 0 && (!dims || !src_strides || !dst_strides)) ||
      (unsigned)src_dtype > 1 || (unsigned)dst_dtype > 1) {
    ptn::set_error("ptn_permute: bad arguments");
    return PTN_EINVAL;
  }
  PermArgs a;
  memset(&a, 0, sizeof(a));
  a.rank = rank;
  a.conj = conj;
  a.total = 1;
  for (int i = 0; i < rank; ++i) {
    if (dims[i] < 0) {
      ptn::set_error("ptn_permute: negative dim");
      return PTN_EINVAL;
    }
    a.dims[i] = dims[i];
    a.ss[i] = src_strides[i];
    a.ds[i] = dst_strides[i];
    a.total *= dims[i];
  }
  if (a.total == 0) return PTN_OK;
  hipStream_t s = static_cast<hipStream_t>(stream);
  const unsigned g = grid_for(a.total);
  if (src_dtype == PTN_C128 && dst_dtype == PTN_C128)
    hipLaunchKernelGGL((permute_kernel<double2, double2>), dim3(g), dim3(256), 0, s, a,
                       static_cast<const double2 *>(src), static_cast<double2 *>(dst));
  else if (src_dtype == PTN_C128)
    hipLaunchKernelGGL((permute_kernel<double2, float2>), dim3(g), dim3(256), 0, s, a,
                       static_cast<const double2 *>(src), static_cast<float2 *>(dst));
  else if (dst_dtype == PTN_C128)
    hipLaunchKernelGGL((permute_kernel<float2, double2>), dim3(g), dim3(256), 0, s, a,
                       static_cast<const float2 *>(src), static_cast<double2 *>(dst));
  else
    hipLaunchKernelGGL((permute_kernel<float2, float2>), dim3(g), dim3(256), 0, s, a,
                       static_cast<const float2 *>(src), static_cast<float2 *>(dst));
  PTN_LAUNCH_CHECK();
  return PTN_OK;
}

extern "C" int ptn_fill(void *stream, void *dst, int64_t n, int32_t dtype, double re, double im) {
  if (n < 0 || !dst || (unsigned)dtype > 1) {
    ptn::set_error("ptn_fill: bad arguments");
    return PTN_EINVAL;
  }
  if (n == 0) return PTN_OK;
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (dtype == PTN_C128)
    hipLaunchKernelGGL((fill_kernel<double2>), dim3(grid_for(n)), dim3(256), 0, s,
                       static_cast<double2 *>(dst), n, re, im);
  else
    hipLaunchKernelGGL((fill_kernel<float2>), dim3(grid_for(n)), dim3(256), 0, s,
                       static_cast<float2 *>(dst), n, re, im);
  PTN_LAUNCH_CHECK();
  return PTN_OK;
}
