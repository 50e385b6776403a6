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

#include <mutex>
#include <vector>

namespace ptn {
static thread_local char g_err[512];
void set_error(const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}
}  // namespace ptn

namespace ptn {
namespace {
struct StageSlot {
  void *host = nullptr;
  size_t cap = 0;
  hipEvent_t ev = nullptr;
  bool pending = false;
};
constexpr int NSLOT = 8;
std::mutex g_stage_mu;
StageSlot g_slots[NSLOT];
int g_next_slot = 0;
}  // namespace

int stage_to_device(hipStream_t s, const void *src, size_t bytes, void *dst) {
  if (bytes == 0) return PTN_OK;
  std::lock_guard<std::mutex> lk(g_stage_mu);
  StageSlot &sl = g_slots[g_next_slot];
  g_next_slot = (g_next_slot + 1) % NSLOT;
  if (sl.pending) {  // the copy that last read this slot has finished
    PTN_HIP_CHECK(hipEventSynchronize(sl.ev));
    sl.pending = false;
  }
  if (sl.cap < bytes) {
    if (sl.host) PTN_HIP_CHECK(hipHostFree(sl.host));
    sl.host = nullptr;
    sl.cap = 0;
    const size_t cap = ((bytes + 65535) / 65536) * 65536;
    PTN_HIP_CHECK(hipHostMalloc(&sl.host, cap, hipHostMallocDefault));
    sl.cap = cap;
  }
  if (!sl.ev) PTN_HIP_CHECK(hipEventCreateWithFlags(&sl.ev, hipEventDisableTiming));
  memcpy(sl.host, src, bytes);
  PTN_HIP_CHECK(hipMemcpyAsync(dst, sl.host, bytes, hipMemcpyHostToDevice, s));
  PTN_HIP_CHECK(hipEventRecord(sl.ev, s));
  sl.pending = true;
  return PTN_OK;
}
}  // namespace ptn

extern "C" const char *ptn_last_error(void) { return ptn::g_err; }
extern "C" int ptn_version(void) { return 100; }

namespace {

constexpr int RT = 256;       // threads per reduction block
constexpr int KV = 8;         // vectors per sweep over w
constexpr int MAXV = 128;     // max basis vectors per call
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

// streaming accesses of the long-vector sweeps: every basis vector and w is
// read once per sweep and the written vector is not read again within it, so
// the loads and stores are nontemporal (the 537 MB vectors of a large local
// space do not fit any cache level anyway; measured 3-10% faster sweeps)
#ifndef PTN_NT_SWEEP
#define PTN_NT_SWEEP 1
#endif
typedef double d2v __attribute__((ext_vector_type(2)));
__device__ __forceinline__ double2 sweep_load(const double2 *p) {
#if PTN_NT_SWEEP
  const d2v v = __builtin_nontemporal_load(reinterpret_cast<const d2v *>(p));
  return make_double2(v.x, v.y);
#else
  return *p;
#endif
}
__device__ __forceinline__ void sweep_store(double2 *p, double2 x) {
#if PTN_NT_SWEEP
  d2v v = {x.x, x.y};
  __builtin_nontemporal_store(v, reinterpret_cast<d2v *>(p));
#else
  *p = x;
#endif
}

// block-wide sum of `cnt` complex values per thread; thread v < cnt stores the
// v-th sum to dst[v].  NT is the block size (a multiple of 64).
template <int NT, int CNT>
__device__ void block_sum(double2 (&acc)[CNT], int cnt, double2 *dst) {
  __shared__ double2 red[NT / 64][CNT];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int v = 0; v < CNT; ++v) {
    if (v < cnt) {  // uniform: cnt is the same in every thread
      double x = acc[v].x, y = acc[v].y;
      for (int off = 32; off > 0; off >>= 1) {
        x += __shfl_down(x, off, 64);
        y += __shfl_down(y, off, 64);
      }
      if (lane == 0) red[wave][v] = make_double2(x, y);
    }
  }
  __syncthreads();
  if (threadIdx.x < cnt) {
    double2 s = make_double2(0, 0);
    for (int w = 0; w < NT / 64; ++w) {
      s.x += red[w][threadIdx.x].x;
      s.y += red[w][threadIdx.x].y;
    }
    dst[threadIdx.x] = s;
  }
  __syncthreads();
}

template <int CNT>
__device__ void block_reduce_store(double2 (&acc)[CNT], int cnt, double2 *out_row) {
  block_sum<RT, CNT>(acc, cnt, out_row);
}

// partial[b][v] = sum over block b's range of conj(Q[v0+v]) * w, v < cnt;
// if self_slot >= 0 the slot `self_slot` gets sum |w|^2 (imag 0).
__device__ __forceinline__ void dots_partial_body(int64_t n, const double2 *Q, int64_t ldq,
                                                  int v0, int cnt, const double2 *w,
                                                  double2 *partial, int pstride, int self_slot,
                                                  int blk, int nblk) {
  double2 acc[KV + 1];
#pragma unroll
  for (int v = 0; v <= KV; ++v) acc[v] = make_double2(0, 0);
  const int64_t stride = (int64_t)nblk * RT;
  for (int64_t i = (int64_t)blk * RT + threadIdx.x; i < n; i += stride) {
    const double2 wi = sweep_load(w + i);
#pragma unroll
    for (int v = 0; v < KV; ++v) {
      if (v < cnt) {
        const double2 q = sweep_load(Q + (int64_t)(v0 + v) * ldq + i);
        const double2 t = cmul_conj_a(q, wi);
        acc[v].x += t.x;
        acc[v].y += t.y;
      }
    }
    if (self_slot >= 0) acc[KV].x += wi.x * wi.x + wi.y * wi.y;
  }
  double2 *row = partial + (int64_t)blk * pstride;
  // move the self term next to the dots so one reduction covers both
  if (self_slot >= 0) {
#pragma unroll
    for (int v = 0; v < KV; ++v)
      if (v == cnt) acc[v] = acc[KV];
  }
  block_reduce_store<KV + 1>(acc, cnt + (self_slot >= 0 ? 1 : 0), row + v0);
}

__global__ __launch_bounds__(RT) void dots_partial(int64_t n, const double2 *Q, int64_t ldq,
                                                   int v0, int cnt, const double2 *w,
                                                   double2 *partial, int pstride,
                                                   int self_slot) {
  dots_partial_body(n, Q, ldq, v0, cnt, w, partial, pstride, self_slot, blockIdx.x, gridDim.x);
}

// out[v] = sum_b partial[b][v]; one block per v, fixed summation order
__device__ __forceinline__ void reduce_cols_body(const double2 *partial, int pstride, int nblk,
                                                 double2 *out, int v) {
  double2 acc[1] = {make_double2(0, 0)};
  for (int b = threadIdx.x; b < nblk; b += RT) {
    const double2 t = partial[(int64_t)b * pstride + v];
    acc[0].x += t.x;
    acc[0].y += t.y;
  }
  block_sum<RT, 1>(acc, 1, out + v);
}

__global__ __launch_bounds__(RT) void reduce_cols(const double2 *partial, int pstride, int nblk,
                                                  double2 *out) {
  reduce_cols_body(partial, pstride, nblk, out, blockIdx.x);
}

// bsq = sum_b partial[b][0].x; writes out[0..nvec) = h1 + h2 and
// out[nvec] = (w0sq, bsq) (nvec > 0), or out[0] = (bsq, 0) (nvec == 0), and
// inv = 1 / sqrt(bsq) (0 when bsq <= 0)
__device__ __forceinline__ void finish_body(const double2 *partial, int pstride, int nblk,
                                            int nvec, const double2 *h1, const double2 *h2,
                                            double2 *out, double2 *inv) {
  __shared__ double2 tot;
  double2 acc[1] = {make_double2(0, 0)};
  for (int b = threadIdx.x; b < nblk; b += RT) acc[0].x += partial[(int64_t)b * pstride].x;
  block_sum<RT, 1>(acc, 1, &tot);
  const double bsq = tot.x;
  for (int v = threadIdx.x; v < nvec; v += RT)
    out[v] = make_double2(h1[v].x + h2[v].x, h1[v].y + h2[v].y);
  if (threadIdx.x == 0) {
    out[nvec] = nvec > 0 ? make_double2(h1[nvec].x, bsq) : make_double2(bsq, 0.0);
    *inv = make_double2(bsq > 0.0 ? 1.0 / sqrt(bsq) : 0.0, 0.0);
  }
}

__global__ __launch_bounds__(RT) void finish_kernel(const double2 *partial, int pstride,
                                                    int nblk, int nvec, const double2 *h1,
                                                    const double2 *h2, double2 *out,
                                                    double2 *inv) {
  finish_body(partial, pstride, nblk, nvec, h1, h2, out, inv);
}

// y = x * inv.x
__device__ __forceinline__ void scale_body(int64_t n, const double2 *x, const double2 *inv,
                                           double2 *y, int blk, int nblk) {
  const double c = inv->x;
  const int64_t stride = (int64_t)nblk * RT;
  for (int64_t i = (int64_t)blk * RT + threadIdx.x; i < n; i += stride) {
    const double2 v = x[i];
    y[i] = make_double2(c * v.x, c * v.y);
  }
}

__global__ __launch_bounds__(RT) void scale_kernel(int64_t n, const double2 *x,
                                                   const double2 *inv, double2 *y) {
  scale_body(n, x, inv, y, blockIdx.x, gridDim.x);
}

// ---- one-workgroup forms for short vectors (n <= SMALL_N) --------------------
// Whole CGS2 step in one launch: h1 = Q^H w, w -= Q h1, h2 = Q^H w, w -= Q h2,
// out = (h1 + h2, (|w_in|^2, |w_out|^2)), qnext = w / |w| when given.  Every
// thread owns the same elements of w in every phase, so w needs no barrier;
// the coefficients pass through LDS.
constexpr int ST = 1024;
constexpr int64_t SMALL_N = 8192;

__device__ __forceinline__ void orth_small_body(int64_t n, const double2 *Q, int64_t ldq,
                                                int nvec, double2 *w, double2 *qnext,
                                                double2 *out) {
  __shared__ double2 h[2][MAXV + 1];
  __shared__ double2 bsq;
  for (int pass = 0; pass < 2; ++pass) {
    for (int v0 = 0; v0 < nvec; v0 += KV) {
      const int cnt = (nvec - v0) < KV ? (nvec - v0) : KV;
      const bool self = pass == 0 && v0 + KV >= nvec;
      double2 acc[KV + 1];
#pragma unroll
      for (int v = 0; v <= KV; ++v) acc[v] = make_double2(0, 0);
      for (int64_t i = threadIdx.x; i < n; i += ST) {
        const double2 wi = w[i];
#pragma unroll
        for (int v = 0; v < KV; ++v) {
          if (v < cnt) {
            const double2 t = cmul_conj_a(Q[(int64_t)(v0 + v) * ldq + i], wi);
            acc[v].x += t.x;
            acc[v].y += t.y;
          }
        }
        if (self) acc[KV].x += wi.x * wi.x + wi.y * wi.y;
      }
      if (self) {
#pragma unroll
        for (int v = 0; v < KV; ++v)
          if (v == cnt) acc[v] = acc[KV];
      }
      block_sum<ST, KV + 1>(acc, cnt + (self ? 1 : 0), &h[pass][v0]);
    }
    double nrm = 0;
    for (int64_t i = threadIdx.x; i < n; i += ST) {
      double2 wi = w[i];
      for (int v = 0; v < nvec; ++v) {
        const double2 q = Q[(int64_t)v * ldq + i];
        const double2 hv = h[pass][v];
        wi.x -= hv.x * q.x - hv.y * q.y;
        wi.y -= hv.x * q.y + hv.y * q.x;
      }
      w[i] = wi;
      nrm += wi.x * wi.x + wi.y * wi.y;
    }
    if (pass == 1) {
      double2 acc[1] = {make_double2(nrm, 0)};
      block_sum<ST, 1>(acc, 1, &bsq);
    }
  }
  for (int v = threadIdx.x; v < nvec; v += ST)
    out[v] = make_double2(h[0][v].x + h[1][v].x, h[0][v].y + h[1][v].y);
  if (threadIdx.x == 0) out[nvec] = make_double2(h[0][nvec].x, bsq.x);
  if (qnext) {
    const double c = bsq.x > 0.0 ? 1.0 / sqrt(bsq.x) : 0.0;
    for (int64_t i = threadIdx.x; i < n; i += ST) {
      const double2 wi = w[i];
      qnext[i] = make_double2(c * wi.x, c * wi.y);
    }
  }
}

__global__ __launch_bounds__(ST) void orth_small(int64_t n, const double2 *Q, int64_t ldq,
                                                 int nvec, double2 *w, double2 *qnext,
                                                 double2 *out) {
  orth_small_body(n, Q, ldq, nvec, w, qnext, out);
}

// y = x / |x|, out[0] = (|x|^2, 0)
__device__ __forceinline__ void normalize_small_body(int64_t n, const double2 *x, double2 *y,
                                                     double2 *out) {
  __shared__ double2 tot;
  double2 acc[1] = {make_double2(0, 0)};
  for (int64_t i = threadIdx.x; i < n; i += ST) {
    const double2 v = x[i];
    acc[0].x += v.x * v.x + v.y * v.y;
  }
  block_sum<ST, 1>(acc, 1, &tot);
  const double c = tot.x > 0.0 ? 1.0 / sqrt(tot.x) : 0.0;
  if (threadIdx.x == 0) out[0] = make_double2(tot.x, 0.0);
  for (int64_t i = threadIdx.x; i < n; i += ST) {
    const double2 v = x[i];
    y[i] = make_double2(c * v.x, c * v.y);
  }
}

__global__ __launch_bounds__(ST) void normalize_small(int64_t n, const double2 *x, double2 *y,
                                                      double2 *out) {
  normalize_small_body(n, x, y, out);
}

// w -= sum_v h[v] Q[v];  optionally partial |w|^2 into partial[b][0]
__device__ __forceinline__ void subtract_proj_body(int64_t n, const double2 *Q, int64_t ldq,
                                                   int nvec, const double2 *h, double2 *w,
                                                   double2 *partial, int pstride, int blk,
                                                   int nblk) {
  __shared__ double2 hs[MAXV];
  for (int v = threadIdx.x; v < nvec; v += RT) hs[v] = h[v];
  __syncthreads();
  double nrm = 0;
  const int64_t stride = (int64_t)nblk * RT;
  for (int64_t i = (int64_t)blk * RT + threadIdx.x; i < n; i += stride) {
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
    block_reduce_store<1>(acc, 1, partial + (int64_t)blk * pstride);
  }
}

__global__ __launch_bounds__(RT) void subtract_proj(int64_t n, const double2 *Q, int64_t ldq,
                                                    int nvec, const double2 *h, double2 *w,
                                                    double2 *partial, int pstride) {
  subtract_proj_body(n, Q, ldq, nvec, h, w, partial, pstride, blockIdx.x, gridDim.x);
}

// The first projection and the second pass's dot products in one sweep (for
// nvec <= FUSE_V): w' = w - sum_v h[v] Q[v] is stored and partial[b][v] =
// sum over block b's range of conj(Q[v]) w', with the Q values of an element
// held in registers between the two uses.  Same per-element order and the same
// block reduction as subtract_proj followed by dots_partial.
constexpr int FUSE_V = 16;
__global__ __launch_bounds__(RT) void subtract_dots(int64_t n, const double2 *Q, int64_t ldq,
                                                    int nvec, const double2 *h, double2 *w,
                                                    double2 *partial, int pstride) {
  __shared__ double2 hs[FUSE_V];
  for (int v = threadIdx.x; v < nvec; v += RT) hs[v] = h[v];
  __syncthreads();
  double2 acc[FUSE_V + 1];
#pragma unroll
  for (int v = 0; v <= FUSE_V; ++v) acc[v] = make_double2(0, 0);
  const int64_t stride = (int64_t)gridDim.x * RT;
  for (int64_t i = (int64_t)blockIdx.x * RT + threadIdx.x; i < n; i += stride) {
    double2 wi = sweep_load(w + i);
    double2 q[FUSE_V];
#pragma unroll
    for (int v = 0; v < FUSE_V; ++v)
      q[v] = v < nvec ? sweep_load(Q + (int64_t)v * ldq + i) : make_double2(0, 0);
#pragma unroll
    for (int v = 0; v < FUSE_V; ++v) {
      if (v < nvec) {
        const double2 hv = hs[v];
        wi.x -= hv.x * q[v].x - hv.y * q[v].y;
        wi.y -= hv.x * q[v].y + hv.y * q[v].x;
      }
    }
    sweep_store(w + i, wi);
#pragma unroll
    for (int v = 0; v < FUSE_V; ++v) {
      if (v < nvec) {
        const double2 t = cmul_conj_a(q[v], wi);
        acc[v].x += t.x;
        acc[v].y += t.y;
      }
    }
    acc[FUSE_V].x += wi.x * wi.x + wi.y * wi.y;
  }
  // |w'|^2 lands in column nvec, next to the dots
#pragma unroll
  for (int v = 0; v < FUSE_V; ++v)
    if (v == nvec) acc[v] = acc[FUSE_V];
  block_reduce_store<FUSE_V + 1>(acc, nvec + 1, partial + (int64_t)blockIdx.x * pstride);
}

// After subtract_dots: h2[0..nvec) = Q^H w', h2[nvec] = (|w'|^2, 0).  The second
// projection removes exactly |h2|^2 from |w'|^2 (Q orthonormal), so
// beta^2 = |w'|^2 - |h2|^2 is known before w'' = w' - Q h2 is formed; out and
// inv = 1 / beta as finish_kernel writes them.
__global__ __launch_bounds__(RT) void finish_pyth_kernel(int nvec, const double2 *h1,
                                                         const double2 *h2, double2 *out,
                                                         double2 *inv) {
  if (threadIdx.x == 0) {
    double bsq = h2[nvec].x;
    for (int v = 0; v < nvec; ++v) bsq -= h2[v].x * h2[v].x + h2[v].y * h2[v].y;
    if (!(bsq > 0.0)) bsq = 0.0;
    out[nvec] = make_double2(h1[nvec].x, bsq);
    *inv = make_double2(bsq > 0.0 ? 1.0 / sqrt(bsq) : 0.0, 0.0);
  }
  for (int v = threadIdx.x; v < nvec; v += RT)
    out[v] = make_double2(h1[v].x + h2[v].x, h1[v].y + h2[v].y);
}

// qnext = (w - sum_v h[v] Q[v]) * inv: the second subtraction and the
// normalisation in one sweep (w itself is left as it was)
__global__ __launch_bounds__(RT) void subtract_scale(int64_t n, const double2 *Q, int64_t ldq,
                                                     int nvec, const double2 *h,
                                                     const double2 *w, const double2 *inv,
                                                     double2 *qnext) {
  __shared__ double2 hs[FUSE_V];
  for (int v = threadIdx.x; v < nvec; v += RT) hs[v] = h[v];
  __syncthreads();
  const double c = inv->x;
  const int64_t stride = (int64_t)gridDim.x * RT;
  for (int64_t i = (int64_t)blockIdx.x * RT + threadIdx.x; i < n; i += stride) {
    double2 wi = sweep_load(w + i);
    // all of an element's basis values are requested before the first is used
    double2 q[FUSE_V];
#pragma unroll
    for (int v = 0; v < FUSE_V; ++v)
      q[v] = v < nvec ? sweep_load(Q + (int64_t)v * ldq + i) : make_double2(0, 0);
#pragma unroll
    for (int v = 0; v < FUSE_V; ++v) {
      if (v < nvec) {
        const double2 hv = hs[v];
        wi.x -= hv.x * q[v].x - hv.y * q[v].y;
        wi.y -= hv.x * q[v].y + hv.y * q[v].x;
      }
    }
    sweep_store(qnext + i, make_double2(c * wi.x, c * wi.y));
  }
}

__device__ __forceinline__ void combine_body(int64_t n, const double2 *Q, int64_t ldq, int nvec,
                                             const double2 *c, double2 *out, int blk, int nblk) {
  __shared__ double2 cs[MAXV];
  for (int v = threadIdx.x; v < nvec; v += RT) cs[v] = c[v];
  __syncthreads();
  const int64_t stride = (int64_t)nblk * RT;
  for (int64_t i = (int64_t)blk * RT + threadIdx.x; i < n; i += stride) {
    double2 s = make_double2(0, 0);
    for (int v = 0; v < nvec; ++v) {
      const double2 q = sweep_load(Q + (int64_t)v * ldq + i);
      s.x += cs[v].x * q.x - cs[v].y * q.y;
      s.y += cs[v].x * q.y + cs[v].y * q.x;
    }
    sweep_store(out + i, s);
  }
}

__global__ __launch_bounds__(RT) void combine_kernel(int64_t n, const double2 *Q, int64_t ldq,
                                                     int nvec, const double2 *c, double2 *out) {
  combine_body(n, Q, ldq, nvec, c, out, blockIdx.x, gridDim.x);
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

// ---- batched forms: one launch over the instances of a shape group ------------
// Every instance has its own vectors (same length n); per-instance pointers
// come from a device array of BVec.  Short vectors: one workgroup per
// instance.  Long vectors: blockIdx.y is the instance, blockIdx.x one of BNB
// slices, with per-instance partial rows and scratch in the work buffer.
struct BVec {
  const double2 *Q;  // basis rows (orthonormalize / combine) or x (normalize)
  int64_t ldq;
  double2 *w;        // vector to orthogonalise (orthonormalize) or y (normalize)
  double2 *qnext;    // may be null
  double2 *out;      // packed scalars
  const double2 *c;  // combine coefficients
};
constexpr int BNB = 32;                      // slices per instance (long vectors)
constexpr int BPSTRIDE = MAXV + 2;           // partial row length
constexpr int BSCR = 2 * (MAXV + 1) + 2;     // h1, h2, inv per instance

__device__ __forceinline__ double2 *bpart(double2 *work, int b) {
  return work + (int64_t)b * (BNB * BPSTRIDE + BSCR);
}

__global__ __launch_bounds__(ST) void orth_small_batched(const BVec *__restrict__ v, int64_t n,
                                                         int nvec) {
  const BVec d = v[blockIdx.x];
  orth_small_body(n, d.Q, d.ldq, nvec, d.w, d.qnext, d.out);
}

__global__ __launch_bounds__(ST) void normalize_small_batched(const BVec *__restrict__ v,
                                                              int64_t n) {
  const BVec d = v[blockIdx.x];
  normalize_small_body(n, d.Q, d.w, d.out);
}

__global__ __launch_bounds__(RT) void dots_partial_batched(const BVec *__restrict__ v, int64_t n,
                                                           int v0, int cnt, int self_slot,
                                                           int use_q, double2 *work) {
  const BVec d = v[blockIdx.y];
  const double2 *x = d.w ? d.w : d.Q;
  dots_partial_body(n, use_q ? d.Q : x, use_q ? d.ldq : n, v0, cnt, x, bpart(work, blockIdx.y),
                    BPSTRIDE, self_slot, blockIdx.x, gridDim.x);
}

__global__ __launch_bounds__(RT) void reduce_cols_batched(int nblk, int which, double2 *work) {
  double2 *p = bpart(work, blockIdx.y);
  double2 *h = p + BNB * BPSTRIDE + (which ? (MAXV + 1) : 0);
  reduce_cols_body(p, BPSTRIDE, nblk, h, blockIdx.x);
}

__global__ __launch_bounds__(RT) void subtract_proj_batched(const BVec *__restrict__ v, int64_t n,
                                                            int nvec, int which, int keep_norm,
                                                            double2 *work) {
  const BVec d = v[blockIdx.y];
  double2 *p = bpart(work, blockIdx.y);
  const double2 *h = p + BNB * BPSTRIDE + (which ? (MAXV + 1) : 0);
  subtract_proj_body(n, d.Q, d.ldq, nvec, h, d.w, keep_norm ? p : nullptr, BPSTRIDE, blockIdx.x,
                     gridDim.x);
}

__global__ __launch_bounds__(RT) void finish_batched(const BVec *__restrict__ v, int nblk, int nvec,
                                                     double2 *work) {
  const BVec d = v[blockIdx.x];
  double2 *p = bpart(work, blockIdx.x);
  double2 *h1 = p + BNB * BPSTRIDE, *h2 = h1 + (MAXV + 1), *inv = h2 + (MAXV + 1);
  finish_body(p, BPSTRIDE, nblk, nvec, h1, h2, d.out, inv);
}

// y = x * inv: orthonormalize (x = w, y = qnext) or normalize (x = Q, y = w)
__global__ __launch_bounds__(RT) void scale_batched(const BVec *__restrict__ v, int64_t n,
                                                    int normalize, double2 *work) {
  const BVec d = v[blockIdx.y];
  const double2 *inv = bpart(work, blockIdx.y) + BNB * BPSTRIDE + 2 * (MAXV + 1);
  if (normalize)
    scale_body(n, d.Q, inv, d.w, blockIdx.x, gridDim.x);
  else if (d.qnext)
    scale_body(n, d.w, inv, d.qnext, blockIdx.x, gridDim.x);
}

__global__ __launch_bounds__(RT) void combine_batched(const BVec *__restrict__ v, int64_t n,
                                                      int nvec) {
  const BVec d = v[blockIdx.y];
  combine_body(n, d.Q, d.ldq, nvec, d.c, d.out, blockIdx.x, gridDim.x);
}

template <typename TD>
__global__ void fill_batched(TD *const *__restrict__ dst, int64_t n, double re, double im) {
  TD *p = dst[blockIdx.y];
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    ptn::cstore(p + i, re, im);
}

template <typename TS, typename TD>
__global__ __launch_bounds__(256) void permute_batched(PermArgs a, const TS *const *__restrict__ src,
                                                       TD *const *__restrict__ dst) {
  const TS *sp = src[blockIdx.y];
  TD *dp = dst[blockIdx.y];
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
    ptn::cload(sp + so, re, im);
    if (a.conj) im = -im;
    ptn::cstore(dp + dof, re, im);
  }
}

}  // namespace

extern "C" int64_t ptn_krylov_work_bytes(int64_t n, int32_t nvec) {
  // partial rows: MAXBLK x (MAXV + 2) complex, plus 2 (MAXV + 1) + 1 complex scratch
  (void)n;
  (void)nvec;
  return (int64_t)sizeof(double2) * ((int64_t)MAXBLK * (MAXV + 2) + 2 * MAXV + 8);
}

extern "C" int ptn_krylov_orthonormalize(void *stream, int64_t n, const void *Qv, int64_t ldq,
                                         int32_t nvec, void *wv, void *qnextv, void *outv,
                                         void *work) {
  if (n <= 0 || nvec < 1 || nvec > MAXV || ldq < n || !Qv || !wv || !outv || !work) {
    ptn::set_error("ptn_krylov_orthonormalize: bad arguments (n=%lld nvec=%d)", (long long)n,
                   nvec);
    return nvec > MAXV ? PTN_ERANGE : PTN_EINVAL;
  }
  hipStream_t s = static_cast<hipStream_t>(stream);
  const double2 *Q = static_cast<const double2 *>(Qv);
  double2 *w = static_cast<double2 *>(wv);
  double2 *qnext = static_cast<double2 *>(qnextv);
  double2 *out = static_cast<double2 *>(outv);
  if (n <= SMALL_N) {
    hipLaunchKernelGGL(orth_small, dim3(1), dim3(ST), 0, s, n, Q, ldq, nvec, w, qnext, out);
    PTN_LAUNCH_CHECK();
    return PTN_OK;
  }
  double2 *partial = static_cast<double2 *>(work);
  const int pstride = MAXV + 2;
  double2 *scratch = partial + (int64_t)MAXBLK * pstride;
  double2 *h1 = scratch, *h2 = scratch + (MAXV + 1), *inv = scratch + 2 * (MAXV + 1);
  const int nb = nblocks_for(n);
  if (nvec <= FUSE_V) {
    // three sweeps: dots (+ |w|^2), subtract fused with the second dots,
    // subtract (+ |w'|^2)
    hipLaunchKernelGGL(dots_partial, dim3(nb), dim3(RT), 0, s, n, Q, ldq, 0,
                       nvec < KV ? nvec : KV, w, partial, pstride, nvec <= KV ? 1 : -1);
    PTN_LAUNCH_CHECK();
    if (nvec > KV) {
      hipLaunchKernelGGL(dots_partial, dim3(nb), dim3(RT), 0, s, n, Q, ldq, KV, nvec - KV, w,
                         partial, pstride, 1);
      PTN_LAUNCH_CHECK();
    }
    hipLaunchKernelGGL(reduce_cols, dim3(nvec + 1), dim3(RT), 0, s, partial, pstride, nb, h1);
    PTN_LAUNCH_CHECK();
    hipLaunchKernelGGL(subtract_dots, dim3(nb), dim3(RT), 0, s, n, Q, ldq, nvec, h1, w, partial,
                       pstride);
    PTN_LAUNCH_CHECK();
    hipLaunchKernelGGL(reduce_cols, dim3(nvec + 1), dim3(RT), 0, s, partial, pstride, nb, h2);
    PTN_LAUNCH_CHECK();
    if (qnext) {
      // beta from |w'|^2 - |h2|^2, then w'' and its normalisation in one sweep
      hipLaunchKernelGGL(finish_pyth_kernel, dim3(1), dim3(RT), 0, s, nvec, h1, h2, out, inv);
      PTN_LAUNCH_CHECK();
      hipLaunchKernelGGL(subtract_scale, dim3(nb), dim3(RT), 0, s, n, Q, ldq, nvec, h2, w, inv,
                         qnext);
      PTN_LAUNCH_CHECK();
      return PTN_OK;
    }
    hipLaunchKernelGGL(subtract_proj, dim3(nb), dim3(RT), 0, s, n, Q, ldq, nvec, h2, w, partial,
                       pstride);
    PTN_LAUNCH_CHECK();
  }
  for (int pass = 0; pass < 2 && nvec > FUSE_V; ++pass) {
    double2 *h = pass == 0 ? h1 : h2;
    for (int v0 = 0; v0 < nvec; v0 += KV) {
      const int cnt = (nvec - v0) < KV ? (nvec - v0) : KV;
      const bool last = v0 + KV >= nvec;
      // pass 0: |w|^2 lands in column nvec of the partial rows
      const int self_slot = (pass == 0 && last) ? 1 : -1;
      hipLaunchKernelGGL(dots_partial, dim3(nb), dim3(RT), 0, s, n, Q, ldq, v0, cnt, w,
                         partial, pstride, self_slot);
      PTN_LAUNCH_CHECK();
    }
    // h1[0..nvec) = Q^H w, h1[nvec] = |w_in|^2;  h2[0..nvec) = Q^H w (second pass)
    hipLaunchKernelGGL(reduce_cols, dim3(nvec + (pass == 0 ? 1 : 0)), dim3(RT), 0, s, partial,
                       pstride, nb, h);
    PTN_LAUNCH_CHECK();
    hipLaunchKernelGGL(subtract_proj, dim3(nb), dim3(RT), 0, s, n, Q, ldq, nvec, h, w,
                       pass == 1 ? partial : nullptr, pstride);
    PTN_LAUNCH_CHECK();
  }
  hipLaunchKernelGGL(finish_kernel, dim3(1), dim3(RT), 0, s, partial, pstride, nb, nvec, h1, h2,
                     out, inv);
  PTN_LAUNCH_CHECK();
  if (qnext) {
    hipLaunchKernelGGL(scale_kernel, dim3(nb), dim3(RT), 0, s, n, w, inv, qnext);
    PTN_LAUNCH_CHECK();
  }
  return PTN_OK;
}

extern "C" int ptn_krylov_orthogonalize(void *stream, int64_t n, const void *Q, int64_t ldq,
                                        int32_t nvec, void *w, void *out, void *work) {
  return ptn_krylov_orthonormalize(stream, n, Q, ldq, nvec, w, nullptr, out, work);
}

// ---- allocation-free orthonormalisation step -----------------------------------
// The work buffer of ptn_krylov_orthonormalize (per-block partial dot products and
// the reduced coefficients) as a file-scope device array: every GPU the library
// runs on holds its own copy (2.1 MB) until the process ends.  Host code never
// passes the array's name to a kernel: the name is a host-side stand-in, and each
// device's address comes from hipGetSymbolAddress, taken once per device with
// that device current and kept in noalloc_ptr.
namespace noalloc {
constexpr int64_t SCRATCH_ELEMS = (int64_t)MAXBLK * (MAXV + 2) + 2 * MAXV + 8;
__device__ __align__(256) double2 scratch[SCRATCH_ELEMS];
constexpr int MAX_DEVICES = 64;
std::mutex mu;
void *noalloc_ptr[MAX_DEVICES] = {};

int scratch_of_current_device(void **out) {
  int dev = -1;
  PTN_HIP_CHECK(hipGetDevice(&dev));
  if (dev < 0 || dev >= MAX_DEVICES) {
    ptn::set_error("ptn_krylov_orthonormalize_noalloc: device %d beyond %d", dev, MAX_DEVICES);
    return PTN_ERANGE;
  }
  std::lock_guard<std::mutex> lk(mu);
  if (!noalloc_ptr[dev]) {
    void *p = nullptr;
    PTN_HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    noalloc_ptr[dev] = p;
  }
  *out = noalloc_ptr[dev];
  return PTN_OK;
}
}  // namespace noalloc

extern "C" int ptn_krylov_orthonormalize_noalloc(void *stream, int64_t n, const void *Q,
                                                 int64_t ldq, int32_t nvec, void *w,
                                                 void *qnext, void *out) {
  if (nvec > MAXV) {
    ptn::set_error("ptn_krylov_orthonormalize_noalloc: nvec=%d beyond the maximum %d "
                   "(the library's scratch holds the partials of %d basis vectors)",
                   nvec, MAXV, MAXV);
    return PTN_ERANGE;
  }
  if (ptn_krylov_work_bytes(n, nvec) > (int64_t)sizeof(noalloc::scratch)) {
    ptn::set_error("ptn_krylov_orthonormalize_noalloc: needs %lld scratch bytes, the "
                   "library holds %lld", (long long)ptn_krylov_work_bytes(n, nvec),
                   (long long)sizeof(noalloc::scratch));
    return PTN_ERANGE;
  }
  void *ptr = nullptr;
  const int rc = noalloc::scratch_of_current_device(&ptr);
  if (rc) return rc;
  // every partial and coefficient the kernels read is written earlier in the same
  // call (no accumulation across calls), so the scratch needs no clearing
  return ptn_krylov_orthonormalize(stream, n, Q, ldq, nvec, w, qnext, out, ptr);
}

extern "C" int ptn_normalize(void *stream, int64_t n, const void *xv, void *yv, void *outv,
                             void *work) {
  if (n <= 0 || !xv || !yv || !outv || !work) {
    ptn::set_error("ptn_normalize: bad arguments");
    return PTN_EINVAL;
  }
  hipStream_t s = static_cast<hipStream_t>(stream);
  const double2 *x = static_cast<const double2 *>(xv);
  double2 *y = static_cast<double2 *>(yv);
  double2 *out = static_cast<double2 *>(outv);
  if (n <= SMALL_N) {
    hipLaunchKernelGGL(normalize_small, dim3(1), dim3(ST), 0, s, n, x, y, out);
    PTN_LAUNCH_CHECK();
    return PTN_OK;
  }
  double2 *partial = static_cast<double2 *>(work);
  const int pstride = MAXV + 2;
  double2 *inv = partial + (int64_t)MAXBLK * pstride + 2 * (MAXV + 1);
  const int nb = nblocks_for(n);
  // cnt = 0 dots: the self norm lands in column 0
  hipLaunchKernelGGL(dots_partial, dim3(nb), dim3(RT), 0, s, n, x, n, 0, 0, x, partial, pstride,
                     1);
  PTN_LAUNCH_CHECK();
  hipLaunchKernelGGL(finish_kernel, dim3(1), dim3(RT), 0, s, partial, pstride, nb, 0, nullptr,
                     nullptr, out, inv);
  PTN_LAUNCH_CHECK();
  hipLaunchKernelGGL(scale_kernel, dim3(nb), dim3(RT), 0, s, n, x, inv, y);
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
  // cnt = 0 dots, self norm only -> lands in column 0
  hipLaunchKernelGGL(dots_partial, dim3(nb), dim3(RT), 0, s, n,
                     static_cast<const double2 *>(x), n, 0, 0,
                     static_cast<const double2 *>(x), partial, pstride, 1);
  PTN_LAUNCH_CHECK();
  hipLaunchKernelGGL(reduce_cols, dim3(1), dim3(RT), 0, s, partial, pstride, nb,
                     static_cast<double2 *>(out));
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
  hipLaunchKernelGGL(reduce_cols, dim3(1), dim3(RT), 0, s, partial, pstride, nb,
                     static_cast<double2 *>(out));
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

// ---- batched entry points ---------------------------------------------------------
namespace {
constexpr int BMAX = 65535;  // instances per batched call (grid y extent)

int64_t bvec_bytes(int32_t count) { return (((int64_t)count * sizeof(BVec) + 255) / 256) * 256; }

int check_batch(int32_t count, const char *who) {
  if (count < 1 || count > BMAX) {
    ptn::set_error("%s: batch count %d outside [1, %d]", who, count, BMAX);
    return PTN_ERANGE;
  }
  return PTN_OK;
}

// stage the per-instance descriptors; returns the device copy
int stage_bvec(hipStream_t s, const std::vector<BVec> &h, void *work, BVec **dev) {
  *dev = static_cast<BVec *>(work);
  return ptn::stage_to_device(s, h.data(), h.size() * sizeof(BVec), work);
}

double2 *batch_scratch(void *work, int32_t count) {
  return reinterpret_cast<double2 *>(static_cast<char *>(work) + bvec_bytes(count));
}
}  // namespace

extern "C" int64_t ptn_krylov_batched_work_bytes(int32_t count) {
  if (count < 1) return 0;
  return bvec_bytes(count) + (int64_t)count * (BNB * BPSTRIDE + BSCR) * sizeof(double2) +
         (int64_t)count * MAXV * sizeof(double2);
}

extern "C" int ptn_krylov_orthonormalize_batched(void *stream, int64_t n, int32_t nvec,
                                                 int32_t count, const void *const *Q,
                                                 const int64_t *ldq, void *const *w,
                                                 void *const *qnext, void *const *out,
                                                 void *work) {
  int rc = check_batch(count, "ptn_krylov_orthonormalize_batched");
  if (rc) return rc;
  if (n <= 0 || nvec < 1 || nvec > MAXV || !Q || !ldq || !w || !out || !work) {
    ptn::set_error("ptn_krylov_orthonormalize_batched: bad arguments (n=%lld nvec=%d)",
                   (long long)n, nvec);
    return nvec > MAXV ? PTN_ERANGE : PTN_EINVAL;
  }
  std::vector<BVec> h(count);
  for (int b = 0; b < count; ++b) {
    if (!Q[b] || !w[b] || !out[b] || ldq[b] < n) {
      ptn::set_error("ptn_krylov_orthonormalize_batched: bad instance %d", b);
      return PTN_EINVAL;
    }
    h[b] = BVec{static_cast<const double2 *>(Q[b]), ldq[b], static_cast<double2 *>(w[b]),
                qnext ? static_cast<double2 *>(qnext[b]) : nullptr,
                static_cast<double2 *>(out[b]), nullptr};
  }
  hipStream_t s = static_cast<hipStream_t>(stream);
  BVec *dv;
  if ((rc = stage_bvec(s, h, work, &dv))) return rc;
  if (n <= SMALL_N) {
    hipLaunchKernelGGL(orth_small_batched, dim3(count), dim3(ST), 0, s, dv, n, nvec);
    PTN_LAUNCH_CHECK();
    return PTN_OK;
  }
  double2 *scr = batch_scratch(work, count);
  int nb = nblocks_for(n);
  if (nb > BNB) nb = BNB;
  for (int pass = 0; pass < 2; ++pass) {
    for (int v0 = 0; v0 < nvec; v0 += KV) {
      const int cnt = (nvec - v0) < KV ? (nvec - v0) : KV;
      const int self_slot = (pass == 0 && v0 + KV >= nvec) ? 1 : -1;
      hipLaunchKernelGGL(dots_partial_batched, dim3(nb, count), dim3(RT), 0, s, dv, n, v0, cnt,
                         self_slot, 1, scr);
      PTN_LAUNCH_CHECK();
    }
    hipLaunchKernelGGL(reduce_cols_batched, dim3(nvec + (pass == 0 ? 1 : 0), count), dim3(RT), 0,
                       s, nb, pass, scr);
    PTN_LAUNCH_CHECK();
    hipLaunchKernelGGL(subtract_proj_batched, dim3(nb, count), dim3(RT), 0, s, dv, n, nvec, pass,
                       pass, scr);
    PTN_LAUNCH_CHECK();
  }
  hipLaunchKernelGGL(finish_batched, dim3(count), dim3(RT), 0, s, dv, nb, nvec, scr);
  PTN_LAUNCH_CHECK();
  if (qnext) {
    hipLaunchKernelGGL(scale_batched, dim3(nb, count), dim3(RT), 0, s, dv, n, 0, scr);
    PTN_LAUNCH_CHECK();
  }
  return PTN_OK;
}

extern "C" int ptn_normalize_batched(void *stream, int64_t n, int32_t count,
                                     const void *const *x, void *const *y, void *const *out,
                                     void *work) {
  int rc = check_batch(count, "ptn_normalize_batched");
  if (rc) return rc;
  if (n <= 0 || !x || !y || !out || !work) {
    ptn::set_error("ptn_normalize_batched: bad arguments");
    return PTN_EINVAL;
  }
  std::vector<BVec> h(count);
  for (int b = 0; b < count; ++b) {
    if (!x[b] || !y[b] || !out[b]) {
      ptn::set_error("ptn_normalize_batched: bad instance %d", b);
      return PTN_EINVAL;
    }
    h[b] = BVec{static_cast<const double2 *>(x[b]), n, static_cast<double2 *>(y[b]), nullptr,
                static_cast<double2 *>(out[b]), nullptr};
  }
  hipStream_t s = static_cast<hipStream_t>(stream);
  BVec *dv;
  if ((rc = stage_bvec(s, h, work, &dv))) return rc;
  if (n <= SMALL_N) {
    hipLaunchKernelGGL(normalize_small_batched, dim3(count), dim3(ST), 0, s, dv, n);
    PTN_LAUNCH_CHECK();
    return PTN_OK;
  }
  double2 *scr = batch_scratch(work, count);
  int nb = nblocks_for(n);
  if (nb > BNB) nb = BNB;
  // self norm of x (= Q of the descriptor, w unused): dots with cnt = 0, column 0
  std::vector<BVec> hx(h);
  for (auto &e : hx) e.w = nullptr;
  BVec *dx = reinterpret_cast<BVec *>(reinterpret_cast<char *>(scr) +
                                      (int64_t)count * (BNB * BPSTRIDE + BSCR) * sizeof(double2));
  // the x-only descriptors go to the coefficient area (count * MAXV complex >= count BVec)
  static_assert(sizeof(BVec) <= MAXV * sizeof(double2), "descriptor size");
  if ((rc = ptn::stage_to_device(s, hx.data(), hx.size() * sizeof(BVec), dx))) return rc;
  hipLaunchKernelGGL(dots_partial_batched, dim3(nb, count), dim3(RT), 0, s, dx, n, 0, 0, 1, 0,
                     scr);
  PTN_LAUNCH_CHECK();
  hipLaunchKernelGGL(finish_batched, dim3(count), dim3(RT), 0, s, dv, nb, 0, scr);
  PTN_LAUNCH_CHECK();
  hipLaunchKernelGGL(scale_batched, dim3(nb, count), dim3(RT), 0, s, dv, n, 1, scr);
  PTN_LAUNCH_CHECK();
  return PTN_OK;
}

extern "C" int ptn_krylov_combine_batched(void *stream, int64_t n, int32_t nvec, int32_t count,
                                          const void *const *Q, const int64_t *ldq,
                                          const void *coeffs, void *const *out, void *work) {
  int rc = check_batch(count, "ptn_krylov_combine_batched");
  if (rc) return rc;
  if (n <= 0 || nvec < 1 || nvec > MAXV || !Q || !ldq || !coeffs || !out || !work) {
    ptn::set_error("ptn_krylov_combine_batched: bad arguments");
    return nvec > MAXV ? PTN_ERANGE : PTN_EINVAL;
  }
  hipStream_t s = static_cast<hipStream_t>(stream);
  double2 *scr = batch_scratch(work, count);
  double2 *cdev = scr + (int64_t)count * (BNB * BPSTRIDE + BSCR);
  if ((rc = ptn::stage_to_device(s, coeffs, (size_t)count * nvec * sizeof(double2), cdev)))
    return rc;
  std::vector<BVec> h(count);
  for (int b = 0; b < count; ++b) {
    if (!Q[b] || !out[b] || ldq[b] < n) {
      ptn::set_error("ptn_krylov_combine_batched: bad instance %d", b);
      return PTN_EINVAL;
    }
    h[b] = BVec{static_cast<const double2 *>(Q[b]), ldq[b], nullptr, nullptr,
                static_cast<double2 *>(out[b]), cdev + (int64_t)b * nvec};
  }
  BVec *dv;
  if ((rc = stage_bvec(s, h, work, &dv))) return rc;
  int nb = nblocks_for(n);
  if (nb > BNB) nb = BNB;
  hipLaunchKernelGGL(combine_batched, dim3(nb, count), dim3(RT), 0, s, dv, n, nvec);
  PTN_LAUNCH_CHECK();
  return PTN_OK;
}

extern "C" int64_t ptn_ptr_batch_work_bytes(int32_t count) {
  return count < 1 ? 0 : 2 * ((((int64_t)count * 8) + 255) / 256) * 256;
}

extern "C" int ptn_fill_batched(void *stream, int32_t count, void *const *dst, int64_t n,
                                int32_t dtype, double re, double im, void *work) {
  int rc = check_batch(count, "ptn_fill_batched");
  if (rc) return rc;
  if (n < 0 || !dst || (unsigned)dtype > 1 || !work) {
    ptn::set_error("ptn_fill_batched: bad arguments");
    return PTN_EINVAL;
  }
  if (n == 0) return PTN_OK;
  for (int b = 0; b < count; ++b)
    if (!dst[b]) {
      ptn::set_error("ptn_fill_batched: null destination %d", b);
      return PTN_EINVAL;
    }
  hipStream_t s = static_cast<hipStream_t>(stream);
  if ((rc = ptn::stage_to_device(s, dst, (size_t)count * sizeof(void *), work))) return rc;
  unsigned g = grid_for(n);
  if (g > 1024) g = 1024;
  if (dtype == PTN_C128)
    hipLaunchKernelGGL((fill_batched<double2>), dim3(g, count), dim3(256), 0, s,
                       static_cast<double2 *const *>(work), n, re, im);
  else
    hipLaunchKernelGGL((fill_batched<float2>), dim3(g, count), dim3(256), 0, s,
                       static_cast<float2 *const *>(work), n, re, im);
  PTN_LAUNCH_CHECK();
  return PTN_OK;
}

extern "C" int ptn_permute_batched(void *stream, int32_t count, int32_t rank, const int64_t *dims,
                                   const void *const *src, const int64_t *src_strides,
                                   int32_t src_dtype, void *const *dst,
                                   const int64_t *dst_strides, int32_t dst_dtype, int32_t conj,
                                   void *work) {
  int rc = check_batch(count, "ptn_permute_batched");
  if (rc) return rc;
  if (rank < 0 || rank > 8 || !src || !dst || !work ||
      (rank > 0 && (!dims || !src_strides || !dst_strides)) || (unsigned)src_dtype > 1 ||
      (unsigned)dst_dtype > 1) {
    ptn::set_error("ptn_permute_batched: bad arguments");
    return PTN_EINVAL;
  }
  PermArgs a;
  memset(&a, 0, sizeof(a));
  a.rank = rank;
  a.conj = conj;
  a.total = 1;
  for (int i = 0; i < rank; ++i) {
    if (dims[i] < 0) {
      ptn::set_error("ptn_permute_batched: negative dim");
      return PTN_EINVAL;
    }
    a.dims[i] = dims[i];
    a.ss[i] = src_strides[i];
    a.ds[i] = dst_strides[i];
    a.total *= dims[i];
  }
  if (a.total == 0) return PTN_OK;
  for (int b = 0; b < count; ++b)
    if (!src[b] || !dst[b]) {
      ptn::set_error("ptn_permute_batched: null pointer in instance %d", b);
      return PTN_EINVAL;
    }
  hipStream_t s = static_cast<hipStream_t>(stream);
  char *wb = static_cast<char *>(work);
  const int64_t half = ptn_ptr_batch_work_bytes(count) / 2;
  if ((rc = ptn::stage_to_device(s, src, (size_t)count * sizeof(void *), wb))) return rc;
  if ((rc = ptn::stage_to_device(s, dst, (size_t)count * sizeof(void *), wb + half))) return rc;
  unsigned g = grid_for(a.total);
  if (g > 1024) g = 1024;
  const dim3 grid(g, count);
  if (src_dtype == PTN_C128 && dst_dtype == PTN_C128)
    hipLaunchKernelGGL((permute_batched<double2, double2>), grid, dim3(256), 0, s, a,
                       reinterpret_cast<const double2 *const *>(wb),
                       reinterpret_cast<double2 *const *>(wb + half));
  else if (src_dtype == PTN_C128)
    hipLaunchKernelGGL((permute_batched<double2, float2>), grid, dim3(256), 0, s, a,
                       reinterpret_cast<const double2 *const *>(wb),
                       reinterpret_cast<float2 *const *>(wb + half));
  else if (dst_dtype == PTN_C128)
    hipLaunchKernelGGL((permute_batched<float2, double2>), grid, dim3(256), 0, s, a,
                       reinterpret_cast<const float2 *const *>(wb),
                       reinterpret_cast<double2 *const *>(wb + half));
  else
    hipLaunchKernelGGL((permute_batched<float2, float2>), grid, dim3(256), 0, s, a,
                       reinterpret_cast<const float2 *const *>(wb),
                       reinterpret_cast<float2 *const *>(wb + half));
  PTN_LAUNCH_CHECK();
  return PTN_OK;
}

// ---- whole Krylov exponential of a short vector in one workgroup -----------------
// For n <= 1024 and at most KMAX Krylov steps: q_0 = x/|x|, then per step
// w = H q_j, two classical Gram-Schmidt passes, beta = |w|, the stop rules of
// krylov_exp_action (breakdown at max(eps n, 1e-12 |H q_j|), Saad estimate
// |x| beta |[exp(f dt T_m) e_1]_{m-1}| < tol), q_{j+1} = w / beta; finally
// y = |x| Q exp(f dt T_m) e_1.
// H is column-major.  The matvec splits k into S slices (S = min(KXT / n, n)):
// thread (row r, slice s) sums H[r + k n] q[k] over k = s, s + S, ... with q_j
// staged in LDS, so a wave loads consecutive rows of one column; the slices
// are added through LDS.
// The small exponential is a degree-KTAYLOR Taylor polynomial of (f dt S)/2^s
// (1-norm <= 1/2, truncation below 3e-17) by Horner in two LDS buffers,
// squared s times (S = (T + T^H)/2 when hermitian, else T): the same column
// as the eigendecomposition form to rounding.
// info[0] = m used, info[1] = 1 when |x| = 0 (then y = x), info[2] = 1 when a
// stop rule ended the recurrence before max_steps.
namespace {
constexpr int KXT = 512;
constexpr int KMAX = 24;
constexpr int KNMAX = 1024;
constexpr int KTAYLOR = 14;
constexpr int KLD = KMAX + 1;

struct SmallLds {
  double2 T[KLD][KLD];
  double2 A[KLD][KLD];
  double2 E0[KLD][KLD];
  double2 E1[KLD][KLD];
  double2 q[KNMAX];
  double2 part[KXT];
  double2 hc[KMAX];
  double red[KXT / 64];
};

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// block sum of one double (every thread gets the result)
__device__ double block_sum1(double v, double *red) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double s = 0;
  for (int w = 0; w < KXT / 64; ++w) s += red[w];
  __syncthreads();
  return s;
}

// exp(c S) of the leading m x m block of S(T); returns the LDS buffer holding
// it (column 0 is the propagator column).  All threads call it.
__device__ double2 (*small_expm(int m, SmallLds &L, int hermitian, double2 c))[KLD] {
  const int tid = threadIdx.x;
  for (int t = tid; t < m * m; t += KXT) {
    const int r = t / m, k = t % m;
    double2 s = L.T[r][k];
    if (hermitian) {
      const double2 u = L.T[k][r];
      s = make_double2(0.5 * (s.x + u.x), 0.5 * (s.y - u.y));
    }
    L.A[r][k] = cmul(c, s);
  }
  __syncthreads();
  double nrm = 0;
  if (tid < m)
    for (int r = 0; r < m; ++r) nrm += sqrt(L.A[r][tid].x * L.A[r][tid].x + L.A[r][tid].y * L.A[r][tid].y);
  for (int off = 32; off > 0; off >>= 1) nrm = fmax(nrm, __shfl_down(nrm, off, 64));
  if ((tid & 63) == 0) L.red[tid >> 6] = nrm;
  __syncthreads();
  double amax = 0;
  for (int w = 0; w < KXT / 64; ++w) amax = fmax(amax, L.red[w]);
  int sq = 0;
  while (amax > 0.5 && sq < 60) {
    amax *= 0.5;
    ++sq;
  }
  const double scale = ldexp(1.0, -sq);
  for (int t = tid; t < m * m; t += KXT) {
    const int r = t / m, k = t % m;
    L.A[r][k].x *= scale;
    L.A[r][k].y *= scale;
    L.E0[r][k] = make_double2(r == k ? 1.0 : 0.0, 0.0);
  }
  __syncthreads();
  double2(*cur)[KLD] = L.E0;
  double2(*nxt)[KLD] = L.E1;
  // Horner: E <- I + A E / k for k = KTAYLOR .. 1
  for (int k = KTAYLOR; k >= 1; --k) {
    const double inv = 1.0 / k;
    for (int t = tid; t < m * m; t += KXT) {
      const int r = t / m, cc = t % m;
      double2 acc = make_double2(0, 0);
      for (int l = 0; l < m; ++l) {
        const double2 p = cmul(L.A[r][l], cur[l][cc]);
        acc.x += p.x;
        acc.y += p.y;
      }
      nxt[r][cc] = make_double2(acc.x * inv + (r == cc ? 1.0 : 0.0), acc.y * inv);
    }
    __syncthreads();
    double2(*t)[KLD] = cur;
    cur = nxt;
    nxt = t;
  }
  for (int q = 0; q < sq; ++q) {
    for (int t = tid; t < m * m; t += KXT) {
      const int r = t / m, cc = t % m;
      double2 acc = make_double2(0, 0);
      for (int l = 0; l < m; ++l) {
        const double2 p = cmul(cur[r][l], cur[l][cc]);
        acc.x += p.x;
        acc.y += p.y;
      }
      nxt[r][cc] = acc;
    }
    __syncthreads();
    double2(*t)[KLD] = cur;
    cur = nxt;
    nxt = t;
  }
  return cur;
}

__global__ __launch_bounds__(KXT) void krylov_small_kernel(
    int n, const double2 *__restrict__ H, const double2 *__restrict__ x, double2 *y,
    double2 *Q, int max_steps, double2 c, double tol, int hermitian, int *info) {
  __shared__ SmallLds L;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  for (int t = tid; t < KLD * KLD; t += KXT) L.T[t / KLD][t % KLD] = make_double2(0, 0);
  double loc = 0;
  for (int i = tid; i < n; i += KXT) loc += x[i].x * x[i].x + x[i].y * x[i].y;
  const double beta0 = sqrt(block_sum1(loc, L.red));
  if (beta0 == 0.0) {
    for (int i = tid; i < n; i += KXT) y[i] = x[i];
    if (tid == 0) {
      info[0] = 0;
      info[1] = 1;
      info[2] = 1;
    }
    return;
  }
  const double inv0 = 1.0 / beta0;
  for (int i = tid; i < n; i += KXT) Q[i] = make_double2(x[i].x * inv0, x[i].y * inv0);
  const double breakdown = 2.220446049250313e-16 * (double)(n > 1 ? n : 1);
  const int S = n <= KXT ? min(KXT / n, n) : 1;
  int m = 0, stopped = 0;
  for (int j = 0; j < max_steps; ++j) {
    __syncthreads();
    const double2 *qj = Q + (int64_t)j * n;
    double2 *w = Q + (int64_t)(j + 1) * n;
    for (int i = tid; i < n; i += KXT) L.q[i] = qj[i];
    __syncthreads();
    // w = H q_j
    if (S == 1) {
      for (int r = tid; r < n; r += KXT) {
        double2 acc = make_double2(0, 0);
#pragma unroll 8
        for (int k = 0; k < n; ++k) {
          const double2 p = cmul(H[r + (int64_t)k * n], L.q[k]);
          acc.x += p.x;
          acc.y += p.y;
        }
        w[r] = acc;
      }
    } else {
      if (tid < S * n) {
        const int r = tid % n, sl = tid / n;
        double2 acc = make_double2(0, 0);
#pragma unroll 8
        for (int k = sl; k < n; k += S) {
          const double2 p = cmul(H[r + (int64_t)k * n], L.q[k]);
          acc.x += p.x;
          acc.y += p.y;
        }
        L.part[tid] = acc;
      }
      __syncthreads();
      for (int r = tid; r < n; r += KXT) {
        double2 acc = make_double2(0, 0);
        for (int sl = 0; sl < S; ++sl) {
          acc.x += L.part[sl * n + r].x;
          acc.y += L.part[sl * n + r].y;
        }
        w[r] = acc;
      }
    }
    __syncthreads();
    loc = 0;
    for (int i = tid; i < n; i += KXT) loc += w[i].x * w[i].x + w[i].y * w[i].y;
    const double w0 = sqrt(block_sum1(loc, L.red));
    for (int pass = 0; pass < 2; ++pass) {
      // h_l = <q_l, w>, one wave per l
      for (int l = wave; l <= j; l += KXT / 64) {
        const double2 *ql = Q + (int64_t)l * n;
        double2 acc = make_double2(0, 0);
        for (int k = lane; k < n; k += 64) {
          const double2 q = ql[k], v = w[k];
          acc.x += q.x * v.x + q.y * v.y;
          acc.y += q.x * v.y - q.y * v.x;
        }
        for (int off = 32; off > 0; off >>= 1) {
          acc.x += __shfl_down(acc.x, off, 64);
          acc.y += __shfl_down(acc.y, off, 64);
        }
        if (lane == 0) L.hc[l] = acc;
      }
      __syncthreads();
      for (int i = tid; i < n; i += KXT) {
        double2 v = w[i];
        for (int l = 0; l <= j; ++l) {
          const double2 p = cmul(L.hc[l], Q[(int64_t)l * n + i]);
          v.x -= p.x;
          v.y -= p.y;
        }
        w[i] = v;
      }
      if (tid <= j) {
        L.T[tid][j].x += L.hc[tid].x;
        L.T[tid][j].y += L.hc[tid].y;
      }
      __syncthreads();
    }
    loc = 0;
    for (int i = tid; i < n; i += KXT) loc += w[i].x * w[i].x + w[i].y * w[i].y;
    const double beta = sqrt(block_sum1(loc, L.red));
    if (tid == 0) L.T[j + 1][j] = make_double2(beta, 0.0);
    m = j + 1;
    int stop = beta <= fmax(breakdown, 1e-12 * w0) ? 1 : 0;
    if (!stop && tol > 0.0) {
      __syncthreads();
      double2(*Ex)[KLD] = small_expm(m, L, hermitian, c);
      const double2 e = Ex[m - 1][0];
      if (beta0 * beta * sqrt(e.x * e.x + e.y * e.y) < tol) stop = 1;
    }
    if (stop) {
      stopped = 1;
      break;
    }
    const double ib = 1.0 / beta;
    for (int i = tid; i < n; i += KXT) w[i] = make_double2(w[i].x * ib, w[i].y * ib);
  }
  __syncthreads();
  double2(*Ex)[KLD] = small_expm(m, L, hermitian, c);
  for (int i = tid; i < n; i += KXT) {
    double2 acc = make_double2(0, 0);
    for (int l = 0; l < m; ++l) {
      const double2 p = cmul(Ex[l][0], Q[(int64_t)l * n + i]);
      acc.x += p.x;
      acc.y += p.y;
    }
    y[i] = make_double2(beta0 * acc.x, beta0 * acc.y);
  }
  if (tid == 0) {
    info[0] = m;
    info[1] = 0;
    info[2] = stopped;
  }
}
}  // namespace

extern "C" int64_t ptn_krylov_small_work_bytes(int64_t n, int32_t max_steps) {
  return (int64_t)sizeof(double2) * n * (max_steps + 1);
}

extern "C" int ptn_krylov_expm_small(void *stream, int64_t n, const void *H, const void *x,
                                     void *y, int32_t max_steps, double time_difference,
                                     double exp_re, double exp_im, double krylov_tol,
                                     int32_t hermitian, void *work, void *info) {
  if (n <= 0 || n > KNMAX || max_steps < 1 || max_steps > KMAX || max_steps > n || !H || !x ||
      !y || !work || !info) {
    ptn::set_error("ptn_krylov_expm_small: bad arguments (n=%lld max_steps=%d; n <= %d, "
                   "max_steps <= min(n, %d))", (long long)n, max_steps, KNMAX, KMAX);
    return (n > KNMAX || max_steps > KMAX) ? PTN_ERANGE : PTN_EINVAL;
  }
  hipStream_t s = static_cast<hipStream_t>(stream);
  const double2 c = make_double2(exp_re * time_difference, exp_im * time_difference);
  hipLaunchKernelGGL(krylov_small_kernel, dim3(1), dim3(KXT), 0, s, (int)n,
                     static_cast<const double2 *>(H), static_cast<const double2 *>(x),
                     static_cast<double2 *>(y), static_cast<double2 *>(work), max_steps, c,
                     krylov_tol, hermitian, static_cast<int *>(info));
  PTN_LAUNCH_CHECK();
  return PTN_OK;
}
