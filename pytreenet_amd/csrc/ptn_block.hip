// Block-sparse operator application: the step of an H_eff / environment
// contraction in which the TTNO node tensor meets a large state-side tensor.
//
//   out[o_r[r] + o_g[g] + o_j[j]] =
//       sum_{t in group g} sum_i W[t][j][i] * Y[y_r[r] + y_t[t] + y_i[i]]
//
// r runs over the large (row) index space, t over the live bond-state blocks of
// the TTNO tensor (sorted by output group g), j / i over the small physical
// indices (out / in).  One thread owns one row: it reads that row's inputs
// once (coalesced across the wavefront) and writes every live output block,
// so the step streams its operands at HBM speed instead of running tiny GEMMs.
#include "ptn_common.h"

namespace {

constexpr int BT = 256;
constexpr int MAXT = 128, MAXG = 128, MAXJ = 4, MAXI = 4;
constexpr int STAGE_BYTES = 60 * 1024;   // LDS for the operator blocks and staged inputs

struct BlockArgs {
  int64_t R;
  int nt, ng, nj, ni;
  int64_t y_i[MAXI];
  int64_t o_j[MAXJ];
  int conj_y, conj_w;
};

template <typename TY, typename TO>
__global__ __launch_bounds__(BT) void block_apply_kernel(BlockArgs a, const TY *Y,
                                                         const int64_t *y_r,
                                                         const int64_t *y_t,
                                                         const int32_t *g_begin,
                                                         const int64_t *o_g,
                                                         const double2 *W, TO *out,
                                                         const int64_t *o_r) {
  __shared__ double2 sW[MAXT * MAXJ * MAXI];
  __shared__ int64_t sYt[MAXT], sOg[MAXG];
  __shared__ int32_t sGb[MAXG + 1];
  const int tid = threadIdx.x;
  const int nw = a.nt * a.nj * a.ni;
  for (int k = tid; k < nw; k += BT) {
    double2 w = W[k];
    if (a.conj_w) w.y = -w.y;
    sW[k] = w;
  }
  for (int k = tid; k < a.nt; k += BT) sYt[k] = y_t[k];
  for (int k = tid; k < a.ng; k += BT) sOg[k] = o_g[k];
  for (int k = tid; k <= a.ng; k += BT) sGb[k] = g_begin[k];
  __syncthreads();
  const int64_t stride = (int64_t)gridDim.x * BT;
  for (int64_t r = (int64_t)blockIdx.x * BT + tid; r < a.R; r += stride) {
    const int64_t yb = y_r[r];
    const int64_t ob = o_r[r];
    for (int g = 0; g < a.ng; ++g) {
      double ar[MAXJ] = {0, 0, 0, 0}, ai[MAXJ] = {0, 0, 0, 0};
      for (int t = sGb[g]; t < sGb[g + 1]; ++t) {
        for (int i = 0; i < a.ni; ++i) {
          double yr, yi;
          ptn::cload(Y + yb + sYt[t] + a.y_i[i], yr, yi);
          if (a.conj_y) yi = -yi;
          const double2 *wt = sW + ((int64_t)t * a.nj) * a.ni + i;
#pragma unroll
          for (int j = 0; j < MAXJ; ++j) {
            if (j < a.nj) {
              const double2 w = wt[j * a.ni];
              ar[j] += w.x * yr - w.y * yi;
              ai[j] += w.x * yi + w.y * yr;
            }
          }
        }
      }
#pragma unroll
      for (int j = 0; j < MAXJ; ++j)
        if (j < a.nj) ptn::cstore(out + ob + sOg[g] + a.o_j[j], ar[j], ai[j]);
    }
  }
}

// The same with the rows' inputs staged through LDS.  Rows come in tiles of
// BT consecutive rows along the innermost row dimension (row r of a tile at
// y_r[r0] + (r - r0) s0); all inputs of a tile lie in one span of `span`
// elements from y_r[r0] + minoff, which the workgroup copies with coalesced
// loads before each thread reads its row from LDS.  Dynamic LDS: the operator
// blocks (nt nj ni complex128) and then the span.
// LDS-only workgroup barrier: this thread's LDS accesses are complete, then
// s_barrier.  Unlike __syncthreads() it does not wait for the global stores in
// flight (vmcnt also counts stores here), so a tile's output stores drain
// while the next tile's inputs load.  The memory clobber keeps the compiler
// from moving memory accesses across it.
constexpr int STAGE_ILP = 16;   // staged loads per thread in flight
template <typename T>
struct StageScalar;
template <>
struct StageScalar<double2> {
  typedef double T;
};
template <>
struct StageScalar<float2> {
  typedef float T;
};

__device__ __forceinline__ void lds_barrier() {
  asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
}

template <typename TY, typename TO>
__global__ __launch_bounds__(BT) void block_apply_staged(BlockArgs a, const TY *Y,
                                                         const int64_t *y_r,
                                                         const int64_t *y_t,
                                                         const int32_t *g_begin,
                                                         const int64_t *o_g,
                                                         const double2 *W, TO *out,
                                                         const int64_t *o_r, int64_t s0,
                                                         int64_t minoff, int span) {
  extern __shared__ double2 dyn[];
  __shared__ int64_t sYt[MAXT], sOg[MAXG];
  __shared__ int32_t sGb[MAXG + 1];
  const int tid = threadIdx.x;
  const int nw = a.nt * a.nj * a.ni;
  double2 *sW = dyn;
  TY *sY = reinterpret_cast<TY *>(dyn + nw);
  for (int k = tid; k < nw; k += BT) {
    double2 w = W[k];
    if (a.conj_w) w.y = -w.y;
    sW[k] = w;
  }
  for (int k = tid; k < a.nt; k += BT) sYt[k] = y_t[k];
  for (int k = tid; k < a.ng; k += BT) sOg[k] = o_g[k];
  for (int k = tid; k <= a.ng; k += BT) sGb[k] = g_begin[k];
  const int64_t ntiles = a.R / BT;
  for (int64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    const int64_t r0 = tile * BT;
    const int64_t base = y_r[r0] + minoff;
    lds_barrier();  // every thread has read the previous tile's span
    // batches of STAGE_ILP loads per thread in flight before the first wait (a
    // plain copy loop waits for each load before its LDS write)
    // (scalar register arrays: an array of vector-typed elements goes to scratch;
    // the loads are unconditional at a clamped index inside the span)
    for (int e0 = 0; e0 < span; e0 += STAGE_ILP * BT) {
      typename StageScalar<TY>::T br[STAGE_ILP], bi[STAGE_ILP];
#pragma unroll
      for (int k = 0; k < STAGE_ILP; ++k) {
        const int e = e0 + k * BT + tid;
        const TY v = Y[base + (e < span ? e : span - 1)];
        br[k] = v.x;
        bi[k] = v.y;
      }
#pragma unroll
      for (int k = 0; k < STAGE_ILP; ++k) {
        const int e = e0 + k * BT + tid;
        if (e < span) ptn::cstore(sY + e, (double)br[k], (double)bi[k]);
      }
    }
    lds_barrier();  // the span is in LDS
    const int64_t r = r0 + tid;
    const int64_t yb = (int64_t)tid * s0 - minoff;
    const int64_t ob = o_r[r];
    for (int g = 0; g < a.ng; ++g) {
      double ar[MAXJ] = {0, 0, 0, 0}, ai[MAXJ] = {0, 0, 0, 0};
      for (int t = sGb[g]; t < sGb[g + 1]; ++t) {
        for (int i = 0; i < a.ni; ++i) {
          double yr, yi;
          ptn::cload(sY + yb + sYt[t] + a.y_i[i], yr, yi);
          if (a.conj_y) yi = -yi;
          const double2 *wt = sW + ((int64_t)t * a.nj) * a.ni + i;
#pragma unroll
          for (int j = 0; j < MAXJ; ++j) {
            if (j < a.nj) {
              const double2 w = wt[j * a.ni];
              ar[j] += w.x * yr - w.y * yi;
              ai[j] += w.x * yi + w.y * yr;
            }
          }
        }
      }
#pragma unroll
      for (int j = 0; j < MAXJ; ++j)
        if (j < a.nj) ptn::cstore(out + ob + sOg[g] + a.o_j[j], ar[j], ai[j]);
    }
  }
}

}  // namespace

extern "C" int ptn_block_apply(void *stream, int64_t R, int32_t nt, int32_t ng, int32_t nj,
                               int32_t ni, const void *Y, int32_t y_dtype, const int64_t *y_r,
                               const int64_t *y_t, const int64_t *y_i, int32_t conj_y,
                               const int32_t *g_begin, const void *W, int32_t conj_w,
                               void *out, int32_t out_dtype, const int64_t *o_r,
                               const int64_t *o_g, const int64_t *o_j) {
  if (R < 0 || nt < 1 || nt > MAXT || ng < 1 || ng > MAXG || nj < 1 || nj > MAXJ || ni < 1 ||
      ni > MAXI || !Y || !y_r || !y_t || !y_i || !g_begin || !W || !out || !o_r || !o_g ||
      !o_j || (unsigned)y_dtype > 1 || (unsigned)out_dtype > 1) {
    ptn::set_error("ptn_block_apply: bad arguments (nt=%d ng=%d nj=%d ni=%d)", nt, ng, nj, ni);
    return (nt > MAXT || ng > MAXG || nj > MAXJ || ni > MAXI) ? PTN_ERANGE : PTN_EINVAL;
  }
  if (R == 0) return PTN_OK;
  BlockArgs a;
  a.R = R;
  a.nt = nt;
  a.ng = ng;
  a.nj = nj;
  a.ni = ni;
  for (int i = 0; i < MAXI; ++i) a.y_i[i] = i < ni ? y_i[i] : 0;
  for (int j = 0; j < MAXJ; ++j) a.o_j[j] = j < nj ? o_j[j] : 0;
  a.conj_y = conj_y;
  a.conj_w = conj_w;
  int64_t g = (R + BT - 1) / BT;
  if (g > 8192) g = 8192;
  hipStream_t s = static_cast<hipStream_t>(stream);
  const double2 *w = static_cast<const double2 *>(W);
  if (y_dtype == PTN_C128 && out_dtype == PTN_C128)
    hipLaunchKernelGGL((block_apply_kernel<double2, double2>), dim3((unsigned)g), dim3(BT), 0, s,
                       a, static_cast<const double2 *>(Y), y_r, y_t, g_begin, o_g, w,
                       static_cast<double2 *>(out), o_r);
  else if (y_dtype == PTN_C128)
    hipLaunchKernelGGL((block_apply_kernel<double2, float2>), dim3((unsigned)g), dim3(BT), 0, s,
                       a, static_cast<const double2 *>(Y), y_r, y_t, g_begin, o_g, w,
                       static_cast<float2 *>(out), o_r);
  else if (out_dtype == PTN_C128)
    hipLaunchKernelGGL((block_apply_kernel<float2, double2>), dim3((unsigned)g), dim3(BT), 0, s,
                       a, static_cast<const float2 *>(Y), y_r, y_t, g_begin, o_g, w,
                       static_cast<double2 *>(out), o_r);
  else
    hipLaunchKernelGGL((block_apply_kernel<float2, float2>), dim3((unsigned)g), dim3(BT), 0, s,
                       a, static_cast<const float2 *>(Y), y_r, y_t, g_begin, o_g, w,
                       static_cast<float2 *>(out), o_r);
  PTN_LAUNCH_CHECK();
  return PTN_OK;
}

extern "C" int ptn_block_apply_staged(void *stream, int64_t R, int32_t nt, int32_t ng,
                                      int32_t nj, int32_t ni, const void *Y, int32_t y_dtype,
                                      const int64_t *y_r, const int64_t *y_t, const int64_t *y_i,
                                      int32_t conj_y, const int32_t *g_begin, const void *W,
                                      int32_t conj_w, void *out, int32_t out_dtype,
                                      const int64_t *o_r, const int64_t *o_g, const int64_t *o_j,
                                      int64_t row_stride, int64_t min_offset, int32_t span) {
  const int64_t ybytes = y_dtype == PTN_C128 ? 16 : 8;
  if (R < 0 || R % BT != 0 || nt < 1 || nt > MAXT || ng < 1 || ng > MAXG || nj < 1 ||
      nj > MAXJ || ni < 1 || ni > MAXI || !Y || !y_r || !y_t || !y_i || !g_begin || !W || !out ||
      !o_r || !o_g || !o_j || (unsigned)y_dtype > 1 || (unsigned)out_dtype > 1 || span < 1 ||
      span * ybytes + (int64_t)nt * nj * ni * 16 > STAGE_BYTES || row_stride < 0) {
    ptn::set_error("ptn_block_apply_staged: bad arguments (R=%lld nt=%d ng=%d nj=%d ni=%d "
                   "span=%d; R a multiple of %d, span within %d bytes)", (long long)R, nt, ng,
                   nj, ni, span, BT, STAGE_BYTES);
    return (nt > MAXT || ng > MAXG || nj > MAXJ || ni > MAXI) ? PTN_ERANGE : PTN_EINVAL;
  }
  if (R == 0) return PTN_OK;
  BlockArgs a;
  a.R = R;
  a.nt = nt;
  a.ng = ng;
  a.nj = nj;
  a.ni = ni;
  for (int i = 0; i < MAXI; ++i) a.y_i[i] = i < ni ? y_i[i] : 0;
  for (int j = 0; j < MAXJ; ++j) a.o_j[j] = j < nj ? o_j[j] : 0;
  a.conj_y = conj_y;
  a.conj_w = conj_w;
  int64_t g = R / BT;
  if (g > 8192) g = 8192;
  const size_t dyn = (size_t)nt * nj * ni * sizeof(double2) + (size_t)span * ybytes;
  hipStream_t s = static_cast<hipStream_t>(stream);
  const double2 *w = static_cast<const double2 *>(W);
  const dim3 grid((unsigned)g), blk(BT);
  if (y_dtype == PTN_C128 && out_dtype == PTN_C128)
    hipLaunchKernelGGL((block_apply_staged<double2, double2>), grid, blk, dyn, s, a,
                       static_cast<const double2 *>(Y), y_r, y_t, g_begin, o_g, w,
                       static_cast<double2 *>(out), o_r, row_stride, min_offset, span);
  else if (y_dtype == PTN_C128)
    hipLaunchKernelGGL((block_apply_staged<double2, float2>), grid, blk, dyn, s, a,
                       static_cast<const double2 *>(Y), y_r, y_t, g_begin, o_g, w,
                       static_cast<float2 *>(out), o_r, row_stride, min_offset, span);
  else if (out_dtype == PTN_C128)
    hipLaunchKernelGGL((block_apply_staged<float2, double2>), grid, blk, dyn, s, a,
                       static_cast<const float2 *>(Y), y_r, y_t, g_begin, o_g, w,
                       static_cast<double2 *>(out), o_r, row_stride, min_offset, span);
  else
    hipLaunchKernelGGL((block_apply_staged<float2, float2>), grid, blk, dyn, s, a,
                       static_cast<const float2 *>(Y), y_r, y_t, g_begin, o_g, w,
                       static_cast<float2 *>(out), o_r, row_stride, min_offset, span);
  PTN_LAUNCH_CHECK();
  return PTN_OK;
}
