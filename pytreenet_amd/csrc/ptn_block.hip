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
