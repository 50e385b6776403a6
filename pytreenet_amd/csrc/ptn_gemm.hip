// Gather-GEMM: one pairwise complex tensor contraction on CDNA4 MFMA.
//
// C[c_off_m[m] + c_off_n[n]] = alpha * sum_k opA(A[..]) * opB(B[..]) (+ C)
//
// The tensor legs of A, B and C are flattened into (M, K), (K, N), (M, N)
// index spaces whose element offsets come from per-mode offset tables, so any
// leg permutation of a tensordot is absorbed by the global->LDS tile loads and
// the epilogue stores.  No transposed copy of an operand is written to HBM.
//
// Tiling: 256 threads = 4 wave64s in a 2x2 grid, block tile 64x64 (complex).
// Each wave owns a 32x32 sub-tile = 2x2 MFMA blocks of 16x16.
//   F64_4M (gemm_direct_f64_kernel): v_mfma_f64_16x16x4_f64, complex product
//           as 4 real MFMAs into (re, im) accumulators.  K step 8, tiles copied
//           global->LDS by global_load_lds_dwordx4 (lane l copies element
//           (k, m0 + l) of a tile row, whatever the operand's layout), three
//           LDS stages, two workgroups per CU.
//   F32_3M (gemm_gather_kernel): v_mfma_f32_16x16x4_f32 with the 3-multiply
//           (Gauss) split: T1 = ar*br, T2 = ai*bi, T3 = (ar+ai)(br+bi);
//           re = T1-T2, im = T3-T1-T2; operands are converted to float while
//           staging through registers (complex128 inputs are read as double and
//           rounded once).  K step 8, two LDS stages, padded LDS rows.
// MFMA operand layout (16x16x4): lane l supplies A[l%16][l/16] and
// B[l/16][l%16]; accumulator register r of lane l holds column l%16 and row
// 4*r + l/16 for the f64 instruction, row 4*(l/16) + r for the f32 one.
// Block -> tile mapping is XCD-aware: the hardware deals workgroups to the 8
// XCDs round-robin, the remap gives each XCD a contiguous tile range so tiles
// sharing an A row panel share one L2.
#include "ptn_common.h"

#include <stdio.h>

#include <vector>

namespace {

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

// register-staged kernel (the F32_3M path): K step, LDS row padding (in
// elements; breaks the bank conflicts of k-major tile stores) and the
// occupancy target in workgroups per CU
#ifndef PTN_GEMM_BK
#define PTN_GEMM_BK 8
#endif
#ifndef PTN_GEMM_PAD
#define PTN_GEMM_PAD 1
#endif
#ifndef PTN_GEMM_OCC
#define PTN_GEMM_OCC 3
#endif
constexpr int BM = 64, BN = 64, BK = PTN_GEMM_BK, NT = 256;
constexpr int LA = BM * BK / NT;  // A (and B) tile elements per thread per stage
static_assert(LA >= 1 && BM * BK % NT == 0, "tile / thread count mismatch");
constexpr int64_t SPLITK_MIN_K = 128;
constexpr int64_t SPLITK_TARGET_BLOCKS = 2048;  // 4 waves of 2 blocks x 256 CUs
constexpr int64_t LARGE_TILES = 2048;
#ifndef PTN_GEMM_GROUP_M
#define PTN_GEMM_GROUP_M 8
#endif
constexpr int64_t GROUP_M = PTN_GEMM_GROUP_M;   // tile rows per raster group

struct KArgs {
  int64_t M, N, K;
  const void *A;
  const int64_t *aom, *aok;
  const void *B;
  const int64_t *bok, *bon;
  void *C;
  const int64_t *com, *con;
  double are, aim;
  int64_t tiles_n;
  int64_t tiles_m;
  int64_t group_m;  // > 1: tile rows raster in groups of group_m (tile_coords)
  int64_t kchunk;   // K range per split (multiple of BK)
  double2 *ws;      // split-K partials [ksplit][M][N], null when gridDim.y == 1
  int aconj, akmaj, bconj, bkmaj, beta_one, remap;
  int lower;        // tiles on and below the diagonal only (square product, tm >= tn)
  int cmaj;         // consecutive m contiguous in C (store-order hint)
};

// (tile row, tile column) of linear tile index bid.  With group_m > 1 the tile
// rows come in groups of group_m and a group's tiles are numbered down its rows
// first: the workgroups in flight on one XCD (a contiguous index range after the
// remap) then share group_m A panels and a few B panels in that XCD's L2, where
// the row-major order cycles every workgroup row through all tiles_n B panels
// (5 MB for N = 1280 at K = 256, more than one XCD's 4 MB L2).
__device__ __forceinline__ void tile_coords(const KArgs &p, int64_t bid, int64_t &tm,
                                            int64_t &tn) {
  if (p.lower) {
    // bid = tm (tm + 1) / 2 + tn with tn <= tm: the lower triangle row by row
    int64_t r = (int64_t)((sqrt(8.0 * (double)bid + 1.0) - 1.0) * 0.5);
    while (r * (r + 1) / 2 > bid) --r;
    while ((r + 1) * (r + 2) / 2 <= bid) ++r;
    tm = r;
    tn = bid - r * (r + 1) / 2;
    return;
  }
  if (p.group_m > 1) {
    const int64_t span = p.group_m * p.tiles_n;
    const int64_t g = bid / span;
    const int64_t local = bid - g * span;
    const int64_t left = p.tiles_m - g * p.group_m;
    const int64_t rows = left < p.group_m ? left : p.group_m;
    tm = g * p.group_m + local % rows;
    tn = local / rows;
  } else {
    tm = bid / p.tiles_n;
    tn = bid - tm * p.tiles_n;
  }
}

template <int MODE>
struct Acc;
template <>
struct Acc<PTN_COMPUTE_F64_4M> {
  typedef double T;
  typedef v4d V;
};
template <>
struct Acc<PTN_COMPUTE_F32_3M> {
  typedef float T;
  typedef v4f V;
};

__device__ __forceinline__ v4d mfma(double a, double b, v4d c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ v4f mfma(float a, float b, v4f c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

// LARGE = 1 (here and in gemm_direct_f64_kernel) is the same kernel
// instantiated separately for launches with a large tile grid (>= LARGE_TILES
// tiles, no split-K), so profilers report the dominant H_eff / environment
// GEMMs apart from the many small ones.
// TR (FP32 only): the transposed block product (MFMA operands swapped), so the
// accumulators run along m and the epilogue stores are contiguous when C is
// m-contiguous.
template <int MODE, typename TA, typename TB, typename TC, bool TR = false>
__device__ __forceinline__ void gather_tile(const KArgs &p, int64_t bid, int64_t split,
                                            bool splitk) {
  static_assert(!TR || MODE == PTN_COMPUTE_F32_3M, "transposed stores: FP32 form only");
  typedef typename Acc<MODE>::T T;
  typedef typename Acc<MODE>::V V;
  __shared__ T sAr[2][BK][BM + PTN_GEMM_PAD];
  __shared__ T sAi[2][BK][BM + PTN_GEMM_PAD];
  __shared__ T sBr[2][BK][BN + PTN_GEMM_PAD];
  __shared__ T sBi[2][BK][BN + PTN_GEMM_PAD];

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;

  int64_t tm, tn;
  tile_coords(p, bid, tm, tn);
  const int64_t m0 = tm * BM, n0 = tn * BN;
  // split-K: this block covers k in [kbeg, kend)
  const int64_t kbeg = split * p.kchunk;
  const int64_t kend = (kbeg + p.kchunk < p.K) ? kbeg + p.kchunk : p.K;
  const TA *A = static_cast<const TA *>(p.A);
  const TB *B = static_cast<const TB *>(p.B);

  // ---- per-thread load coordinates ----------------------------------------
  // A: m-major -> (ml = tid%64, kl = tid/64 + 4i); k-major -> (kl = tid%BK, ml = tid/BK + (NT/BK)i)
  int64_t a_moff[LA];
  bool a_mval[LA];
  int a_kl[LA], a_ml[LA];
#pragma unroll
  for (int i = 0; i < LA; ++i) {
    if (p.akmaj) {
      a_kl[i] = tid % BK;
      a_ml[i] = tid / BK + (NT / BK) * i;
    } else {
      a_kl[i] = (tid >> 6) + 4 * i;
      a_ml[i] = tid & 63;
    }
    const int64_t gm = m0 + a_ml[i];
    a_mval[i] = gm < p.M;
    a_moff[i] = a_mval[i] ? p.aom[gm] : 0;
  }
  int64_t b_noff[LA];
  bool b_nval[LA];
  int b_kl[LA], b_nl[LA];
#pragma unroll
  for (int i = 0; i < LA; ++i) {
    if (p.bkmaj) {
      b_kl[i] = tid % BK;
      b_nl[i] = tid / BK + (NT / BK) * i;
    } else {
      b_kl[i] = (tid >> 6) + 4 * i;
      b_nl[i] = tid & 63;
    }
    const int64_t gn = n0 + b_nl[i];
    b_nval[i] = gn < p.N;
    b_noff[i] = b_nval[i] ? p.bon[gn] : 0;
  }

  // Software pipeline: the k-offsets of stage s+1 are fetched while stage s's
  // data loads are in flight, and the loaded values are only consumed (sign of
  // the conjugation applied) when they are written to LDS after the MFMA block.
  // Loads are branch-free: an element outside the matrix reads element 0 of the
  // operand and is replaced by zero.
  T ra_r[LA], ra_i[LA], rb_r[LA], rb_i[LA];
  int64_t a_koff[LA], b_koff[LA];
  const T asg = p.aconj ? T(-1) : T(1);
  const T bsg = p.bconj ? T(-1) : T(1);

  auto kload = [&](int64_t k0) {
#pragma unroll
    for (int i = 0; i < LA; ++i) {
      const int64_t gk = k0 + a_kl[i];
      a_koff[i] = (gk < kend) ? p.aok[gk] : -1;
    }
#pragma unroll
    for (int i = 0; i < LA; ++i) {
      const int64_t gk = k0 + b_kl[i];
      b_koff[i] = (gk < kend) ? p.bok[gk] : -1;
    }
  };
  auto gload = [&]() {
#pragma unroll
    for (int i = 0; i < LA; ++i) {
      const bool ok = a_mval[i] && a_koff[i] >= 0;
      T re, im;
      ptn::cload(A + (ok ? a_moff[i] + a_koff[i] : 0), re, im);
      ra_r[i] = ok ? re : T(0);
      ra_i[i] = ok ? im : T(0);
    }
#pragma unroll
    for (int i = 0; i < LA; ++i) {
      const bool ok = b_nval[i] && b_koff[i] >= 0;
      T re, im;
      ptn::cload(B + (ok ? b_koff[i] + b_noff[i] : 0), re, im);
      rb_r[i] = ok ? re : T(0);
      rb_i[i] = ok ? im : T(0);
    }
  };
  auto sstore = [&](int buf) {
#pragma unroll
    for (int i = 0; i < LA; ++i) {
      sAr[buf][a_kl[i]][a_ml[i]] = ra_r[i];
      sAi[buf][a_kl[i]][a_ml[i]] = asg * ra_i[i];
      sBr[buf][b_kl[i]][b_nl[i]] = rb_r[i];
      sBi[buf][b_kl[i]][b_nl[i]] = bsg * rb_i[i];
    }
  };

  V c0[2][2], c1[2][2], c2[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      c0[i][j] = V{0, 0, 0, 0};
      c1[i][j] = V{0, 0, 0, 0};
      c2[i][j] = V{0, 0, 0, 0};
    }

  const int64_t KT = kend > kbeg ? (kend - kbeg + BK - 1) / BK : 0;
  if (KT > 0) {
    kload(kbeg);
    gload();
    sstore(0);
    if (KT > 1) kload(kbeg + BK);
  }
  __syncthreads();
  const int fr = lane & 15, fk = lane >> 4;
  for (int64_t kt = 0; kt < KT; ++kt) {
    const int buf = kt & 1;
    if (kt + 1 < KT) {
      gload();                                   // data of stage kt+1
      if (kt + 2 < KT) kload(kbeg + (kt + 2) * BK);  // offsets of stage kt+2
    }
    // all MFMA operand fragments of this K step are read from LDS first, so the
    // reads of step kk+1 are in flight while step kk's MFMAs issue
    T ar[BK / 4][2], ai[BK / 4][2], br[BK / 4][2], bi[BK / 4][2];
#pragma unroll
    for (int kk = 0; kk < BK / 4; ++kk) {
      const int kr = kk * 4 + fk;
#pragma unroll
      for (int i = 0; i < 2; ++i) {
        ar[kk][i] = sAr[buf][kr][wm * 32 + i * 16 + fr];
        ai[kk][i] = sAi[buf][kr][wm * 32 + i * 16 + fr];
        br[kk][i] = sBr[buf][kr][wn * 32 + i * 16 + fr];
        bi[kk][i] = sBi[buf][kr][wn * 32 + i * 16 + fr];
      }
    }
#pragma unroll
    for (int kk = 0; kk < BK / 4; ++kk) {
      if constexpr (MODE == PTN_COMPUTE_F64_4M) {
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
          for (int j = 0; j < 2; ++j) {
            c0[i][j] = mfma(ar[kk][i], br[kk][j], c0[i][j]);
            c1[i][j] = mfma(ar[kk][i], bi[kk][j], c1[i][j]);
            c0[i][j] = mfma(-ai[kk][i], bi[kk][j], c0[i][j]);
            c1[i][j] = mfma(ai[kk][i], br[kk][j], c1[i][j]);
          }
      } else {
        T as[2], bs[2];
#pragma unroll
        for (int i = 0; i < 2; ++i) {
          as[i] = ar[kk][i] + ai[kk][i];
          bs[i] = br[kk][i] + bi[kk][i];
        }
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
          for (int j = 0; j < 2; ++j) {
            if constexpr (TR) {
              c0[i][j] = mfma(br[kk][j], ar[kk][i], c0[i][j]);
              c1[i][j] = mfma(bi[kk][j], ai[kk][i], c1[i][j]);
              c2[i][j] = mfma(bs[j], as[i], c2[i][j]);
            } else {
              c0[i][j] = mfma(ar[kk][i], br[kk][j], c0[i][j]);
              c1[i][j] = mfma(ai[kk][i], bi[kk][j], c1[i][j]);
              c2[i][j] = mfma(as[i], bs[j], c2[i][j]);
            }
          }
      }
    }
    if (kt + 1 < KT) sstore(buf ^ 1);
    __syncthreads();
  }

  // ---- epilogue -------------------------------------------------------------
  TC *C = static_cast<TC *>(p.C);
#pragma unroll
  for (int j = 0; j < 2; ++j) {
    const int64_t gn0 = n0 + wn * 32 + j * 16 + fr;
    if (!TR && gn0 >= p.N) continue;
    const int64_t noff0 = (!TR) ? p.con[gn0] : 0;
#pragma unroll
    for (int i = 0; i < 2; ++i) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        // accumulator row of (lane, r): f64 16x16x4 -> 4*r + lane/16,
        // f32 16x16x4 -> 4*(lane/16) + r; with TR the block is C^T: that row is
        // an n index and the column lane%16 an m index
        const int rr = (MODE == PTN_COMPUTE_F64_4M) ? (r * 4 + fk) : (fk * 4 + r);
        const int64_t gm = TR ? m0 + wm * 32 + i * 16 + fr : m0 + wm * 32 + i * 16 + rr;
        const int64_t gn = TR ? n0 + wn * 32 + j * 16 + rr : gn0;
        if (gm >= p.M || gn >= p.N) continue;
        const int64_t noff = TR ? p.con[gn] : noff0;
        double re, im;
        if constexpr (MODE == PTN_COMPUTE_F64_4M) {
          re = c0[i][j][r];
          im = c1[i][j][r];
        } else {
          const float t1 = c0[i][j][r], t2 = c1[i][j][r], t3 = c2[i][j][r];
          re = (double)(t1 - t2);
          im = (double)(t3 - t1 - t2);
        }
        if (splitk) {
          // raw partial sum; the reduction kernel applies alpha, beta and the scatter
          p.ws[(split * p.M + gm) * p.N + gn] = make_double2(re, im);
          continue;
        }
        double ore = p.are * re - p.aim * im;
        double oim = p.are * im + p.aim * re;
        TC *dst = C + p.com[gm] + noff;
        if (p.beta_one) {
          double xr, xi;
          ptn::cload(dst, xr, xi);
          ore += xr;
          oim += xi;
        }
        ptn::cstore(dst, ore, oim);
      }
    }
  }
}

// ---- BF16 (3M) register-staged kernel ---------------------------------------------
// The lowest precision of the relaxed Krylov form.  Operands (complex64 or
// complex128) are read through the offset tables as in gather_tile, rounded to
// bf16 (round to nearest even) while staging, and multiplied on
// v_mfma_f32_16x16x16_bf16 with FP32 accumulation in the three-product form.  The
// sums ar + ai and br + bi of the third product are formed in FP32 before
// rounding.  bf16 keeps FP32's exponent range, so unit-norm Krylov vectors and
// environments need no scaling.  LDS planes are [m][k] with k contiguous: lane l
// reads its four k values of a 16x16x16 operand (rows l%16, k = 4 (l/16) + 0..3)
// as one 8-byte load.  K step 16 (one MFMA slab per stage; a K step of 32
// spills the per-element offsets), two stages.
constexpr int HBK = 16, HKP = HBK + 4, HLA = BM * HBK / NT;
typedef short v4s __attribute__((ext_vector_type(4)));

__device__ __forceinline__ unsigned short bf16_rne(float x) {
  const unsigned u = __float_as_uint(x);
  return (unsigned short)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}

template <typename TA, typename TB, typename TC>
__device__ __forceinline__ void gather_tile_bf16(const KArgs &p, int64_t bid, int64_t split,
                                                 bool splitk) {
  // planes: 0 re, 1 im, 2 re + im
  __shared__ __attribute__((aligned(16))) unsigned short sA[2][3][BM][HKP];
  __shared__ __attribute__((aligned(16))) unsigned short sB[2][3][BN][HKP];
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  int64_t tm, tn;
  tile_coords(p, bid, tm, tn);
  const int64_t m0 = tm * BM, n0 = tn * BN;
  const int64_t kbeg = split * p.kchunk;
  const int64_t kend = (kbeg + p.kchunk < p.K) ? kbeg + p.kchunk : p.K;
  const TA *A = static_cast<const TA *>(p.A);
  const TB *B = static_cast<const TB *>(p.B);

  int64_t a_moff[HLA], b_noff[HLA];
  bool a_mval[HLA], b_nval[HLA];
  int a_kl[HLA], a_ml[HLA], b_kl[HLA], b_nl[HLA];
#pragma unroll
  for (int i = 0; i < HLA; ++i) {
    if (p.akmaj) {
      a_kl[i] = tid % HBK;
      a_ml[i] = tid / HBK + (NT / HBK) * i;
    } else {
      a_kl[i] = (tid >> 6) + 4 * i;
      a_ml[i] = tid & 63;
    }
    const int64_t gm = m0 + a_ml[i];
    a_mval[i] = gm < p.M;
    a_moff[i] = a_mval[i] ? p.aom[gm] : 0;
    if (p.bkmaj) {
      b_kl[i] = tid % HBK;
      b_nl[i] = tid / HBK + (NT / HBK) * i;
    } else {
      b_kl[i] = (tid >> 6) + 4 * i;
      b_nl[i] = tid & 63;
    }
    const int64_t gn = n0 + b_nl[i];
    b_nval[i] = gn < p.N;
    b_noff[i] = b_nval[i] ? p.bon[gn] : 0;
  }

  float ra_r[HLA], ra_i[HLA], rb_r[HLA], rb_i[HLA];
  int64_t a_koff[HLA], b_koff[HLA];
  const float asg = p.aconj ? -1.0f : 1.0f, bsg = p.bconj ? -1.0f : 1.0f;
  auto kload = [&](int64_t k0) {
#pragma unroll
    for (int i = 0; i < HLA; ++i) {
      const int64_t ga = k0 + a_kl[i];
      a_koff[i] = (ga < kend) ? p.aok[ga] : -1;
      const int64_t gb = k0 + b_kl[i];
      b_koff[i] = (gb < kend) ? p.bok[gb] : -1;
    }
  };
  auto gload = [&]() {
#pragma unroll
    for (int i = 0; i < HLA; ++i) {
      const bool ok = a_mval[i] && a_koff[i] >= 0;
      float re, im;
      ptn::cload(A + (ok ? a_moff[i] + a_koff[i] : 0), re, im);
      ra_r[i] = ok ? re : 0.0f;
      ra_i[i] = ok ? im : 0.0f;
    }
#pragma unroll
    for (int i = 0; i < HLA; ++i) {
      const bool ok = b_nval[i] && b_koff[i] >= 0;
      float re, im;
      ptn::cload(B + (ok ? b_koff[i] + b_noff[i] : 0), re, im);
      rb_r[i] = ok ? re : 0.0f;
      rb_i[i] = ok ? im : 0.0f;
    }
  };
  auto sstore = [&](int buf) {
#pragma unroll
    for (int i = 0; i < HLA; ++i) {
      const float ai = asg * ra_i[i], bi = bsg * rb_i[i];
      sA[buf][0][a_ml[i]][a_kl[i]] = bf16_rne(ra_r[i]);
      sA[buf][1][a_ml[i]][a_kl[i]] = bf16_rne(ai);
      sA[buf][2][a_ml[i]][a_kl[i]] = bf16_rne(ra_r[i] + ai);
      sB[buf][0][b_nl[i]][b_kl[i]] = bf16_rne(rb_r[i]);
      sB[buf][1][b_nl[i]][b_kl[i]] = bf16_rne(bi);
      sB[buf][2][b_nl[i]][b_kl[i]] = bf16_rne(rb_r[i] + bi);
    }
  };

  v4f c0[2][2], c1[2][2], c2[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      c0[i][j] = v4f{0, 0, 0, 0};
      c1[i][j] = v4f{0, 0, 0, 0};
      c2[i][j] = v4f{0, 0, 0, 0};
    }

  const int64_t KT = kend > kbeg ? (kend - kbeg + HBK - 1) / HBK : 0;
  if (KT > 0) {
    kload(kbeg);
    gload();
    sstore(0);
    if (KT > 1) kload(kbeg + HBK);
  }
  __syncthreads();
  const int fr = lane & 15, fk = lane >> 4;
  for (int64_t kt = 0; kt < KT; ++kt) {
    const int buf = kt & 1;
    if (kt + 1 < KT) {
      gload();
      if (kt + 2 < KT) kload(kbeg + (kt + 2) * HBK);
    }
#pragma unroll
    for (int sl = 0; sl < HBK / 16; ++sl) {
      v4s a[3][2], b[3][2];
#pragma unroll
      for (int q = 0; q < 3; ++q)
#pragma unroll
        for (int i = 0; i < 2; ++i) {
          a[q][i] = *reinterpret_cast<const v4s *>(
              &sA[buf][q][wm * 32 + i * 16 + fr][sl * 16 + fk * 4]);
          b[q][i] = *reinterpret_cast<const v4s *>(
              &sB[buf][q][wn * 32 + i * 16 + fr][sl * 16 + fk * 4]);
        }
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
          c0[i][j] =
              __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[0][i], b[0][j], c0[i][j], 0, 0, 0);
          c1[i][j] =
              __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[1][i], b[1][j], c1[i][j], 0, 0, 0);
          c2[i][j] =
              __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[2][i], b[2][j], c2[i][j], 0, 0, 0);
        }
    }
    if (kt + 1 < KT) sstore(buf ^ 1);
    __syncthreads();
  }

  // ---- epilogue (accumulator row of (lane, r) = 4*(lane/16) + r) -------------
  TC *C = static_cast<TC *>(p.C);
#pragma unroll
  for (int j = 0; j < 2; ++j) {
    const int64_t gn = n0 + wn * 32 + j * 16 + fr;
    if (gn >= p.N) continue;
    const int64_t noff = p.con[gn];
#pragma unroll
    for (int i = 0; i < 2; ++i) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t gm = m0 + wm * 32 + i * 16 + fk * 4 + r;
        if (gm >= p.M) continue;
        const float t1 = c0[i][j][r], t2 = c1[i][j][r], t3 = c2[i][j][r];
        const double re = (double)(t1 - t2), im = (double)(t3 - t1 - t2);
        if (splitk) {
          p.ws[(split * p.M + gm) * p.N + gn] = make_double2(re, im);
          continue;
        }
        double ore = p.are * re - p.aim * im;
        double oim = p.are * im + p.aim * re;
        TC *dst = C + p.com[gm] + noff;
        if (p.beta_one) {
          double xr, xi;
          ptn::cload(dst, xr, xi);
          ore += xr;
          oim += xi;
        }
        ptn::cstore(dst, ore, oim);
      }
    }
  }
}

// tile index of a launch with the XCD-aware remap applied
__device__ __forceinline__ int64_t remapped_tile(const KArgs &p) {
  int64_t bid = blockIdx.x;
  if (p.remap) {
    const int64_t per = gridDim.x >> 3;
    bid = (bid & 7) * per + (bid >> 3);
  }
  return bid;
}

template <int MODE, typename TA, typename TB, typename TC, int LARGE, bool TR = false>
__global__ __launch_bounds__(NT, PTN_GEMM_OCC) void gemm_gather_kernel(KArgs p) {
  gather_tile<MODE, TA, TB, TC, TR>(p, remapped_tile(p), blockIdx.y, gridDim.y > 1);
}

template <typename TA, typename TB, typename TC, int LARGE>
__global__ __launch_bounds__(NT, 2) void gemm_bf16_kernel(KArgs p) {
  gather_tile_bf16<TA, TB, TC>(p, remapped_tile(p), blockIdx.y, gridDim.y > 1);
}

// ---- BF16X6_3M: FP32-accurate 3M products on the bf16 MFMA -------------------------
// Every FP32 operand value x is split into three bf16 pieces x = h + m + l (+ a
// remainder below 2^-24 |x|): h = bf16(x), m = bf16(x - h), l = bf16(x - h - m),
// each difference exact in FP32.  A real product a b is accumulated as the six
// piece products of order >= 2^-16: hh + hm + mh + mm + hl + lh, with FP32
// accumulation (v_mfma_f32_16x16x32_bf16, 8 bf16 pairs per lane per MFMA), so
// the GEMM carries the error of an FP32 GEMM (Henry, Tang and Heinecke, ARITH
// 2019, "bf16x6").  The complex product uses Gauss's three real products
// (T1 = ar br, T2 = ai bi, T3 = (ar + ai)(br + bi); the sums formed in FP32
// before splitting): 18 MFMAs per 16 x 16 block product and K step 32.
// Operands are gathered through the offset tables into registers (two
// consecutive k per element pair), split, and stored as nine bf16 LDS planes
// [re, im, re + im][h, m, l][row][k] per operand; 16-byte k chunks of a row are
// XOR-swizzled by row bits 2-3, so the MFMA operand reads of 16 rows
// (lane l: row l % 16, k = 8 (l / 16) + 0..7) hit distinct banks.  One LDS
// stage (72 KB, two workgroups per CU) with the next stage's values prefetched
// into registers during the MFMAs.
constexpr int XBK = 32;
typedef __bf16 v8bf16 __attribute__((ext_vector_type(8)));

typedef __bf16 v2bf16 __attribute__((ext_vector_type(2)));

// the three-piece splits of two values x0, x1 (consecutive k), each piece pair
// packed as one 32-bit word (x0's piece in the low half): one v_cvt_pk_bf16_f32
// (round to nearest even) per piece pair
__device__ __forceinline__ void split3_pair(float x0, float x1, unsigned &h, unsigned &m,
                                            unsigned &l) {
  const v2bf16 hv = {(__bf16)x0, (__bf16)x1};
  h = __builtin_bit_cast(unsigned, hv);
  float r0 = x0 - __uint_as_float(h << 16), r1 = x1 - __uint_as_float(h & 0xFFFF0000u);
  const v2bf16 mv = {(__bf16)r0, (__bf16)r1};
  m = __builtin_bit_cast(unsigned, mv);
  r0 -= __uint_as_float(m << 16);
  r1 -= __uint_as_float(m & 0xFFFF0000u);
  const v2bf16 lv = {(__bf16)r0, (__bf16)r1};
  l = __builtin_bit_cast(unsigned, lv);
}

__device__ __forceinline__ int xswz(int row, int k) {  // element index within a row
  return ((((k >> 3) ^ ((row >> 2) & 3))) << 3) | (k & 7);
}

template <typename TA, typename TB, typename TC>
__device__ __forceinline__ void gather_tile_bf16x6(const KArgs &p, int64_t bid, int64_t split,
                                                   bool splitk) {
  // [plane re/im/sum][piece h/m/l][row][k], as 32-bit words of two k
  __shared__ __attribute__((aligned(16))) unsigned sA[3][3][BM][XBK / 2];
  __shared__ __attribute__((aligned(16))) unsigned sB[3][3][BN][XBK / 2];
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  int64_t tm, tn;
  tile_coords(p, bid, tm, tn);
  const int64_t m0 = tm * BM, n0 = tn * BN;
  const int64_t kbeg = split * p.kchunk;
  const int64_t kend = (kbeg + p.kchunk < p.K) ? kbeg + p.kchunk : p.K;
  const TA *A = static_cast<const TA *>(p.A);
  const TB *B = static_cast<const TB *>(p.B);
  // each thread stages four (row, k pair) slots of A and of B: 4 x 2 = 8 elements.
  // k-major operand: the pair k = 2 (tid % 16) + {0,1}, rows tid / 16 + 16 i;
  // m-major operand: row tid % 64, pairs k = 8 i + 2 (tid / 64) + {0,1}
  int a_row[4], a_k[4], b_row[4], b_k[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    if (p.akmaj) {
      a_k[i] = 2 * (tid & 15);
      a_row[i] = (tid >> 4) + 16 * i;
    } else {
      a_row[i] = tid & 63;
      a_k[i] = 8 * i + 2 * (tid >> 6);
    }
    if (p.bkmaj) {
      b_k[i] = 2 * (tid & 15);
      b_row[i] = (tid >> 4) + 16 * i;
    } else {
      b_row[i] = tid & 63;
      b_k[i] = 8 * i + 2 * (tid >> 6);
    }
  }
  int64_t a_moff[4], b_noff[4];
  bool a_mval[4], b_nval[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int64_t gm = m0 + a_row[i];
    a_mval[i] = gm < p.M;
    a_moff[i] = a_mval[i] ? p.aom[gm] : 0;
    const int64_t gn = n0 + b_row[i];
    b_nval[i] = gn < p.N;
    b_noff[i] = b_nval[i] ? p.bon[gn] : 0;
  }
  float ra[4][2][2], rb[4][2][2];  // [slot][k of the pair][re, im]
  const float asg = p.aconj ? -1.0f : 1.0f, bsg = p.bconj ? -1.0f : 1.0f;
  auto gload = [&](int64_t k0) {
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int e = 0; e < 2; ++e) {
        const int64_t ka = k0 + a_k[i] + e;
        const bool oka = a_mval[i] && ka < kend;
        const int64_t offa = oka ? a_moff[i] + p.aok[ka] : 0;
        float re, im;
        ptn::cload(A + offa, re, im);
        ra[i][e][0] = oka ? re : 0.0f;
        ra[i][e][1] = oka ? asg * im : 0.0f;
        const int64_t kb = k0 + b_k[i] + e;
        const bool okb = b_nval[i] && kb < kend;
        const int64_t offb = okb ? p.bok[kb] + b_noff[i] : 0;
        ptn::cload(B + offb, re, im);
        rb[i][e][0] = okb ? re : 0.0f;
        rb[i][e][1] = okb ? bsg * im : 0.0f;
      }
  };
  auto sstore = [&]() {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      unsigned h, m, l;
      const int w = xswz(a_row[i], a_k[i]) >> 1;
#pragma unroll
      for (int q = 0; q < 3; ++q) {
        const float x0 = q < 2 ? ra[i][0][q] : ra[i][0][0] + ra[i][0][1];
        const float x1 = q < 2 ? ra[i][1][q] : ra[i][1][0] + ra[i][1][1];
        split3_pair(x0, x1, h, m, l);
        sA[q][0][a_row[i]][w] = h;
        sA[q][1][a_row[i]][w] = m;
        sA[q][2][a_row[i]][w] = l;
      }
      const int wb = xswz(b_row[i], b_k[i]) >> 1;
#pragma unroll
      for (int q = 0; q < 3; ++q) {
        const float x0 = q < 2 ? rb[i][0][q] : rb[i][0][0] + rb[i][0][1];
        const float x1 = q < 2 ? rb[i][1][q] : rb[i][1][0] + rb[i][1][1];
        split3_pair(x0, x1, h, m, l);
        sB[q][0][b_row[i]][wb] = h;
        sB[q][1][b_row[i]][wb] = m;
        sB[q][2][b_row[i]][wb] = l;
      }
    }
  };

  v4f acc[3][2][2];
#pragma unroll
  for (int q = 0; q < 3; ++q)
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) acc[q][i][j] = v4f{0, 0, 0, 0};

  const int64_t KT = kend > kbeg ? (kend - kbeg + XBK - 1) / XBK : 0;
  const int fr = lane & 15, fk = lane >> 4;  // operand row, 8-k chunk
  if (KT > 0) gload(kbeg);
  for (int64_t kt = 0; kt < KT; ++kt) {
    sstore();
    __syncthreads();
    if (kt + 1 < KT) gload(kbeg + (kt + 1) * XBK);  // in flight during the MFMAs
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      v8bf16 a[3][2], b[3][2];
#pragma unroll
      for (int pc = 0; pc < 3; ++pc)
#pragma unroll
        for (int i = 0; i < 2; ++i) {
          const int ra_ = wm * 32 + i * 16 + fr, rb_ = wn * 32 + i * 16 + fr;
          a[pc][i] = *reinterpret_cast<const v8bf16 *>(&sA[q][pc][ra_][xswz(ra_, fk * 8) >> 1]);
          b[pc][i] = *reinterpret_cast<const v8bf16 *>(&sB[q][pc][rb_][xswz(rb_, fk * 8) >> 1]);
        }
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
          v4f c = acc[q][i][j];
          // smallest pieces first
          c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[2][i], b[0][j], c, 0, 0, 0);
          c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0][i], b[2][j], c, 0, 0, 0);
          c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[1][i], b[1][j], c, 0, 0, 0);
          c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[1][i], b[0][j], c, 0, 0, 0);
          c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0][i], b[1][j], c, 0, 0, 0);
          c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0][i], b[0][j], c, 0, 0, 0);
          acc[q][i][j] = c;
        }
    }
    __syncthreads();
  }

  // ---- epilogue (accumulator row of (lane, r) = 4 (lane / 16) + r) -------------------
  TC *C = static_cast<TC *>(p.C);
#pragma unroll
  for (int j = 0; j < 2; ++j) {
    const int64_t gn = n0 + wn * 32 + j * 16 + fr;
    if (gn >= p.N) continue;
    const int64_t noff = p.con[gn];
#pragma unroll
    for (int i = 0; i < 2; ++i) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t gm = m0 + wm * 32 + i * 16 + fk * 4 + r;
        if (gm >= p.M) continue;
        const double t1 = acc[0][i][j][r], t2 = acc[1][i][j][r], t3 = acc[2][i][j][r];
        const double re = t1 - t2, im = t3 - t1 - t2;
        if (splitk) {
          p.ws[(split * p.M + gm) * p.N + gn] = make_double2(re, im);
          continue;
        }
        double ore = p.are * re - p.aim * im;
        double oim = p.are * im + p.aim * re;
        TC *dst = C + p.com[gm] + noff;
        if (p.beta_one) {
          double xr, xi;
          ptn::cload(dst, xr, xi);
          ore += xr;
          oim += xi;
        }
        ptn::cstore(dst, ore, oim);
      }
    }
  }
}

template <typename TA, typename TB, typename TC, int LARGE>
__global__ __launch_bounds__(NT, 2) void gemm_bf16x6_kernel(KArgs p) {
  gather_tile_bf16x6<TA, TB, TC>(p, remapped_tile(p), blockIdx.y, gridDim.y > 1);
}

// ---- FP64 kernel with direct global->LDS tile loads ---------------------------
// Each wave fills whole LDS rows of a tile with global_load_lds_dwordx4: lane l
// copies element (k, m0 + l) straight into LDS, no register staging.  With m
// (n) contiguous in memory a row is one coalesced 1 KB read; with k contiguous
// the eight rows of a tile read the same cache lines in turn.  Three LDS stages form a ring: the
// loads of stage kt+2 are issued right after the barrier that retires stage
// kt-1, so each load has two K steps of MFMA work to arrive.  The k-offsets of
// a tile row are wave-uniform (scalar loads).  Elements outside the matrix or
// past the K range are read from a zero page.
constexpr int DBK = 8, DST = 3;
#ifndef PTN_F64_3M
#define PTN_F64_3M 1
#endif
#ifndef PTN_DIRECT_OCC
#define PTN_DIRECT_OCC 2
#endif
__device__ double2 g_zero_page[64];

__device__ __forceinline__ void lds_dma16(const double2 *src, double2 *lds_row) {
  __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void *)src,
                                   (__attribute__((address_space(3))) void *)lds_row, 16, 0, 0);
}

// TR computes the transposed block product (MFMA operands swapped, same
// registers): the accumulators then run along m, so the epilogue stores are
// contiguous when C is m-contiguous.
template <bool TR>
__device__ __forceinline__ void direct_tile(const KArgs &p, const int64_t *__restrict__ aok,
                                            const int64_t *__restrict__ bok, int64_t bid,
                                            int64_t split, bool splitk) {
  // one LDS array per ring stage, so every LDS read names the stage it reads
  // and the compiler need not wait for the DMA of the other stages
  __shared__ double2 sA0[DBK][BM], sA1[DBK][BM], sA2[DBK][BM];
  __shared__ double2 sB0[DBK][BN], sB1[DBK][BN], sB2[DBK][BN];
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int wm = wave >> 1, wn = wave & 1;
  int64_t tm, tn;
  tile_coords(p, bid, tm, tn);
  const int64_t m0 = tm * BM, n0 = tn * BN;
  const int64_t kbeg = split * p.kchunk;
  const int64_t kend = (kbeg + p.kchunk < p.K) ? kbeg + p.kchunk : p.K;
  const double2 *A = static_cast<const double2 *>(p.A);
  const double2 *B = static_cast<const double2 *>(p.B);
  const int64_t gm = m0 + lane, gn = n0 + lane;
  const bool mval = gm < p.M, nval = gn < p.N;
  const int64_t moff = mval ? p.aom[gm] : 0;
  const int64_t noff = nval ? p.bon[gn] : 0;
  const double2 *zero = g_zero_page + lane;
  const int64_t klast = kend > kbeg ? kend - 1 : kbeg;

  // rows wave and wave + 4 of the A and B tiles of stage kt into (sa, sb)
  auto issue = [&](int64_t kt, double2 (*sa)[BM], double2 (*sb)[BN]) {
#pragma unroll
    for (int r = 0; r < 2; ++r) {
      const int kl = wave + 4 * r;
      const int64_t gk = kbeg + kt * DBK + kl;
      const bool kval = gk < kend;
      const int64_t gkc = kval ? gk : klast;
      const int64_t ka = aok[gkc];
      const int64_t kb = bok[gkc];
      lds_dma16((kval && mval) ? A + moff + ka : zero, &sa[kl][0]);
      lds_dma16((kval && nval) ? B + kb + noff : zero, &sb[kl][0]);
    }
  };

  typedef double v4d_t __attribute__((ext_vector_type(4)));
  // PTN_F64_3M: Gauss's three-product form, T1 = sum ar br, T2 = sum ai bi,
  // T3 = sum (ar + ai)(br + bi); re = T1 - T2, im = T3 - T1 - T2 (3 MFMAs per
  // block product instead of 4; normwise error of the same order as the
  // four-product form).  Otherwise re and im accumulate directly (4 MFMAs).
  v4d_t c0[2][2], c1[2][2];
#if PTN_F64_3M
  v4d_t c2[2][2];
#endif
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      c0[i][j] = v4d_t{0, 0, 0, 0};
      c1[i][j] = v4d_t{0, 0, 0, 0};
#if PTN_F64_3M
      c2[i][j] = v4d_t{0, 0, 0, 0};
#endif
    }
  const double asg = p.aconj ? -1.0 : 1.0, bsg = p.bconj ? -1.0 : 1.0;
  const int fr = lane & 15, fk = lane >> 4;

  // x supplies the A side of the block product, y the B side
  auto mf = [](double x, double y, v4d_t c) {
    if constexpr (TR)
      return __builtin_amdgcn_mfma_f64_16x16x4f64(y, x, c, 0, 0, 0);
    else
      return __builtin_amdgcn_mfma_f64_16x16x4f64(x, y, c, 0, 0, 0);
  };
  auto compute = [&](double2 (*sa)[BM], double2 (*sb)[BN]) {
#pragma unroll
    for (int kk = 0; kk < DBK / 4; ++kk) {
      const int kr = kk * 4 + fk;
      double ar[2], ai[2], br[2], bi[2];
#pragma unroll
      for (int i = 0; i < 2; ++i) {
        const double2 a = sa[kr][wm * 32 + i * 16 + fr];
        const double2 b = sb[kr][wn * 32 + i * 16 + fr];
        ar[i] = a.x;
        ai[i] = asg * a.y;
        br[i] = b.x;
        bi[i] = bsg * b.y;
      }
#if PTN_F64_3M
      double as[2], bs[2];
#pragma unroll
      for (int i = 0; i < 2; ++i) {
        as[i] = ar[i] + ai[i];
        bs[i] = br[i] + bi[i];
      }
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
          c0[i][j] = mf(ar[i], br[j], c0[i][j]);
          c1[i][j] = mf(ai[i], bi[j], c1[i][j]);
          c2[i][j] = mf(as[i], bs[j], c2[i][j]);
        }
#else
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
          c0[i][j] = mf(ar[i], br[j], c0[i][j]);
          c1[i][j] = mf(ar[i], bi[j], c1[i][j]);
          c0[i][j] = mf(-ai[i], bi[j], c0[i][j]);
          c1[i][j] = mf(ai[i], br[j], c1[i][j]);
        }
#endif
    }
  };

  // one K step: wait for this wave's loads of stage kt (the 4 loads of stage
  // kt+1 may still fly), barrier (all rows of stage kt visible, stage kt-1
  // retired), refill the retired buffer with stage kt+2, compute stage kt
  const int64_t KT = kend > kbeg ? (kend - kbeg + DBK - 1) / DBK : 0;
  auto step = [&](int64_t kt, double2 (*ca)[BM], double2 (*cb)[BN], double2 (*na)[BM],
                  double2 (*nb)[BN]) {
    if (kt + 1 < KT)
      __builtin_amdgcn_s_waitcnt(0x0F74);  // vmcnt(4)
    else
      __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0)
    // a plain s_barrier: __syncthreads() would add a workgroup fence, which
    // waits for every load in flight (vmcnt(0)) including stage kt+1.  Each
    // wave has waited for its own stage-kt rows above and its LDS reads of
    // stage kt-1 completed before their MFMAs; the empty asm statements keep
    // the compiler from moving LDS accesses across the barrier.
    asm volatile("" ::: "memory");
    __builtin_amdgcn_s_barrier();
    asm volatile("" ::: "memory");
    if (kt + 2 < KT) issue(kt + 2, na, nb);
    compute(ca, cb);
  };
  if (KT > 0) issue(0, sA0, sB0);
  if (KT > 1) issue(1, sA1, sB1);
  // whole ring turns first, with no exit inside the unrolled body (an exit
  // between stages makes the compiler copy the accumulators at every stage and
  // wait for the MFMA pipeline to drain), then the 0-2 remaining stages
  int64_t kt = 0;
  for (; kt + 3 <= KT; kt += 3) {
    step(kt, sA0, sB0, sA2, sB2);
    step(kt + 1, sA1, sB1, sA0, sB0);
    step(kt + 2, sA2, sB2, sA1, sB1);
  }
  if (kt < KT) step(kt, sA0, sB0, sA2, sB2);
  if (kt + 1 < KT) step(kt + 1, sA1, sB1, sA0, sB0);

  // ---- epilogue (accumulator row of (lane, r) = 4*r + lane/16) ----------------
  double2 *C = static_cast<double2 *>(p.C);
#pragma unroll
  for (int j = 0; j < 2; ++j) {
#pragma unroll
    for (int i = 0; i < 2; ++i) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        // accumulator register r of lane l: row 4 r + l/16, column l%16 of the
        // block of C (of C^T when TR)
        const int64_t cm = TR ? m0 + wm * 32 + i * 16 + fr : m0 + wm * 32 + i * 16 + r * 4 + fk;
        const int64_t cn = TR ? n0 + wn * 32 + j * 16 + r * 4 + fk : n0 + wn * 32 + j * 16 + fr;
        if (cm >= p.M || cn >= p.N) continue;
#if PTN_F64_3M
        const double t1 = c0[i][j][r], t2 = c1[i][j][r];
        const double re = t1 - t2, im = c2[i][j][r] - t1 - t2;
#else
        const double re = c0[i][j][r], im = c1[i][j][r];
#endif
        if (splitk) {
          p.ws[(split * p.M + cm) * p.N + cn] = make_double2(re, im);
          continue;
        }
        double ore = p.are * re - p.aim * im;
        double oim = p.are * im + p.aim * re;
        double2 *dst = C + p.com[cm] + p.con[cn];
        if (p.beta_one) {
          const double2 x = *dst;
          ore += x.x;
          oim += x.y;
        }
        *dst = make_double2(ore, oim);
      }
    }
  }
}

template <int LARGE, bool TR>
__global__ __launch_bounds__(NT, PTN_DIRECT_OCC) void gemm_direct_f64_kernel(
    KArgs p, const int64_t *__restrict__ aok, const int64_t *__restrict__ bok) {
  direct_tile<TR>(p, aok, bok, remapped_tile(p), blockIdx.y, gridDim.y > 1);
}

// ---- grouped launches -----------------------------------------------------------
// One launch runs the tiles of `count` independent contractions (one per batch
// instance): workgroup b finds its descriptor d with prefix[d] <= b <
// prefix[d+1] (binary search over scalar loads) and runs tile b - prefix[d] of
// it.  No split-K and no XCD remap inside a grouped launch.
__device__ __forceinline__ int find_desc(const int64_t *__restrict__ prefix, int count,
                                         int64_t b) {
  int lo = 0, hi = count;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (prefix[mid] <= b)
      lo = mid;
    else
      hi = mid;
  }
  return lo;
}

__global__ __launch_bounds__(NT, PTN_DIRECT_OCC) void gemm_direct_f64_grouped(const KArgs *__restrict__ descs,
                                                                 const int64_t *__restrict__ prefix,
                                                                 int count) {
  const int64_t b = blockIdx.x;
  const int d = find_desc(prefix, count, b);
  const KArgs p = descs[d];
  direct_tile<false>(p, p.aok, p.bok, b - prefix[d], 0, false);
}

template <int MODE, typename TA, typename TB, typename TC>
__global__ __launch_bounds__(NT, PTN_GEMM_OCC) void gemm_gather_grouped(
    const KArgs *__restrict__ descs, const int64_t *__restrict__ prefix, int count) {
  const int64_t b = blockIdx.x;
  const int d = find_desc(prefix, count, b);
  const KArgs p = descs[d];
  gather_tile<MODE, TA, TB, TC>(p, b - prefix[d], 0, false);
}

// sum the split-K partials in split order, then alpha / beta / scatter store
template <typename TC>
__global__ __launch_bounds__(256) void splitk_reduce(KArgs p, int ksplit) {
  const int64_t total = p.M * p.N;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  TC *C = static_cast<TC *>(p.C);
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += stride) {
    double re = 0, im = 0;
    for (int s = 0; s < ksplit; ++s) {
      const double2 v = p.ws[(int64_t)s * total + t];
      re += v.x;
      im += v.y;
    }
    const int64_t m = t / p.N, n = t - m * p.N;
    if (p.lower && m / BM < n / BN) continue;   // a skipped tile: no partials
    double ore = p.are * re - p.aim * im;
    double oim = p.are * im + p.aim * re;
    TC *dst = C + p.com[m] + p.con[n];
    if (p.beta_one) {
      double xr, xi;
      ptn::cload(dst, xr, xi);
      ore += xr;
      oim += xi;
    }
    ptn::cstore(dst, ore, oim);
  }
}

// the same sum for few outputs and many splits: one wave per output, lane l
// adds splits l, l + 64, ... in order, then a fixed shuffle tree
template <typename TC>
__global__ __launch_bounds__(256) void splitk_reduce_wide(KArgs p, int ksplit) {
  const int64_t total = p.M * p.N;
  const int lane = threadIdx.x & 63;
  const int64_t t = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (t >= total) return;
  double re = 0, im = 0;
  for (int s = lane; s < ksplit; s += 64) {
    const double2 v = p.ws[(int64_t)s * total + t];
    re += v.x;
    im += v.y;
  }
  for (int off = 32; off > 0; off >>= 1) {
    re += __shfl_down(re, off, 64);
    im += __shfl_down(im, off, 64);
  }
  if (lane != 0) return;
  const int64_t m = t / p.N, n = t - m * p.N;
  double ore = p.are * re - p.aim * im;
  double oim = p.are * im + p.aim * re;
  TC *dst = static_cast<TC *>(p.C) + p.com[m] + p.con[n];
  if (p.beta_one) {
    double xr, xi;
    ptn::cload(dst, xr, xi);
    ore += xr;
    oim += xi;
  }
  ptn::cstore(dst, ore, oim);
}

// ---- wave-level FP64 GEMM for small and thin products -------------------------
// Each wave computes one (16 BMB) x (16 BNB) block of C over one K range with
// the operands loaded straight into MFMA registers: for each 4-wide k slab,
// lane l supplies A[m0 + 16 i + l%16][k + l/16] and B[k + l/16][n0 + 16 j + l%16].
// No LDS and no barriers, so a wave's cost is its loads and MFMAs alone; the
// 64 x 64 workgroup tile would spend 16x the MFMA work on a 16 x 16 product
// and a full prologue / epilogue on a K of 16.  Waves are numbered tile-major
// (the splits of one tile are consecutive).  ksplit == 1: the wave writes C;
// otherwise raw partial sums go to ws[split][M][N] for the reduction kernel.
// The k-offsets of the next group of WU slabs are fetched while the current
// group's data loads are in flight.  TR computes the transposed product
// (MFMA operands swapped, same registers): the accumulators then hold C with
// lanes along m, which makes the stores contiguous when C is m-contiguous.
constexpr int WU = 4;
// 16-blocks per wave along the long side of a thin product (1, 2 or 4): one
// block keeps the most waves in flight (18 vs 28 us at M=131072, N=K=16)
#ifndef PTN_WAVE_THIN
#define PTN_WAVE_THIN 1
#endif

template <int BMB, int BNB, bool TR, typename TA, typename TB, typename TC>
__global__ __launch_bounds__(256) void wave_gemm_f64(KArgs p, int64_t wtiles_n, int64_t nwaves,
                                                     int ksplit) {
  const int lane = threadIdx.x & 63;
  const int64_t gw = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (gw >= nwaves) return;
  const int64_t tile = gw / ksplit, split = gw - tile * ksplit;
  const int64_t tm = tile / wtiles_n, tn = tile - tm * wtiles_n;
  const int64_t m0 = tm * 16 * BMB, n0 = tn * 16 * BNB;
  const int64_t kbeg = split * p.kchunk;
  const int64_t kend = (kbeg + p.kchunk < p.K) ? kbeg + p.kchunk : p.K;
  const int fr = lane & 15, fk = lane >> 4;
  const TA *A = static_cast<const TA *>(p.A);
  const TB *B = static_cast<const TB *>(p.B);
  int64_t aoff[BMB], boff[BNB];
  bool aval[BMB], bval[BNB];
#pragma unroll
  for (int i = 0; i < BMB; ++i) {
    const int64_t gm = m0 + 16 * i + fr;
    aval[i] = gm < p.M;
    aoff[i] = aval[i] ? p.aom[gm] : 0;
  }
#pragma unroll
  for (int j = 0; j < BNB; ++j) {
    const int64_t gn = n0 + 16 * j + fr;
    bval[j] = gn < p.N;
    boff[j] = bval[j] ? p.bon[gn] : 0;
  }
  typedef double v4d_t __attribute__((ext_vector_type(4)));
  v4d_t c0[BMB][BNB], c1[BMB][BNB];
#pragma unroll
  for (int i = 0; i < BMB; ++i)
#pragma unroll
    for (int j = 0; j < BNB; ++j) {
      c0[i][j] = v4d_t{0, 0, 0, 0};
      c1[i][j] = v4d_t{0, 0, 0, 0};
    }
  const double asg = p.aconj ? -1.0 : 1.0, bsg = p.bconj ? -1.0 : 1.0;
  int64_t ka[WU], kb[WU];
  auto koffs = [&](int64_t k) {
#pragma unroll
    for (int u = 0; u < WU; ++u) {
      const int64_t gk = k + 4 * u + fk;
      ka[u] = gk < kend ? p.aok[gk] : -1;
      kb[u] = gk < kend ? p.bok[gk] : -1;
    }
  };
  if (kbeg < kend) koffs(kbeg);
  for (int64_t k = kbeg; k < kend; k += 4 * WU) {
    double ar[WU][BMB], ai[WU][BMB], br[WU][BNB], bi[WU][BNB];
#pragma unroll
    for (int u = 0; u < WU; ++u) {
#pragma unroll
      for (int i = 0; i < BMB; ++i) {
        const bool ok = aval[i] && ka[u] >= 0;
        double vr, vi;
        ptn::cload(A + (ok ? aoff[i] + ka[u] : 0), vr, vi);
        ar[u][i] = ok ? vr : 0.0;
        ai[u][i] = ok ? asg * vi : 0.0;
      }
#pragma unroll
      for (int j = 0; j < BNB; ++j) {
        const bool ok = bval[j] && kb[u] >= 0;
        double vr, vi;
        ptn::cload(B + (ok ? kb[u] + boff[j] : 0), vr, vi);
        br[u][j] = ok ? vr : 0.0;
        bi[u][j] = ok ? bsg * vi : 0.0;
      }
    }
    if (k + 4 * WU < kend) koffs(k + 4 * WU);
#pragma unroll
    for (int u = 0; u < WU; ++u) {
      if (k + 4 * u >= kend) continue;  // slabs past the range (wave-uniform)
#pragma unroll
      for (int i = 0; i < BMB; ++i)
#pragma unroll
        for (int j = 0; j < BNB; ++j) {
          if constexpr (TR) {
            c0[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(br[u][j], ar[u][i], c0[i][j], 0, 0, 0);
            c1[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(bi[u][j], ar[u][i], c1[i][j], 0, 0, 0);
            c0[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(-bi[u][j], ai[u][i], c0[i][j], 0, 0, 0);
            c1[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(br[u][j], ai[u][i], c1[i][j], 0, 0, 0);
          } else {
            c0[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[u][i], br[u][j], c0[i][j], 0, 0, 0);
            c1[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[u][i], bi[u][j], c1[i][j], 0, 0, 0);
            c0[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(-ai[u][i], bi[u][j], c0[i][j], 0, 0, 0);
            c1[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(ai[u][i], br[u][j], c1[i][j], 0, 0, 0);
          }
        }
    }
  }
  // accumulator register r of lane l: row 4 r + l/16, column l%16 of the
  // block (of C, or of C^T when TR)
  TC *C = static_cast<TC *>(p.C);
#pragma unroll
  for (int j = 0; j < BNB; ++j) {
#pragma unroll
    for (int i = 0; i < BMB; ++i) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t cm = TR ? m0 + 16 * i + fr : m0 + 16 * i + 4 * r + fk;
        const int64_t cn = TR ? n0 + 16 * j + 4 * r + fk : n0 + 16 * j + fr;
        if (cm >= p.M || cn >= p.N) continue;
        const double re = c0[i][j][r], im = c1[i][j][r];
        if (ksplit > 1) {
          p.ws[(split * p.M + cm) * p.N + cn] = make_double2(re, im);
          continue;
        }
        double ore = p.are * re - p.aim * im;
        double oim = p.are * im + p.aim * re;
        TC *dst = C + p.com[cm] + p.con[cn];
        if (p.beta_one) {
          double xr, xi;
          ptn::cload(dst, xr, xi);
          ore += xr;
          oim += xi;
        }
        ptn::cstore(dst, ore, oim);
      }
    }
  }
}

// Shapes the wave kernel takes: both sides within two 16-blocks; one side
// within one 16-block (the other then runs in blocks of PTN_WAVE_THIN); or,
// for K <= 256, N within three or five 16-blocks (blocks of 1 x 3 or 1 x 5:
// the 64 x 64 tile would waste most of its second column of tiles and spend a
// whole prologue on a few K steps), or M within three with N longer (3 x 1).  Returns 0 for shapes that stay
// with the 64 x 64 workgroup tile.
int wave_shape(int64_t M, int64_t N, int64_t K, int &bmb, int &bnb) {
  const int64_t bm = (M + 15) / 16, bn = (N + 15) / 16;
  if (bm <= 2 && bn <= 2) {
    bmb = (int)bm;
    bnb = (int)bn;
    return 1;
  }
  if (bn == 1) {
    bmb = PTN_WAVE_THIN;
    bnb = 1;
    return 1;
  }
  if (bm == 1) {
    bmb = 1;
    bnb = PTN_WAVE_THIN;
    return 1;
  }
  if (K <= 256 && (bn == 3 || bn == 5) && bm > bn) {
    bmb = 1;
    bnb = (int)bn;
    return 1;
  }
  if (K <= 256 && bm == 3 && bn > bm) {
    bmb = 3;
    bnb = 1;
    return 1;
  }
  return 0;
}

// launch the wave kernel for a validated descriptor; splits K over waves when
// the tiles cannot fill the chip (each split >= 64 of K, partial traffic no
// more than half the operand traffic, partials within the workspace)
template <bool TR, typename TA, typename TB, typename TC>
void launch_wave_kernel(hipStream_t s, const KArgs &a, int bmb, int bnb, dim3 g, int64_t tn,
                        int64_t nwaves, int ks) {
  const dim3 b(256);
  if (bmb == 1 && bnb == 1)
    hipLaunchKernelGGL((wave_gemm_f64<1, 1, TR, TA, TB, TC>), g, b, 0, s, a, tn, nwaves, ks);
  else if (bmb == 2 && bnb == 1)
    hipLaunchKernelGGL((wave_gemm_f64<2, 1, TR, TA, TB, TC>), g, b, 0, s, a, tn, nwaves, ks);
  else if (bmb == 1 && bnb == 2)
    hipLaunchKernelGGL((wave_gemm_f64<1, 2, TR, TA, TB, TC>), g, b, 0, s, a, tn, nwaves, ks);
  else if (bmb == 2 && bnb == 2)
    hipLaunchKernelGGL((wave_gemm_f64<2, 2, TR, TA, TB, TC>), g, b, 0, s, a, tn, nwaves, ks);
  else if (bmb == 4 && bnb == 1)
    hipLaunchKernelGGL((wave_gemm_f64<4, 1, TR, TA, TB, TC>), g, b, 0, s, a, tn, nwaves, ks);
  else if (bmb == 1 && bnb == 4)
    hipLaunchKernelGGL((wave_gemm_f64<1, 4, TR, TA, TB, TC>), g, b, 0, s, a, tn, nwaves, ks);
  else if (bmb == 1 && bnb == 3)
    hipLaunchKernelGGL((wave_gemm_f64<1, 3, TR, TA, TB, TC>), g, b, 0, s, a, tn, nwaves, ks);
  else if (bmb == 3 && bnb == 1)
    hipLaunchKernelGGL((wave_gemm_f64<3, 1, TR, TA, TB, TC>), g, b, 0, s, a, tn, nwaves, ks);
  else
    hipLaunchKernelGGL((wave_gemm_f64<1, 5, TR, TA, TB, TC>), g, b, 0, s, a, tn, nwaves, ks);
}

template <typename TA, typename TB, typename TC>
void launch_wave_typed(hipStream_t s, const KArgs &a, int bmb, int bnb, dim3 g, int64_t tn,
                       int64_t nwaves, int ks, bool tr) {
  if (tr)
    launch_wave_kernel<true, TA, TB, TC>(s, a, bmb, bnb, g, tn, nwaves, ks);
  else
    launch_wave_kernel<false, TA, TB, TC>(s, a, bmb, bnb, g, tn, nwaves, ks);
}

// the operand dtypes the library's contractions use: all complex128 (FP64), or
// complex64 with at most one complex128 side (the mixed form)
// ---- skinny products: short K, narrow N, long M -----------------------------------
// C (M x N) = A (M x K) B (K x N) with K <= SK_MAXK and N <= SK_MAXN: a stream over
// the rows of A and C with B in LDS.  Each thread owns rows: it loads the row's K
// values of A into registers and writes its N outputs, with FP64 vector FMAs (the
// product has 4 K N flops per 16 (K + N) bytes of a row, far below the MFMA's
// balance).  The MFMA kernels spend most of a 16 x 16 (or 64 x 64) block on
// padding and prologue here: a 10-deep K runs in 12 (or 16) and every block reloads
// its offsets.  Rows are dealt out grid-stride; consecutive threads take
// consecutive rows, so an m-contiguous A or C is read / written in full lines.
constexpr int SK_MAXK = 16;
constexpr int SK_MAXN = 64;
constexpr int SK_NT = 256;
constexpr int64_t SK_MIN_M = 8192;   // rows from which the rule takes the skinny kernel
                                     // (with A and C m-contiguous)

template <typename TA, typename TB, typename TC>
__global__ __launch_bounds__(SK_NT) void skinny_gemm(KArgs p) {
  __shared__ double2 sB[SK_MAXK][SK_MAXN];
  __shared__ int64_t s_aok[SK_MAXK], s_con[SK_MAXN];
  const int K = (int)p.K, N = (int)p.N;
  const TA *A = static_cast<const TA *>(p.A);
  const TB *B = static_cast<const TB *>(p.B);
  TC *C = static_cast<TC *>(p.C);
  const double asg = p.aconj ? -1.0 : 1.0, bsg = p.bconj ? -1.0 : 1.0;
  for (int e = threadIdx.x; e < K * N; e += SK_NT) {
    const int k = e / N, n = e - k * N;
    double vr, vi;
    ptn::cload(B + p.bok[k] + p.bon[n], vr, vi);
    sB[k][n] = make_double2(vr, bsg * vi);
  }
  for (int k = threadIdx.x; k < K; k += SK_NT) s_aok[k] = p.aok[k];
  for (int n = threadIdx.x; n < N; n += SK_NT) s_con[n] = p.con[n];
  __syncthreads();
  const int64_t stride = (int64_t)gridDim.x * SK_NT;
  for (int64_t m = (int64_t)blockIdx.x * SK_NT + threadIdx.x; m < p.M; m += stride) {
    const int64_t ao = p.aom[m], co = p.com[m];
    double ar[SK_MAXK], ai[SK_MAXK];
#pragma unroll
    for (int k = 0; k < SK_MAXK; ++k) {
      double vr = 0.0, vi = 0.0;
      if (k < K) ptn::cload(A + ao + s_aok[k], vr, vi);
      ar[k] = vr;
      ai[k] = asg * vi;
    }
    for (int n = 0; n < N; ++n) {
      double re = 0.0, im = 0.0;
#pragma unroll
      for (int k = 0; k < SK_MAXK; ++k) {
        if (k < K) {
          const double2 b = sB[k][n];
          re = fma(ar[k], b.x, re);
          re = fma(-ai[k], b.y, re);
          im = fma(ar[k], b.y, im);
          im = fma(ai[k], b.x, im);
        }
      }
      double ore = p.are * re - p.aim * im;
      double oim = p.are * im + p.aim * re;
      TC *dst = C + co + s_con[n];
      if (p.beta_one) {
        double xr, xi;
        ptn::cload(dst, xr, xi);
        ore += xr;
        oim += xi;
      }
      ptn::cstore(dst, ore, oim);
    }
  }
}

bool skinny_ok(int64_t M, int64_t N, int64_t K) {
  return K >= 1 && K <= SK_MAXK && N >= 1 && N <= SK_MAXN && M >= 1;
}

int launch_skinny(hipStream_t s, const KArgs &a, int adt, int bdt, int cdt) {
  int64_t blocks = (a.M + SK_NT - 1) / SK_NT;
  if (blocks > 256 * 16) blocks = 256 * 16;   // grid-stride beyond 16 workgroups per CU
  const dim3 g((unsigned)blocks), b(SK_NT);
  const int code = adt * 4 + bdt * 2 + cdt;  // 0: complex128, 1: complex64
  switch (code) {
    case 0: hipLaunchKernelGGL((skinny_gemm<double2, double2, double2>), g, b, 0, s, a); break;
    case 7: hipLaunchKernelGGL((skinny_gemm<float2, float2, float2>), g, b, 0, s, a); break;
    case 6: hipLaunchKernelGGL((skinny_gemm<float2, float2, double2>), g, b, 0, s, a); break;
    case 3: hipLaunchKernelGGL((skinny_gemm<double2, float2, float2>), g, b, 0, s, a); break;
    case 5: hipLaunchKernelGGL((skinny_gemm<float2, double2, float2>), g, b, 0, s, a); break;
    default:
      ptn::set_error("ptn_gemm_gather: operand dtype combination %d not supported", code);
      return PTN_EINVAL;
  }
  PTN_LAUNCH_CHECK();
  return PTN_OK;
}

int launch_wave(hipStream_t s, KArgs a, int bmb, int bnb, int64_t ws_bytes, bool c_mmajor,
                int adt, int bdt, int cdt) {
  const int64_t tn = (a.N + 16 * bnb - 1) / (16 * bnb);
  const int64_t tiles = ((a.M + 16 * bmb - 1) / (16 * bmb)) * tn;
  int64_t ksplit = 1;
  if (tiles < 1024 && a.K >= 128 && a.ws && ws_bytes > 0) {
    ksplit = (4096 + tiles - 1) / tiles;
    const int64_t by_k = a.K / 64;
    const int64_t by_traffic = a.K * (a.M + a.N) / (2 * a.M * a.N);
    const int64_t by_ws = ws_bytes / ((int64_t)sizeof(double2) * a.M * a.N);
    if (ksplit > by_k) ksplit = by_k;
    if (ksplit > by_traffic) ksplit = by_traffic;
    if (ksplit > by_ws) ksplit = by_ws;
    if (ksplit < 1) ksplit = 1;
  }
  const int64_t kc = (a.K + ksplit - 1) / ksplit;
  a.kchunk = a.K > 0 ? ((kc + 4 * WU - 1) / (4 * WU)) * (4 * WU) : 4 * WU;
  ksplit = a.K > 0 ? (a.K + a.kchunk - 1) / a.kchunk : 1;
  if (ksplit > 1) {
    if (!a.ws) {
      ptn::set_error("ptn_gemm_gather: split-K needs a workspace");
      return PTN_EINVAL;
    }
  } else {
    a.ws = nullptr;
  }
  const int64_t nwaves = tiles * ksplit;
  const int64_t blocks = (nwaves + 3) / 4;
  if (blocks > 0x7fffffff) {
    ptn::set_error("ptn_gemm_gather: %lld waves exceed one launch", (long long)nwaves);
    return PTN_ERANGE;
  }
  const dim3 g((unsigned)blocks);
  const int ks = (int)ksplit;
  // lanes along m for the stores when C is m-contiguous (partials: n-contiguous)
  const bool tr = c_mmajor && ksplit == 1;
  const int code = adt * 4 + bdt * 2 + cdt;  // 0: complex128, 1: complex64
  switch (code) {
    case 0: launch_wave_typed<double2, double2, double2>(s, a, bmb, bnb, g, tn, nwaves, ks, tr);
            break;
    case 7: launch_wave_typed<float2, float2, float2>(s, a, bmb, bnb, g, tn, nwaves, ks, tr);
            break;
    case 6: launch_wave_typed<float2, float2, double2>(s, a, bmb, bnb, g, tn, nwaves, ks, tr);
            break;
    case 3: launch_wave_typed<double2, float2, float2>(s, a, bmb, bnb, g, tn, nwaves, ks, tr);
            break;
    case 5: launch_wave_typed<float2, double2, float2>(s, a, bmb, bnb, g, tn, nwaves, ks, tr);
            break;
    default:
      ptn::set_error("ptn_gemm_gather: operand dtype combination %d not supported", code);
      return PTN_EINVAL;
  }
  PTN_LAUNCH_CHECK();
  if (ksplit > 1) {
    const int64_t total = a.M * a.N;
    const bool c64 = cdt == PTN_C64;
    if (total <= 16384 && ksplit >= 16) {
      const dim3 gr((unsigned)((total + 3) / 4));
      if (c64)
        hipLaunchKernelGGL((splitk_reduce_wide<float2>), gr, dim3(256), 0, s, a, ks);
      else
        hipLaunchKernelGGL((splitk_reduce_wide<double2>), gr, dim3(256), 0, s, a, ks);
    } else {
      int64_t gr = (total + 255) / 256;
      if (gr > 4096) gr = 4096;
      if (c64)
        hipLaunchKernelGGL((splitk_reduce<float2>), dim3((unsigned)gr), dim3(256), 0, s, a, ks);
      else
        hipLaunchKernelGGL((splitk_reduce<double2>), dim3((unsigned)gr), dim3(256), 0, s, a, ks);
    }
    PTN_LAUNCH_CHECK();
  }
  return PTN_OK;
}

template <int MODE, typename TA, typename TB, typename TC>
int launch(hipStream_t s, const KArgs &a, int64_t tiles, int ksplit) {
  const bool large = ksplit == 1 && tiles >= LARGE_TILES;
  if constexpr (MODE == PTN_COMPUTE_F64_4M) {
    // stores along m when C is m-contiguous (not for split-K partials)
    const bool tr = a.cmaj && ksplit == 1;
    if (large && tr)
      hipLaunchKernelGGL((gemm_direct_f64_kernel<1, true>), dim3((unsigned)tiles, 1), dim3(NT),
                         0, s, a, a.aok, a.bok);
    else if (large)
      hipLaunchKernelGGL((gemm_direct_f64_kernel<1, false>), dim3((unsigned)tiles, 1), dim3(NT),
                         0, s, a, a.aok, a.bok);
    else if (tr)
      hipLaunchKernelGGL((gemm_direct_f64_kernel<0, true>), dim3((unsigned)tiles, 1), dim3(NT),
                         0, s, a, a.aok, a.bok);
    else
      hipLaunchKernelGGL((gemm_direct_f64_kernel<0, false>), dim3((unsigned)tiles, ksplit),
                         dim3(NT), 0, s, a, a.aok, a.bok);
  } else {
    if constexpr (MODE == PTN_COMPUTE_BF16_3M) {
      if (large)
        hipLaunchKernelGGL((gemm_bf16_kernel<TA, TB, TC, 1>), dim3((unsigned)tiles, 1),
                           dim3(NT), 0, s, a);
      else
        hipLaunchKernelGGL((gemm_bf16_kernel<TA, TB, TC, 0>), dim3((unsigned)tiles, ksplit),
                           dim3(NT), 0, s, a);
    } else if constexpr (MODE == PTN_COMPUTE_BF16X6_3M) {
      if (large)
        hipLaunchKernelGGL((gemm_bf16x6_kernel<TA, TB, TC, 1>), dim3((unsigned)tiles, 1),
                           dim3(NT), 0, s, a);
      else
        hipLaunchKernelGGL((gemm_bf16x6_kernel<TA, TB, TC, 0>), dim3((unsigned)tiles, ksplit),
                           dim3(NT), 0, s, a);
    } else if (large && a.cmaj)
      hipLaunchKernelGGL((gemm_gather_kernel<MODE, TA, TB, TC, 1, true>),
                         dim3((unsigned)tiles, 1), dim3(NT), 0, s, a);
    else if (large)
      hipLaunchKernelGGL((gemm_gather_kernel<MODE, TA, TB, TC, 1>), dim3((unsigned)tiles, 1),
                         dim3(NT), 0, s, a);
    else if (a.cmaj && ksplit == 1)
      hipLaunchKernelGGL((gemm_gather_kernel<MODE, TA, TB, TC, 0, true>),
                         dim3((unsigned)tiles, 1), dim3(NT), 0, s, a);
    else
      hipLaunchKernelGGL((gemm_gather_kernel<MODE, TA, TB, TC, 0>),
                         dim3((unsigned)tiles, ksplit), dim3(NT), 0, s, a);
  }
  PTN_LAUNCH_CHECK();
  if (ksplit > 1) {
    int64_t g = (a.M * a.N + 255) / 256;
    if (g > 4096) g = 4096;
    hipLaunchKernelGGL((splitk_reduce<TC>), dim3((unsigned)g), dim3(256), 0, s, a, ksplit);
    PTN_LAUNCH_CHECK();
  }
  return PTN_OK;
}

template <int MODE, typename TA, typename TB>
int dispatch_c(hipStream_t s, const KArgs &a, int64_t tiles, int cdt, int ks) {
  if (cdt == PTN_C128) return launch<MODE, TA, TB, double2>(s, a, tiles, ks);
  return launch<MODE, TA, TB, float2>(s, a, tiles, ks);
}

template <int MODE, typename TA>
int dispatch_b(hipStream_t s, const KArgs &a, int64_t tiles, int bdt, int cdt, int ks) {
  if (bdt == PTN_C128) return dispatch_c<MODE, TA, double2>(s, a, tiles, cdt, ks);
  return dispatch_c<MODE, TA, float2>(s, a, tiles, cdt, ks);
}

}  // namespace

namespace {
// validate a descriptor and fill the kernel arguments (no split-K, no remap);
// returns PTN_OK, or 1 for an empty product (M or N zero), or an error code
int to_kargs(const ptn_gemm_desc *d, KArgs &a, int64_t &tiles) {
  if (!d) {
    ptn::set_error("ptn_gemm_gather: null descriptor");
    return PTN_EINVAL;
  }
  if (d->M < 0 || d->N < 0 || d->K < 0) {
    ptn::set_error("ptn_gemm_gather: negative extent");
    return PTN_EINVAL;
  }
  if (d->M == 0 || d->N == 0) return 1;
  const bool dt_ok = (unsigned)d->a_dtype <= 1 && (unsigned)d->b_dtype <= 1 &&
                     (unsigned)d->c_dtype <= 1;
  if (!dt_ok || (d->compute != PTN_COMPUTE_F64_4M && d->compute != PTN_COMPUTE_F32_3M &&
                 d->compute != PTN_COMPUTE_BF16_3M && d->compute != PTN_COMPUTE_BF16X6_3M)) {
    ptn::set_error("ptn_gemm_gather: bad dtype/compute code");
    return PTN_EINVAL;
  }
  if (d->compute == PTN_COMPUTE_F64_4M &&
      (d->a_dtype != PTN_C128 || d->b_dtype != PTN_C128 || d->c_dtype != PTN_C128)) {
    ptn::set_error("ptn_gemm_gather: F64_4M requires complex128 operands");
    return PTN_EINVAL;
  }
  if (!d->A || !d->B || !d->C || !d->a_off_m || !d->b_off_n || !d->c_off_m || !d->c_off_n ||
      (d->K > 0 && (!d->a_off_k || !d->b_off_k))) {
    ptn::set_error("ptn_gemm_gather: null pointer");
    return PTN_EINVAL;
  }
  a.M = d->M;
  a.N = d->N;
  a.K = d->K;
  a.A = d->A;
  a.aom = d->a_off_m;
  a.aok = d->a_off_k;
  a.B = d->B;
  a.bok = d->b_off_k;
  a.bon = d->b_off_n;
  a.C = d->C;
  a.com = d->c_off_m;
  a.con = d->c_off_n;
  a.are = d->alpha_re;
  a.aim = d->alpha_im;
  a.aconj = d->a_conj;
  a.akmaj = d->a_kmajor;
  a.bconj = d->b_conj;
  a.bkmaj = d->b_kmajor;
  a.beta_one = d->beta_one;
  const int64_t tiles_m = (d->M + BM - 1) / BM;
  a.tiles_n = (d->N + BN - 1) / BN;
  a.tiles_m = tiles_m;
  a.group_m = 0;
  tiles = tiles_m * a.tiles_n;
  if (tiles > (int64_t)(0xffffffffu / NT)) {
    ptn::set_error("ptn_gemm_gather: %lld tiles exceed one launch", (long long)tiles);
    return PTN_ERANGE;
  }
  a.remap = 0;
  a.lower = 0;
  a.cmaj = d->c_mmajor != 0;
  a.ws = nullptr;
  a.kchunk = d->K > 0 ? ((d->K + BK - 1) / BK) * BK : BK;
  return PTN_OK;
}
}  // namespace

extern "C" int ptn_gemm_gather(void *stream, const ptn_gemm_desc *d) {
  KArgs a;
  int64_t tiles = 0;
  const int rc0 = to_kargs(d, a, tiles);
  if (rc0 == 1) return PTN_OK;
  if (rc0) return rc0;
  hipStream_t s = static_cast<hipStream_t>(stream);
  int bmb = 0, bnb = 0;
  // small and thin products run on the wave kernel in both compute modes (FP64
  // MFMA, complex64 operands widened on load): they are bound by memory and
  // latency, where the wider arithmetic costs nothing
  const int code = d->a_dtype * 4 + d->b_dtype * 2 + d->c_dtype;
  const bool wave_types = code == 0 || code == 7 || code == 6 || code == 3 || code == 5;
  // (rows coalesce only when A and C are m-contiguous: consecutive threads take
  // consecutive rows)
  if (wave_types && d->M >= SK_MIN_M && !d->a_kmajor && d->c_mmajor &&
      skinny_ok(d->M, d->N, d->K))
    return launch_skinny(s, a, d->a_dtype, d->b_dtype, d->c_dtype);
  if (wave_types && wave_shape(d->M, d->N, d->K, bmb, bnb)) {
    a.ws = static_cast<double2 *>(d->ws);
    return launch_wave(s, a, bmb, bnb, d->ws ? d->ws_bytes : 0, d->c_mmajor != 0, d->a_dtype,
                       d->b_dtype, d->c_dtype);
  }
  a.remap = (tiles % 8 == 0) ? 1 : 0;
  // grouped raster for the LARGE grids with many tile columns
  if (tiles >= LARGE_TILES && a.tiles_n >= 8) a.group_m = GROUP_M;
  // split-K when the tile grid cannot fill the chip and K is long: each split
  // keeps >= SPLITK_MIN_K of K, the partials (ksplit M N) stay within the
  // operand traffic (K (M + N), at least two splits' worth) and the caller's
  // workspace
  int ksplit = 1;
  if (tiles < 256 && d->K >= 2 * SPLITK_MIN_K && d->ws && d->ws_bytes > 0) {
    int64_t want = (SPLITK_TARGET_BLOCKS + tiles - 1) / tiles;
    const int64_t by_k = d->K / SPLITK_MIN_K;
    int64_t by_traffic = d->K * (d->M + d->N) / (d->M * d->N);
    if (by_traffic < 2) by_traffic = 2;
    if (want > by_k) want = by_k;
    if (want > by_traffic) want = by_traffic;
    const int64_t by_ws = d->ws_bytes / ((int64_t)sizeof(double2) * d->M * d->N);
    if (want > by_ws) want = by_ws;
    if (want > 1024) want = 1024;
    if (want > 1) ksplit = (int)want;
  }
  const int64_t kc = (d->K + ksplit - 1) / ksplit;
  a.kchunk = ((kc + BK - 1) / BK) * BK;
  ksplit = (int)((d->K + a.kchunk - 1) / a.kchunk);
  if (ksplit < 1) ksplit = 1;
  a.ws = ksplit > 1 ? static_cast<double2 *>(d->ws) : nullptr;
  if (d->K == 0) a.kchunk = BK;
  if (d->compute == PTN_COMPUTE_F64_4M) {
    // lower-triangle hint: the same tiles and K splits as the full product,
    // launched for the tiles with tm >= tn only
    if (d->c_lower && d->M == d->N && a.tiles_m == a.tiles_n) {
      a.lower = 1;
      tiles = a.tiles_m * (a.tiles_m + 1) / 2;
      a.remap = (tiles % 8 == 0) ? 1 : 0;
    }
    return launch<PTN_COMPUTE_F64_4M, double2, double2, double2>(s, a, tiles, ksplit);
  }
  if (d->compute == PTN_COMPUTE_BF16_3M) {
    if (d->a_dtype == PTN_C128)
      return dispatch_b<PTN_COMPUTE_BF16_3M, double2>(s, a, tiles, d->b_dtype, d->c_dtype,
                                                      ksplit);
    return dispatch_b<PTN_COMPUTE_BF16_3M, float2>(s, a, tiles, d->b_dtype, d->c_dtype, ksplit);
  }
  if (d->compute == PTN_COMPUTE_BF16X6_3M) {
    if (d->a_dtype == PTN_C128)
      return dispatch_b<PTN_COMPUTE_BF16X6_3M, double2>(s, a, tiles, d->b_dtype, d->c_dtype,
                                                        ksplit);
    return dispatch_b<PTN_COMPUTE_BF16X6_3M, float2>(s, a, tiles, d->b_dtype, d->c_dtype,
                                                     ksplit);
  }
  if (d->a_dtype == PTN_C128)
    return dispatch_b<PTN_COMPUTE_F32_3M, double2>(s, a, tiles, d->b_dtype, d->c_dtype, ksplit);
  return dispatch_b<PTN_COMPUTE_F32_3M, float2>(s, a, tiles, d->b_dtype, d->c_dtype, ksplit);
}

// the (BMB, BNB) block shapes the wave kernel is instantiated for
static bool wave_block_ok(int bmb, int bnb) {
  static const int ok[][2] = {{1, 1}, {2, 1}, {1, 2}, {2, 2}, {4, 1}, {1, 4}, {1, 3}, {3, 1}, {1, 5}};
  for (const auto &b : ok)
    if (b[0] == bmb && b[1] == bnb) return true;
  return false;
}

extern "C" int ptn_gemm_gather_config(void *stream, const ptn_gemm_desc *d, int32_t kind,
                                      int32_t bmb, int32_t bnb, int32_t ksplit) {
  if (kind == PTN_GEMM_RULE) return ptn_gemm_gather(stream, d);
  KArgs a;
  int64_t tiles = 0;
  const int rc0 = to_kargs(d, a, tiles);
  if (rc0 == 1) return PTN_OK;
  if (rc0) return rc0;
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (kind == PTN_GEMM_WAVE) {
    const int code = d->a_dtype * 4 + d->b_dtype * 2 + d->c_dtype;
    const bool wave_types = code == 0 || code == 7 || code == 6 || code == 3 || code == 5;
    if (!wave_types || !wave_block_ok(bmb, bnb)) {
      ptn::set_error("ptn_gemm_gather_config: no wave kernel for blocks %dx%d (dtypes %d)", bmb,
                     bnb, code);
      return PTN_EINVAL;
    }
    a.ws = static_cast<double2 *>(d->ws);
    return launch_wave(s, a, bmb, bnb, d->ws ? d->ws_bytes : 0, d->c_mmajor != 0, d->a_dtype,
                       d->b_dtype, d->c_dtype);
  }
  if (kind == PTN_GEMM_SKINNY) {
    const int code = d->a_dtype * 4 + d->b_dtype * 2 + d->c_dtype;
    const bool types = code == 0 || code == 7 || code == 6 || code == 3 || code == 5;
    if (!types || !skinny_ok(d->M, d->N, d->K)) {
      ptn::set_error("ptn_gemm_gather_config: no skinny kernel for %lld x %lld x %lld",
                     (long long)d->M, (long long)d->N, (long long)d->K);
      return PTN_EINVAL;
    }
    return launch_skinny(s, a, d->a_dtype, d->b_dtype, d->c_dtype);
  }
  if (kind != PTN_GEMM_TILE || ksplit < 1 || ksplit > 1024) {
    ptn::set_error("ptn_gemm_gather_config: bad kind %d / ksplit %d", kind, ksplit);
    return PTN_EINVAL;
  }
  const int64_t kc = (d->K + ksplit - 1) / ksplit;
  a.kchunk = d->K > 0 ? ((kc + BK - 1) / BK) * BK : BK;
  int ks = d->K > 0 ? (int)((d->K + a.kchunk - 1) / a.kchunk) : 1;
  if (ks < 1) ks = 1;
  if (ks > 1 && (!d->ws || d->ws_bytes < (int64_t)sizeof(double2) * ks * d->M * d->N)) {
    ptn::set_error("ptn_gemm_gather_config: %d K splits need %lld workspace bytes", ks,
                   (long long)((int64_t)sizeof(double2) * ks * d->M * d->N));
    return PTN_EINVAL;
  }
  a.ws = ks > 1 ? static_cast<double2 *>(d->ws) : nullptr;
  a.remap = (tiles % 8 == 0) ? 1 : 0;
  if (d->compute == PTN_COMPUTE_F64_4M)
    return launch<PTN_COMPUTE_F64_4M, double2, double2, double2>(s, a, tiles, ks);
  if (d->compute == PTN_COMPUTE_BF16_3M) {
    if (d->a_dtype == PTN_C128)
      return dispatch_b<PTN_COMPUTE_BF16_3M, double2>(s, a, tiles, d->b_dtype, d->c_dtype, ks);
    return dispatch_b<PTN_COMPUTE_BF16_3M, float2>(s, a, tiles, d->b_dtype, d->c_dtype, ks);
  }
  if (d->compute == PTN_COMPUTE_BF16X6_3M) {
    if (d->a_dtype == PTN_C128)
      return dispatch_b<PTN_COMPUTE_BF16X6_3M, double2>(s, a, tiles, d->b_dtype, d->c_dtype,
                                                        ks);
    return dispatch_b<PTN_COMPUTE_BF16X6_3M, float2>(s, a, tiles, d->b_dtype, d->c_dtype, ks);
  }
  if (d->a_dtype == PTN_C128)
    return dispatch_b<PTN_COMPUTE_F32_3M, double2>(s, a, tiles, d->b_dtype, d->c_dtype, ks);
  return dispatch_b<PTN_COMPUTE_F32_3M, float2>(s, a, tiles, d->b_dtype, d->c_dtype, ks);
}

extern "C" const char *ptn_last_error(void);

extern "C" int ptn_gemm_gather_many(void *stream, const ptn_gemm_desc *descs, int32_t count) {
  if (count < 0 || (count > 0 && !descs)) {
    ptn::set_error("ptn_gemm_gather_many: bad arguments (count=%d)", count);
    return PTN_EINVAL;
  }
  for (int32_t i = 0; i < count; ++i) {
    const int rc = ptn_gemm_gather(stream, descs + i);
    if (rc != PTN_OK) {
      char msg[400];
      snprintf(msg, sizeof(msg), "%s", ptn_last_error());
      ptn::set_error("ptn_gemm_gather_many: descriptor %d: %s", i, msg);
      return rc;
    }
  }
  return PTN_OK;
}

extern "C" int64_t ptn_gemm_grouped_work_bytes(int32_t count) {
  if (count < 1) return 0;
  return ((((int64_t)count * sizeof(KArgs) + 255) / 256) * 256) + ((int64_t)count + 1) * 8;
}

extern "C" int ptn_gemm_gather_grouped(void *stream, const ptn_gemm_desc *descs, int32_t count,
                                       void *work) {
  if (count < 1 || count > (1 << 20) || !descs || !work) {
    ptn::set_error("ptn_gemm_gather_grouped: bad arguments (count=%d)", count);
    return PTN_EINVAL;
  }
  std::vector<KArgs> ka;
  std::vector<int64_t> prefix(1, 0);
  ka.reserve(count);
  const ptn_gemm_desc &d0 = descs[0];
  if (d0.compute == PTN_COMPUTE_BF16_3M || d0.compute == PTN_COMPUTE_BF16X6_3M) {
    ptn::set_error("ptn_gemm_gather_grouped: no grouped form of the bf16 compute modes");
    return PTN_EINVAL;
  }
  for (int i = 0; i < count; ++i) {
    const ptn_gemm_desc &d = descs[i];
    if (d.compute != d0.compute || d.a_dtype != d0.a_dtype || d.b_dtype != d0.b_dtype ||
        d.c_dtype != d0.c_dtype) {
      ptn::set_error("ptn_gemm_gather_grouped: descriptor %d differs in compute mode or dtype", i);
      return PTN_EINVAL;
    }
    KArgs a;
    int64_t tiles = 0;
    const int rc = to_kargs(&d, a, tiles);
    if (rc == 1) continue;
    if (rc) return rc;
    ka.push_back(a);
    prefix.push_back(prefix.back() + tiles);
  }
  const int n = (int)ka.size();
  if (n == 0) return PTN_OK;
  const int64_t total = prefix.back();
  // the dispatch counts work-items in 32 bits
  if (total > (int64_t)(0xffffffffu / NT)) {
    ptn::set_error("ptn_gemm_gather_grouped: %lld tiles exceed one launch; split the batch",
                   (long long)total);
    return PTN_ERANGE;
  }
  hipStream_t s = static_cast<hipStream_t>(stream);
  char *w = static_cast<char *>(work);
  const int64_t off = (((int64_t)count * sizeof(KArgs) + 255) / 256) * 256;
  int rc = ptn::stage_to_device(s, ka.data(), ka.size() * sizeof(KArgs), w);
  if (rc) return rc;
  if ((rc = ptn::stage_to_device(s, prefix.data(), prefix.size() * sizeof(int64_t), w + off)))
    return rc;
  const KArgs *dk = reinterpret_cast<const KArgs *>(w);
  const int64_t *dp = reinterpret_cast<const int64_t *>(w + off);
  const dim3 grid((unsigned)total);
  if (d0.compute == PTN_COMPUTE_F64_4M) {
    hipLaunchKernelGGL(gemm_direct_f64_grouped, grid, dim3(NT), 0, s, dk, dp, n);
  } else {
    const bool a64 = d0.a_dtype == PTN_C128, b64 = d0.b_dtype == PTN_C128,
               c64 = d0.c_dtype == PTN_C128;
#define PTN_GROUPED(TA, TB, TC) \
  hipLaunchKernelGGL((gemm_gather_grouped<PTN_COMPUTE_F32_3M, TA, TB, TC>), grid, dim3(NT), 0, s, dk, dp, n)
    if (a64 && b64 && c64) PTN_GROUPED(double2, double2, double2);
    else if (a64 && b64) PTN_GROUPED(double2, double2, float2);
    else if (a64 && c64) PTN_GROUPED(double2, float2, double2);
    else if (a64) PTN_GROUPED(double2, float2, float2);
    else if (b64 && c64) PTN_GROUPED(float2, double2, double2);
    else if (b64) PTN_GROUPED(float2, double2, float2);
    else if (c64) PTN_GROUPED(float2, float2, double2);
    else PTN_GROUPED(float2, float2, float2);
#undef PTN_GROUPED
  }
  PTN_LAUNCH_CHECK();
  return PTN_OK;
}
