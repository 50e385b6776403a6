/*
 * ptn_hip.h -- C ABI of libptnhip.so, the MI355X (gfx950) kernels behind the
 * PyTreeNet time-evolution hot path.
 *
 * Conventions
 *   - Every buffer is caller-owned device memory; pointers are plain device
 *     addresses, sizes are element counts unless named *_bytes.
 *   - Complex data is interleaved (re, im): dtype PTN_C128 = 2 x double,
 *     PTN_C64 = 2 x float.
 *   - Every call is ordered on the given HIP stream (hipStream_t passed as
 *     void*, NULL = default stream) and returns 0 on success, a negative
 *     PTN_E* code on failure.  ptn_last_error() gives a message.
 *   - The library keeps no mutable global state except the rocSOLVER handle
 *     (one per device, created lazily under a mutex).
 *
 * Each entry point names the reference (PyTreeNet, pure Python/NumPy) function
 * whose arithmetic it replaces.  The reference has no FFI: its boundary is the
 * Python call surface, which pytreenet_amd/ mirrors and which reaches these
 * symbols through ctypes (pytreenet_amd/_lib.py).
 */
#ifndef PTN_HIP_H
#define PTN_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { PTN_C128 = 0, PTN_C64 = 1 };
/* Compute modes of ptn_gemm_gather: FP64 (complex128 operands, FP64 MFMA),
 * 3M FP32 (FP32 MFMA, FP32 accumulation; complex64 or complex128 operands
 * rounded on load), 3M bf16 (operands rounded to bf16 on load, FP32
 * accumulation) and 3M bf16x6 (every FP32 operand value split into three bf16
 * pieces, six piece products per real product on the bf16 MFMA, FP32
 * accumulation: the accuracy of the FP32 mode at the bf16 MFMA rate).  The
 * reduced modes serve the relaxed Krylov form, which the reference does not
 * have.  Small and thin products run on the FP64 wave kernel in every mode;
 * grouped launches take FP64 and FP32 only.                                  */
enum { PTN_COMPUTE_F64_4M = 0, PTN_COMPUTE_F32_3M = 1, PTN_COMPUTE_BF16_3M = 2,
       PTN_COMPUTE_BF16X6_3M = 3 };
enum {
  PTN_OK = 0,
  PTN_EINVAL = -1,     /* bad argument / shape mismatch                       */
  PTN_EHIP = -2,       /* HIP runtime error                                    */
  PTN_ESOLVER = -3,    /* rocSOLVER/rocBLAS error                              */
  PTN_ENOCONV = -4,    /* decomposition did not converge                       */
  PTN_ERANGE = -5      /* size beyond a documented limit                       */
};

/* One pairwise tensor contraction executed as a complex GEMM whose operands
 * are gathered through per-mode offset tables:
 *
 *   C[c_off_m[m] + c_off_n[n]] (beta_one ? += : =)
 *       alpha * sum_k opA(A[a_off_m[m] + a_off_k[k]]) * opB(B[b_off_k[k] + b_off_n[n]])
 *
 * opX is the identity or complex conjugation (x_conj).  Offsets are in
 * elements of the operand's dtype and live in device memory.  x_kmajor selects
 * the coalesced load orientation (1: consecutive k are contiguous; 0:
 * consecutive m (A) / n (B) are contiguous).  The leg permutations of the
 * contraction are thereby fused into the LDS-staged tile loads; no transposed
 * copy is written to HBM.  c_lower = 1 marks a square product (M == N) of
 * which only the part on and below the diagonal is read (the Gram matrix of a
 * CholeskyQR split): the library may then skip the 64 x 64 output tiles that
 * lie strictly above the diagonal, whose elements of C are left unwritten.
 * The elements it writes are those of the full product bit for bit.
 *
 * Replaces every np.tensordot on the hot path:
 *   contractions/state_operator_contraction.py:227-273 (contract_ket_ham_with_envs),
 *   :407-694 (contract_any / contract_leaf / contract_subtrees_using_dictionary),
 *   :984-1040 (contract_bond_tensor), contractions/contraction_util.py:82-290,
 *   contractions/node_contraction.py:12-141 (contract_nodes), and
 *   contractions/state_state_contraction.py:29-66 (contract_two_ttns).
 */
typedef struct ptn_gemm_desc {
  int64_t M, N, K;
  const void *A;
  const int64_t *a_off_m, *a_off_k;
  int32_t a_dtype, a_conj, a_kmajor, c_lower;
  const void *B;
  const int64_t *b_off_k, *b_off_n;
  int32_t b_dtype, b_conj, b_kmajor, _pad1;
  void *C;
  const int64_t *c_off_m, *c_off_n;
  int32_t c_dtype, beta_one;
  double alpha_re, alpha_im;
  int32_t compute;
  int32_t c_mmajor;     /* layout hint: 1 if consecutive m are contiguous in C */
  void *ws;            /* optional scratch for split-K partials (may be NULL) */
  int64_t ws_bytes;
} ptn_gemm_desc;

int ptn_gemm_gather(void *stream, const ptn_gemm_desc *desc);

/* descs[0..count) issued in order on `stream`, each exactly as by
 * ptn_gemm_gather: one call for a run of pairwise contractions, the
 * consecutive np.tensordot calls of one network contraction in the reference
 * (contractions/state_operator_contraction.py:227-273 and :407-694, the H_eff
 * and environment networks replayed inside the Krylov loop and the sweeps).
 * Stops at the first failing descriptor and returns its error, whose message
 * names the descriptor's index; the launches before it stay issued.        */
int ptn_gemm_gather_many(void *stream, const ptn_gemm_desc *descs, int32_t count);

/* The same contraction on a kernel configuration chosen by the caller (the
 * project's autotuner times the candidates once per shape and replays the
 * fastest).  No counterpart in the reference, whose np.tensordot has no
 * launch configuration.  kind:
 *   PTN_GEMM_RULE  the library's own choice (exactly ptn_gemm_gather);
 *   PTN_GEMM_WAVE  the wave kernel with (bmb, bnb) 16-blocks per wave, one of
 *                  1x1 2x1 1x2 2x2 4x1 1x4 1x3 3x1 1x5 (K split over waves by
 *                  the library's rule when desc->ws allows);
 *   PTN_GEMM_TILE  the 64 x 64 workgroup kernel with K cut into `ksplit`
 *                  ranges (1..1024; > 1 needs ksplit*M*N*16 workspace bytes);
 *   PTN_GEMM_SKINNY the row-streaming kernel for K <= 16 and N <= 64 (FP64
 *                  vector arithmetic, B in LDS; the rule takes it from M >= 8192).
 * Every configuration computes the same result up to rounding; PTN_EINVAL for
 * a configuration that does not exist or does not fit the workspace.       */
enum { PTN_GEMM_RULE = 0, PTN_GEMM_WAVE = 1, PTN_GEMM_TILE = 2, PTN_GEMM_SKINNY = 3 };
int ptn_gemm_gather_config(void *stream, const ptn_gemm_desc *desc, int32_t kind,
                           int32_t bmb, int32_t bnb, int32_t ksplit);

/* The same contraction when every offset table is an arithmetic progression
 * (the legs of each index space are already grouped in memory):
 *
 *   C[m c_sm + n c_sn] (beta_one ? += : =)
 *       alpha * sum_k opA(A[m a_sm + k a_sk]) * opB(B[k b_sk + n b_sn])
 *
 * run as one rocblas_zgemm (dtype PTN_C128) or rocblas_cgemm (PTN_C64) call
 * on `stream`.  A, B, C point at element (0, 0).  Each operand and C needs a
 * unit stride on one of its two index spaces, and a conjugated operand must
 * end up transposed in the column-major call; ptn_gemm_strided_ok() tells
 * whether a descriptor qualifies (1) or not (0) without launching anything,
 * and ptn_gemm_strided returns PTN_EINVAL for one that does not.  Replaces
 * the same np.tensordot calls as ptn_gemm_gather.                            */
typedef struct ptn_gemm_strided_desc {
  int64_t M, N, K;
  const void *A;
  int64_t a_sm, a_sk;
  int32_t a_conj, _pad0;
  const void *B;
  int64_t b_sk, b_sn;
  int32_t b_conj, _pad1;
  void *C;
  int64_t c_sm, c_sn;
  int32_t dtype, beta_one;
  double alpha_re, alpha_im;
} ptn_gemm_strided_desc;

int ptn_gemm_strided_ok(const ptn_gemm_strided_desc *desc);
int ptn_gemm_strided(void *stream, const ptn_gemm_strided_desc *desc);

/* ---- batched forms: independent problem instances in one launch ------------
 * Many instances (disorder realisations, parameter sweeps) run the same step
 * of the same algorithm at once.  The reference runs such instances one after
 * another, each through the single-instance functions cited above; these
 * entry points replace that loop for one step.  Per-instance descriptors and
 * pointers are host arrays; the library stages them into the caller's device
 * `work` buffer (size from the matching *_work_bytes function) on `stream`,
 * so a `work` buffer may be reused by the next call on the same stream.
 *
 * ptn_gemm_gather_grouped: descs[0..count) as for ptn_gemm_gather (shapes
 * may differ; compute mode and dtypes must agree; ws is ignored, no split-K);
 * workgroup b of the launch runs one output tile of the descriptor whose tile
 * range holds b.                                                             */
int64_t ptn_gemm_grouped_work_bytes(int32_t count);
int ptn_gemm_gather_grouped(void *stream, const ptn_gemm_desc *descs, int32_t count,
                            void *work);

/* ptn_krylov_orthonormalize / ptn_normalize / ptn_krylov_combine over
 * `count` instances of one vector length n (nvec <= 128): instance b uses Q[b] (row stride
 * ldq[b]), w[b], qnext[b] (qnext itself may be NULL), out[b]; combine takes
 * coeffs as a host array of count x nvec complex128 (instance-major).      */
int64_t ptn_krylov_batched_work_bytes(int32_t count);
int ptn_krylov_orthonormalize_batched(void *stream, int64_t n, int32_t nvec, int32_t count,
                                      const void *const *Q, const int64_t *ldq,
                                      void *const *w, void *const *qnext, void *const *out,
                                      void *work);
int ptn_normalize_batched(void *stream, int64_t n, int32_t count, const void *const *x,
                          void *const *y, void *const *out, void *work);
int ptn_krylov_combine_batched(void *stream, int64_t n, int32_t nvec, int32_t count,
                               const void *const *Q, const int64_t *ldq, const void *coeffs,
                               void *const *out, void *work);

/* ptn_fill / ptn_permute over `count` instances of one shape (dst[b], src[b]
 * per instance).                                                            */
int64_t ptn_ptr_batch_work_bytes(int32_t count);
int ptn_fill_batched(void *stream, int32_t count, void *const *dst, int64_t n, int32_t dtype,
                     double re, double im, void *work);
int ptn_permute_batched(void *stream, int32_t count, int32_t rank, const int64_t *dims,
                        const void *const *src, const int64_t *src_strides, int32_t src_dtype,
                        void *const *dst, const int64_t *dst_strides, int32_t dst_dtype,
                        int32_t conj, void *work);

/* Strided copy with optional conjugation and dtype conversion:
 *   dst[sum_i idx_i * dst_strides[i]] = op(src[sum_i idx_i * src_strides[i]])
 * over the box dims[0..rank-1] (rank <= 8).  Strides are in elements.
 * Replaces np.transpose/reshape copies of util/tensor_util.py:44-81
 * (tensor_matricization) and the KEEP zero-padding of
 * util/tensor_splitting.py:98-106.                                          */
int ptn_permute(void *stream, int32_t rank, const int64_t *dims,
                const void *src, const int64_t *src_strides, int32_t src_dtype,
                void *dst, const int64_t *dst_strides, int32_t dst_dtype,
                int32_t conj);

/* Fill a complex128 / complex64 buffer with a constant. */
int ptn_fill(void *stream, void *dst, int64_t n, int32_t dtype, double re, double im);

/* Krylov (Arnoldi) building blocks; all vectors complex128, contiguous.
 * Replaces the re-orthogonalisation and norms of krylov_exp_action,
 * time_evolution/time_evolution.py:721-821.
 *
 * ptn_krylov_orthogonalize: with Q = [q_0..q_{nvec-1}] (row stride ldq),
 *   w0sq        = ||w||^2 before projection,
 *   h1 = Q^H w;  w -= Q h1;  h2 = Q^H w;  w -= Q h2   (two classical passes),
 *   beta_sq     = ||w||^2 after projection.
 * Writes out[0..nvec) = h1 + h2 (complex128), out[nvec] = (w0sq, beta_sq).
 * nvec <= 128 per call (PTN_ERANGE beyond); longer bases are projected out in
 * chunks by the caller.
 * `work` must hold at least ptn_krylov_work_bytes(n, nvec) bytes.           */
int64_t ptn_krylov_work_bytes(int64_t n, int32_t nvec);
int ptn_krylov_orthogonalize(void *stream, int64_t n, const void *Q, int64_t ldq,
                             int32_t nvec, void *w, void *out, void *work);

/* ptn_krylov_orthogonalize followed by the Arnoldi normalisation
 * q_{j+1} = w / ||w|| (time_evolution/time_evolution.py:818) written to
 * `qnext` (n complex128; not written when NULL; zero when ||w|| = 0), so the
 * next H_eff application can be queued before the host reads `out`.  Short
 * vectors (n <= 8192) take one single-workgroup launch.  With `qnext` given,
 * w's content on return is unspecified: for bases of up to 16 vectors the
 * second projection is written straight into qnext, with beta_sq =
 * ||w'||^2 - ||h2||^2 (w' after the first pass; Q orthonormal).            */
int ptn_krylov_orthonormalize(void *stream, int64_t n, const void *Q, int64_t ldq,
                              int32_t nvec, void *w, void *qnext, void *out, void *work);

/* ptn_krylov_orthonormalize with its work buffer taken from a device array
 * inside the library instead of an argument, so the call allocates nothing.
 * No counterpart in the reference: it is for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  The array holds the partials of at most 128 basis
 * vectors (ptn_krylov_work_bytes(n, 128) = 2.1 MB) on every GPU the library
 * runs on, until the process ends; nvec > 128 returns PTN_ERANGE.  Every call
 * on one GPU shares that GPU's array and leaves its contents behind: calls
 * that overlap on one GPU (other streams or threads) race.                 */
int ptn_krylov_orthonormalize_noalloc(void *stream, int64_t n, const void *Q, int64_t ldq,
                                      int32_t nvec, void *w, void *qnext, void *out);

/* y = x / ||x|| (y may alias x; zero when ||x|| = 0) and out[0] = (||x||^2, 0)
 * (complex128, device): the start vector q_0 = psi / ||psi|| of
 * krylov_exp_action (time_evolution/time_evolution.py:775-781).             */
int ptn_normalize(void *stream, int64_t n, const void *x, void *y, void *out, void *work);

/* out = sum_i coeffs[i] * Q[i]   (coeffs: device complex128[nvec]);
 * also used for q_{j+1} = w / beta with nvec = 1.                           */
int ptn_krylov_combine(void *stream, int64_t n, const void *Q, int64_t ldq,
                       int32_t nvec, const void *coeffs, void *out);

/* The whole Krylov exponential y = exp(f dt H) x of a short vector in one launch
 * (replaces krylov_exp_action, time_evolution/time_evolution.py:721-830, for
 * the small local spaces of a sweep): H is the dense n x n effective
 * Hamiltonian (complex128, device, column-major: element (r, k) at H[r + k n]),
 * f = exp_re + i exp_im,
 * n <= 1024, max_steps <= min(n, 24).  Same recurrence and stop rules as the
 * ptn_normalize / ptn_krylov_orthonormalize loop (CGS2, breakdown at
 * max(eps n, 1e-12 |H q_j|), Saad estimate < krylov_tol when krylov_tol > 0);
 * the small propagator column is exp(f dt S) e_1 with S = (T + T^H)/2 when
 * hermitian != 0, else T.  work: ptn_krylov_small_work_bytes(n, max_steps)
 * bytes of device memory; info (device int32[3]): steps used, 1 when |x| = 0
 * (y = x), 1 when a stop rule ended the recurrence before max_steps.        */
int64_t ptn_krylov_small_work_bytes(int64_t n, int32_t max_steps);
int ptn_krylov_expm_small(void *stream, int64_t n, const void *H, const void *x, void *y,
                          int32_t max_steps, double time_difference, double exp_re,
                          double exp_im, double krylov_tol, int32_t hermitian, void *work,
                          void *info);

/* ||x||^2 of a complex128 vector into out[0] (double, device).              */
int ptn_norm_sq(void *stream, int64_t n, const void *x, void *out, void *work);

/* <x|y> = sum conj(x) y into out (complex128, device).                      */
int ptn_dotc(void *stream, int64_t n, const void *x, const void *y, void *out,
             void *work);

/* Cholesky factor of a Hermitian positive definite n x n complex128 matrix
 * (n <= 1024): L (n x n row-major, lower, zero upper part) with
 * L L^H = G + s I, s = shift_factor * trace(G) (the shift of shifted
 * CholeskyQR).  *status (device int) is set to 1 if a pivot is <= min_pivot
 * (the panel is numerically rank deficient), else 0.  One workgroup.
 * Together with ptn_trtri_lower and gather-GEMM Gram/update products this
 * forms the CholeskyQR3 node split that replaces np.linalg.qr in
 * tensor_qr_decomposition (util/tensor_splitting.py:60-107).                */
int ptn_potrf_lower(void *stream, int32_t n, const void *G, int64_t ldg, void *L,
                    double shift_factor, double min_pivot, void *status);

/* X = L^{-1} for a lower triangular n x n complex128 matrix (row-major,
 * n <= 1024).                                                              */
int ptn_trtri_lower(void *stream, int32_t n, const void *L, void *X);

/* Reduced QR of a column-major m x n complex128 matrix A (lda = m):
 * on return the first k = min(m,n) columns of A hold Q (column-major, lda = m)
 * and R (k x n, column-major, ldr = k) is written.  tau: k complex128.
 * Replaces np.linalg.qr in tensor_qr_decomposition,
 * util/tensor_splitting.py:60-107.                                          */
int ptn_qr(void *stream, int64_t m, int64_t n, void *A, void *R, void *tau);

/* QR(KEEP) split of a small tensor in one launch: T (rank <= 8, dims, strides
 * in elements, any layout) matricised as (other legs, row-major in leg order)
 * x (leg `leg`), m x n with m n <= 3072 and n <= 256.  Q: contiguous
 * row-major tensor of T's dims (the new bond at `leg`; columns >= min(m, n)
 * zero), R: n x n row-major (rows >= min(m, n) zero), Q R = T, diag(R) real
 * and non-negative.  Householder reflectors, so rank-deficient panels need no
 * fallback.  Replaces tensor_qr_decomposition with SplitMode.KEEP
 * (util/tensor_splitting.py:60-106) as the orthogonality-centre moves
 * (core/ttn.py:1530) and the one-site link updates
 * (time_evolution/tdvp_algorithms/onesitetdvp.py:97) call it, for the
 * small nodes of a tree.                                                    */
int ptn_qr_split_small(void *stream, int32_t rank, const int64_t *dims, const int64_t *strides,
                       int32_t leg, const void *T, void *Q, void *R);

/* Thin SVD of a column-major m x n complex128 matrix A (destroyed):
 * S (double, k = min(m,n), descending), U (m x k col-major), Vh (k x n
 * col-major).  Replaces np.linalg.svd / scipy gesvd in tensor_svd,
 * util/tensor_splitting.py:122-177.                                         */
int ptn_svd(void *stream, int64_t m, int64_t n, void *A, void *S, void *U,
            void *Vh, void *E);

/* Thin SVD of a small m x n complex128 matrix by one-sided Jacobi (Hestenes):
 * one workgroup up to 64 rotated columns; from 65 columns one launch per
 * rotation round over the whole chip, and then the call waits on the stream
 * (it reads the per-sweep convergence flags back) before it returns.
 * A (column-major, leading dimension lda; not modified) =
 * U diag(S) Vh with k = min(m, n), S (double, k, descending), U (m x k,
 * column-major, ld m), Vh (k x n, column-major, ld k).  on_rows = 0 rotates the
 * columns of A (needs m >= n), on_rows = 1 the columns of A^H (needs m <= n; for
 * the triangle R of a QR factorisation this is the preconditioned form, which
 * converges in about ten sweeps also for graded spectra).  Limits: short side
 * <= 256, long side <= 4096 (PTN_ERANGE beyond).  work: ptn_svd_jacobi_work_bytes
 * (m, n) bytes of device memory; info (device int32): 0 converged, 1 when 64
 * sweeps did not converge.  Small singular values come out with high relative
 * accuracy.  Replaces np.linalg.svd / scipy gesvd in tensor_svd
 * (util/tensor_splitting.py:122-177) for the panels of truncated splits.      */
int64_t ptn_svd_jacobi_work_bytes(int64_t m, int64_t n);
int ptn_svd_jacobi(void *stream, int64_t m, int64_t n, const void *A, int64_t lda,
                   int32_t on_rows, void *S, void *U, void *Vh, void *work, void *info);

/* ptn_svd_jacobi over `count` independent matrices in one launch (one workgroup
 * each; shapes may differ): per-matrix arguments are host arrays, ws[b] holds
 * ptn_svd_jacobi_work_bytes(m[b], n[b]) bytes.  The library stages the job
 * descriptors into `work` (ptn_svd_jacobi_batched_work_bytes(count) bytes).  The
 * batched two-site TDVP replaces the per-instance SVDs of a split point
 * (util/tensor_splitting.py:122-177, run once per instance by the reference)
 * with this one call.                                                         */
int64_t ptn_svd_jacobi_batched_work_bytes(int32_t count);
int ptn_svd_jacobi_batched(void *stream, int32_t count, const int64_t *m, const int64_t *n,
                           const void *const *A, const int64_t *lda, const int32_t *on_rows,
                           void *const *S, void *const *U, void *const *Vh,
                           void *const *ws, void *const *info, void *work);

/* Block-sparse operator application (the step of an H_eff / environment
 * contraction where the TTNO node tensor meets a large tensor):
 *   out[o_r[r] + o_g[g] + o_j[j]] =
 *       sum_{t in group g} sum_i W[t][j][i] * Y[y_r[r] + y_t[t] + y_i[i]]
 * for r < R.  Blocks t are sorted by group; group g owns t in
 * [g_begin[g], g_begin[g+1]).  W: nt x nj x ni complex128 (device).
 * y_r, o_r (R entries), y_t (nt), o_g (ng), g_begin (ng+1, int32) are device
 * arrays; y_i (ni) and o_j (nj) are host arrays.  nt, ng <= 128, nj, ni <= 4.
 * Replaces the operator-tensor np.tensordot of
 * contractions/state_operator_contraction.py:263-272 and
 * contraction_util.py:757-817 (contract_operator_tensor_ignoring_one_leg),
 * skipping the TTNO's structurally zero bond-state blocks.                 */
int ptn_block_apply(void *stream, int64_t R, int32_t nt, int32_t ng, int32_t nj, int32_t ni,
                    const void *Y, int32_t y_dtype, const int64_t *y_r, const int64_t *y_t,
                    const int64_t *y_i, int32_t conj_y, const int32_t *g_begin,
                    const void *W, int32_t conj_w, void *out, int32_t out_dtype,
                    const int64_t *o_r, const int64_t *o_g, const int64_t *o_j);

/* ptn_block_apply with the inputs staged through LDS: rows come in tiles of
 * 256 consecutive rows along the innermost row dimension, row r of a tile at
 * y_r[r0] + (r - r0) * row_stride, and every input of a tile lies within
 * [y_r[r0] + min_offset, y_r[r0] + min_offset + span).  Requires R a multiple
 * of 256 and span * sizeof(Y element) + nt nj ni * 16 <= 60 KiB.  Same results
 * as ptn_block_apply; the workgroup reads the span with coalesced loads.   */
int ptn_block_apply_staged(void *stream, int64_t R, int32_t nt, int32_t ng, int32_t nj,
                           int32_t ni, const void *Y, int32_t y_dtype, const int64_t *y_r,
                           const int64_t *y_t, const int64_t *y_i, int32_t conj_y,
                           const int32_t *g_begin, const void *W, int32_t conj_w, void *out,
                           int32_t out_dtype, const int64_t *o_r, const int64_t *o_g,
                           const int64_t *o_j, int64_t row_stride, int64_t min_offset,
                           int32_t span);

const char *ptn_last_error(void);
int ptn_version(void);

#ifdef __cplusplus
}
#endif
#endif /* PTN_HIP_H */
