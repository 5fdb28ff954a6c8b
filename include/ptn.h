/*
 * ptn.h -- C ABI of the gfx950 (MI355X) library behind pytenet_amd.
 *
 * The reference (cmendl/pytenet 1.3.0) is pure Python/NumPy; its swap point is
 * the set of local operations that the sweep drivers call
 * (pytenet/tdvp.py:223-238, pytenet/dmrg.py:181-189 and the functions listed
 * per entry point below). Each entry point here replaces one of them.
 *
 * Conventions
 *   - All tensors are device pointers to row-major complex128 (interleaved
 *     re, im; the layout of torch.complex128 and numpy.complex128) unless
 *     stated otherwise. Shapes follow the reference:
 *       MPS tensor A (Dl, d, Dr); MPO tensor W (Ml, d_out, d_in, Mr);
 *       left block L (Dl_ket, Ml, Dl_bra); right block R (Dr_ket, Mr, Dr_bra).
 *   - `precision` selects the arithmetic of the dense contractions:
 *       PTN_FP64   complex128 operands, FP64 MFMA (the parity mode);
 *       PTN_FP32   complex64 operands, FP32 MFMA, FP32 accumulation;
 *       PTN_BF16X3 each FP32 operand split into two bf16 terms, three bf16
 *                  MFMA products (hi*hi + hi*lo + lo*hi), FP32 accumulation;
 *       PTN_BF16X6 each FP32 operand split into three bf16 terms (the whole
 *                  FP32 significand), six bf16 MFMA products, FP32
 *                  accumulation: FP32 accuracy on the bf16 MFMA (products too
 *                  small to fill the chip run on the FP32 MFMA).
 *     Inputs and outputs stay complex128 in every mode; Krylov vectors,
 *     inner products, norms and the small tridiagonal problems are FP64.
 *   - Every call is asynchronous on `stream` (a hipStream_t), returns a status
 *     code and never throws. Scratch memory is passed in by the caller
 *     (`ws`, `ws_bytes`); the matching *_workspace() function gives the size.
 *   - Scalars written by the library (norms, energies, status words) go to
 *     device memory given by the caller, so no call synchronises the stream.
 */
#ifndef PTN_H
#define PTN_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define PTN_OK               0
#define PTN_ERR_ARG          1   /* invalid shape or argument */
#define PTN_ERR_WORKSPACE    2   /* workspace too small */
#define PTN_ERR_HIP          3   /* HIP runtime error (see ptn_last_hip_error) */
#define PTN_ERR_UNSUPPORTED  4   /* precision / shape combination not built */
#define PTN_ERR_NONFINITE    5   /* a singular value was not finite (ptn_svd_split) */

/* precision modes */
#define PTN_FP64    0
#define PTN_FP32    1
#define PTN_BF16X3  2
#define PTN_BF16X6  3
/* flag OR-ed into a precision code: the operands are block sparse (U(1)
 * quantum numbers, the reference's dense arrays with zero blocks;
 * pytenet/block_sparse_util.py:56-64). The contractions then build per-tile
 * nonzero masks of their operands and skip every tile and k-stage whose
 * product is identically zero; results equal the dense products. The gain
 * needs bonds whose quantum numbers are grouped into contiguous sectors (as the
 * bond-splitting SVD produces them). */
#define PTN_BLOCK_SPARSE 0x100

/* local operator kinds for the Krylov solvers */
#define PTN_OP_HEFF 0   /* L.W.A.R (one or two sites; for two sites pass d = d0*d1 and W = merged pair) */
#define PTN_OP_BOND 1   /* L.C.R ("zero-site" bond operator) */

/* Krylov status words (int32[4] in device memory):
 *   [0] number of Lanczos vectors used (numiter, or j+1 after a breakdown at step j)
 *   [1] j+1 if beta[j] < 100*n*eps stopped the iteration at step j, else 0; with
 *       PTN_FP32 / PTN_BF16X6 / PTN_BF16X3 matvecs also if beta[j] < rel * |A v_j|
 *       (rel = 128 eps_fp32, 128 eps_fp32, 2^-12): a residual below the matvec's
 *       rounding noise is treated as a breakdown
 *   [2] bit 0: a norm or inner product was not finite;
 *       bit 1: the tridiagonal QL iteration did not converge (the result is unusable)
 *   [3] 1 if the start vector has norm zero (the reference asserts, krylov.py:30), else 0
 * After a stop (breakdown, non-finite value, zero start) the remaining
 * iterations launch no further work on the stopped problem. */

int ptn_version(void);
const char* ptn_status_string(int status);
const char* ptn_last_hip_error(void);

/* Shape rules of K1-K4: virtual bonds Dl, Dr >= 0 (a zero-width bond, as a
 * split that keeps no singular value leaves it, pytenet/bond_ops.py:28-29,
 * gives an empty output or zeros, as NumPy's tensordot does); d, Ml, Mr >= 1
 * with d*Ml <= 32 and d*Mr <= 32 (the MPO mode mix holds one W slice in LDS;
 * two-site operators pass d = d0*d1). Other shapes return PTN_ERR_ARG. */

/* ---- K1: effective Hamiltonian, replaces pytenet/chain_ops.py:237-279 apply_local_hamiltonian(a, w, l, r)
 * out[lb, sb, rb] = sum L[l,m,lb] W[m,sb,s,m'] A[l,s,r] R[r,m',rb] */
size_t ptn_heff_workspace(int precision, int64_t Dl, int64_t d, int64_t Dr, int64_t Ml, int64_t Mr);
int ptn_heff_apply(int precision, int64_t Dl, int64_t d, int64_t Dr, int64_t Ml, int64_t Mr,
                   const void* l, const void* w, const void* r, const void* a, void* out,
                   void* ws, size_t ws_bytes, void* stream);

/* ---- K2: bond operator, replaces pytenet/chain_ops.py:282-317 apply_local_bond_contraction(c, l, r)
 * out[lb, rb] = sum L[l,m,lb] C[l,r] R[r,m,rb] */
size_t ptn_bond_workspace(int precision, int64_t Dl, int64_t Dr, int64_t M);
int ptn_bond_apply(int precision, int64_t Dl, int64_t Dr, int64_t M,
                   const void* l, const void* r, const void* c, void* out,
                   void* ws, size_t ws_bytes, void* stream);

/* ---- K3: left environment step, replaces pytenet/chain_ops.py:60-99 contraction_operator_step_left(a, b, w, l)
 * out[r, m', rb] = sum A[l,s,r] W[m,sb,s,m'] L[l,m,lb] conj(B[lb,sb,rb]); A and B are (Dl, d, Dr) */
size_t ptn_env_workspace(int precision, int64_t Dl, int64_t d, int64_t Dr, int64_t Ml, int64_t Mr);
int ptn_env_left(int precision, int64_t Dl, int64_t d, int64_t Dr, int64_t Ml, int64_t Mr,
                 const void* a, const void* b, const void* w, const void* l, void* out,
                 void* ws, size_t ws_bytes, void* stream);

/* ---- K4: right environment step, replaces pytenet/chain_ops.py:16-57 contraction_operator_step_right(a, b, w, r)
 * out[l, m, lb] = sum A[l,s,r] W[m,sb,s,m'] R[r,m',rb] conj(B[lb,sb,rb]) */
int ptn_env_right(int precision, int64_t Dl, int64_t d, int64_t Dr, int64_t Ml, int64_t Mr,
                  const void* a, const void* b, const void* w, const void* r, void* out,
                  void* ws, size_t ws_bytes, void* stream);

/* ---- K5/K7: strided complex128 matrix product, replaces the tensordot/einsum
 * pair products pytenet/mps.py:528-535 (merge), pytenet/tdvp.py:84,112 and
 * pytenet/mps.py:472,490 (absorbing R or C into a neighbour).
 * C[m*c_m + n*c_n] (+)= sum_k opA(m,k) opB(k,n), opA(m,k) = A[m*a_m + k*a_k] (conj if conjA),
 * opB(k,n) = B[k*b_k + n*b_n] (conj if conjB). Strides in complex elements.
 * The workspace is 0 bytes (ws may be NULL; it is not used). */
size_t ptn_zgemm_workspace(int64_t M, int64_t N, int64_t K);
/* Tile autotuner of all complex products (no counterpart in the reference,
 * whose products are BLAS calls): with on != 0, the first product of each
 * shape times the tile configurations on the caller's stream (synchronising
 * it once) and keeps the fastest for that shape; products that accumulate
 * into C or skip zero tiles keep the fixed choice. C must not alias A or B
 * while it is on. Process-wide; returns the previous setting. */
int ptn_set_autotune(int on);
int ptn_zgemm(int64_t M, int64_t N, int64_t K,
              const void* A, int64_t a_m, int64_t a_k, int conjA,
              const void* B, int64_t b_k, int64_t b_n, int conjB,
              void* C, int64_t c_m, int64_t c_n, int accumulate,
              void* ws, size_t ws_bytes, void* stream);

/* ---- K8-K10: device Lanczos with the local operator applied in place of the
 * reference's matvec closure (pytenet/tdvp.py:223-238, pytenet/dmrg.py:181-189). */
typedef struct {
  int kind;                 /* PTN_OP_HEFF or PTN_OP_BOND */
  int precision;            /* PTN_FP64 / PTN_FP32 / PTN_BF16X3 / PTN_BF16X6, optionally | PTN_BLOCK_SPARSE */
  int64_t Dl, d, Dr, Ml, Mr;/* bond operator: d = 1, Ml = Mr = M */
  const void* l;            /* L (Dl, Ml, Dl) */
  const void* w;            /* W (Ml, d, d, Mr); NULL for the bond operator */
  const void* r;            /* R (Dr, Mr, Dr) */
  /* Optional U(1) quantum numbers (host int64 arrays; NULL = none; zero-
   * initialise the struct when not used): left / right virtual bond (Dl / Dr
   * labels), physical (d; two sites: the flattened pair sums,
   * pytenet/block_sparse_util.py:40-44) and MPO virtual bonds (Ml / Mr), with
   * the reference's sign conventions (A: q_l + q_s = q_r, pytenet/mps.py:71;
   * W[m, s', s, m']: q_ml + q_s' = q_s + q_mr, pytenet/mpo.py:71,154; L, R:
   * q + q_m = q', pytenet/tdvp.py:62). With PTN_BLOCK_SPARSE, PTN_FP64,
   * PTN_FP32 or PTN_BF16X6 (which takes the FP32 products here), an H_eff
   * operator, batch 1 and sector-contiguous virtual bonds (every label in one
   * run of indices, as the reference's QR and SVD splits leave them), the two
   * products of H_eff (and of the environment steps, ptn_env_apply) run as
   * one grouped launch each over the admissible sector blocks only (the block
   * descriptors are built on the first call with a given set of labels and
   * cached; that first call copies them to the device synchronously).
   * Otherwise the tile-mask products run.
   * Results are the same either way: the grouped sums run over exactly the
   * nonzero terms. The labels are read during the call only. */
  const int64_t* q_l;
  const int64_t* q_s;
  const int64_t* q_r;
  const int64_t* q_ml;
  const int64_t* q_mr;
} ptn_local_op;

/* One application of a local operator (the matvec of the local solves):
 * out = H_eff a (PTN_OP_HEFF) or the bond operator on a (PTN_OP_BOND), with
 * the sector-grouped products when `op` carries quantum numbers as above.
 * Workspace: ptn_heff_workspace / ptn_bond_workspace of the operator's
 * precision and shapes. Replaces the matvec closures pytenet/tdvp.py:223-238
 * and pytenet/dmrg.py:181-189 (one call of apply_local_hamiltonian,
 * pytenet/chain_ops.py:237-279, or apply_local_bond_contraction,
 * pytenet/chain_ops.py:282-317). *grouped (optional, host) receives 1 if the
 * grouped products ran, else 0. */
int ptn_local_apply(const ptn_local_op* op, const void* a, void* out, int32_t* grouped,
                    void* ws, size_t ws_bytes, void* stream);

/* One environment step of the site described by `op` (kind PTN_OP_HEFF: Dl,
 * d, Dr, Ml, Mr, W, and the block being extended: op->l for PTN_ENV_LEFT,
 * op->r for PTN_ENV_RIGHT), with the sector-grouped products (and the
 * labels' conditions) as for ptn_local_apply, otherwise the tile-mask
 * products; results as ptn_env_left / ptn_env_right (the same arguments a, b,
 * out and workspace, ptn_env_workspace). Replaces
 * pytenet/chain_ops.py:60-99 contraction_operator_step_left and
 * pytenet/chain_ops.py:16-57 contraction_operator_step_right. */
#define PTN_ENV_LEFT  0
#define PTN_ENV_RIGHT 1
int ptn_env_apply(const ptn_local_op* op, int side, const void* a, const void* b, void* out,
                  int32_t* grouped, void* ws, size_t ws_bytes, void* stream);

size_t ptn_krylov_workspace(const ptn_local_op* op, int numiter);

/* hipGraph capture of the local solves (no counterpart in the reference):
 * with graphs on (the default), a solve whose vectors have at most 2^20
 * elements (all problems of a batch together) is captured once per shape
 * bucket and replayed as one graph launch with the new pointers; breakdowns
 * are predicated on the device. on < 0 only queries. Returns the previous
 * setting. ptn_graph_launches() counts the graph launches of the process. */
int ptn_set_graphs(int on);
int64_t ptn_graph_launches(void);

/* expm(dt*H) v by `numiter` Lanczos steps, replaces pytenet/krylov.py:122-136
 * (Hermitian branch) with pytenet/krylov.py:12-57 and :142-150. dt is complex.
 * alpha_beta (device double[2*numiter], optional): alpha then beta. */
int ptn_krylov_expm(const ptn_local_op* op, const void* v, double dt_re, double dt_im,
                    int numiter, void* out, int32_t* info, double* alpha_beta,
                    void* ws, size_t ws_bytes, void* stream);

/* Allocation-free variant of ptn_krylov_expm. No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same arguments and results as ptn_krylov_expm without the workspace: the
 * kernels use a fixed device scratch array of PTN_NOALLOC_SCRATCH_BYTES per
 * GPU (held until the process ends); problems whose ptn_krylov_workspace()
 * exceeds it are rejected with PTN_ERR_WORKSPACE. The call launches its
 * kernels directly (no graph, no launch timing, no tile autotuning) and
 * allocates nothing. Calls that overlap on one GPU share
 * the scratch array and race; serialise them (one stream, or the caller's
 * own ordering). */
#define PTN_NOALLOC_SCRATCH_BYTES (64u << 20)
int ptn_krylov_expm_noalloc(const ptn_local_op* op, const void* v, double dt_re, double dt_im,
                            int numiter, void* out, int32_t* info, double* alpha_beta,
                            void* stream);

/* lowest Ritz pair, replaces pytenet/krylov.py:110-119 with numeig = 1:
 * out = Ritz vector, *energy (device double) = lowest Ritz value. */
int ptn_krylov_eigh(const ptn_local_op* op, const void* v, int numiter, void* out,
                    double* energy, int32_t* info, double* alpha_beta,
                    void* ws, size_t ws_bytes, void* stream);

/* ---- K11: Householder QR with LAPACK's zgeqrf + zungqr conventions (beta =
 * -sign(Re alpha) * |(alpha, x)|, tau = 0 for an exactly zero sub-column),
 * replaces np.linalg.qr(mode="reduced") at pytenet/block_sparse_util.py:157.
 * A (m x n, row stride lda) -> Q (m x k, ldq), R (k x n, ldr), k = min(m, n). */
size_t ptn_qr_workspace(int64_t m, int64_t n);
int ptn_qr(int64_t m, int64_t n, const void* a, int64_t lda, void* q, int64_t ldq,
           void* r, int64_t ldr, void* ws, size_t ws_bytes, void* stream);

/* ---- K12: thin SVD A = U diag(s) Vh (one-sided Jacobi, FP64), singular values
 * in descending order; replaces np.linalg.svd(full_matrices=False) at
 * pytenet/block_sparse_util.py:294. U (m x k, ldu), s (k doubles), Vh (k x n, ldvh).
 * info (device int32[4]): [0] sweeps used, [1] 1 if not converged, [2] non-finite flag. */
size_t ptn_svd_workspace(int64_t m, int64_t n);
int ptn_svd(int64_t m, int64_t n, const void* a, int64_t lda, void* u, int64_t ldu,
            double* s, void* vh, int64_t ldvh, int32_t* info,
            void* ws, size_t ws_bytes, void* stream);

/* ---- K12 + K13: truncating two-site split of one quantum-number sector (or
 * of a tensor with zero quantum numbers); replaces mps_split_tensor_svd with
 * svd_distr "left" / "right" (pytenet/mps.py:538-568) together with
 * split_block_sparse_matrix_svd / retained_bond_indices
 * (pytenet/bond_ops.py:23-54) and the SVD at pytenet/block_sparse_util.py:294.
 * Theta (m x n row-major, leading dimension lda) -> k, the number of singular
 * values kept by the reference's discarded-weight rule with `tol` and then at
 * most `max_bond` of them (max_bond <= 0: no cap), and
 *   distr PTN_SPLIT_RIGHT: a0 = U_k (m x k), a1 = Sigma_k V_k^H (k x n);
 *   distr PTN_SPLIT_LEFT:  a0 = U_k Sigma_k (m x k), a1 = V_k^H (k x n),
 * row-major; a0 / a1 need room for m * kmax / kmax * n elements, kmax =
 * min(m, n, max_bond). s_host (host double[min(m, n)]) receives all singular
 * values in descending order; *k_out the kept count; info_host (optional host
 * int32[4]): FP64 Jacobi sweeps, not converged, non-finite, FP32 sweeps.
 * Only the isometric factor's singular vectors are computed (one-sided Jacobi
 * on the QR-preconditioned matrix without accumulating the rotations); the
 * other factor is the projection of Theta onto them (one product).
 * precond32 != 0 starts the Jacobi iteration with FP32 sweeps (the FP64 output
 * accuracy is kept). The call synchronises the stream (k sizes the outputs).
 * Returns PTN_ERR_NONFINITE when a singular value is not finite. */
#define PTN_SPLIT_LEFT  0
#define PTN_SPLIT_RIGHT 1
size_t ptn_svd_split_workspace(int64_t m, int64_t n, int distr, int precond32);
int ptn_svd_split(int64_t m, int64_t n, const void* theta, int64_t lda, int distr, double tol,
                  int64_t max_bond, int precond32, void* a0, void* a1, double* s_host,
                  int64_t* k_out, int32_t* info_host, void* ws, size_t ws_bytes, void* stream);

/* ---- K14: multiply a (n0, n1, n2) complex tensor by real factors along axis 0
 * or axis 2, in place; replaces the broadcasts at pytenet/mps.py:558-566. */
int ptn_scale_axis(int64_t n0, int64_t n1, int64_t n2, void* a, const double* s, int axis,
                   void* stream);

/* ---- a17: quantum-number sparsity of a 3-index tensor (n0, n1, n2):
 * *out (device double) = max |a[i,j,k]| over entries with q0[i]+q1[j]+q2[k] != 0
 * (pytenet/block_sparse_util.py:47-53); with zero_them != 0 those entries are
 * also set to zero (block_sparse_util.py:56-64). q* are device int64 arrays. */
int ptn_qsparse3(int64_t n0, int64_t n1, int64_t n2, const int64_t* q0, const int64_t* q1,
                 const int64_t* q2, void* a, int zero_them, double* out, void* stream);

/* ---- Batched forms for many independent chains (no counterpart in the
 * reference, which evolves one chain per call; they replace a loop of the
 * reference's tdvp_singlesite over chains, pytenet/tdvp.py:26-118). `batch`
 * problems of identical shape, every tensor stacked contiguously along a
 * leading batch axis; status words / energies per problem. */
int ptn_zgemm_batched(int64_t batch, int64_t M, int64_t N, int64_t K,
                      const void* A, int64_t a_m, int64_t a_k, int64_t bz_a, int conjA,
                      const void* B, int64_t b_k, int64_t b_n, int64_t bz_b, int conjB,
                      void* C, int64_t c_m, int64_t c_n, int64_t bz_c, int accumulate,
                      void* stream);
size_t ptn_krylov_workspace_batched(const ptn_local_op* op, int numiter, int64_t batch);
int ptn_krylov_expm_batched(const ptn_local_op* op, int64_t batch, const void* v, double dt_re,
                            double dt_im, int numiter, void* out, int32_t* info,
                            void* ws, size_t ws_bytes, void* stream);
int ptn_krylov_eigh_batched(const ptn_local_op* op, int64_t batch, const void* v, int numiter,
                            void* out, double* energy, int32_t* info,
                            void* ws, size_t ws_bytes, void* stream);
size_t ptn_env_workspace_batched(int precision, int64_t batch, int64_t Dl, int64_t d, int64_t Dr,
                                 int64_t Ml, int64_t Mr);
int ptn_env_left_batched(int precision, int64_t batch, int64_t Dl, int64_t d, int64_t Dr,
                         int64_t Ml, int64_t Mr, const void* a, const void* b, const void* w,
                         const void* l, void* out, void* ws, size_t ws_bytes, void* stream);
int ptn_env_right_batched(int precision, int64_t batch, int64_t Dl, int64_t d, int64_t Dr,
                          int64_t Ml, int64_t Mr, const void* a, const void* b, const void* w,
                          const void* r, void* out, void* ws, size_t ws_bytes, void* stream);
/* m x n row-major matrices, Q m x k (ldq = k), R k x n (ldr = n) */
size_t ptn_qr_workspace_batched(int64_t batch, int64_t m, int64_t n);
int ptn_qr_batched(int64_t batch, int64_t m, int64_t n, const void* a, void* q, void* r,
                   void* ws, size_t ws_bytes, void* stream);
/* Thin SVDs of `batch` m x n row-major matrices: U (m x k), s (k), Vh (k x n)
 * per problem, k = min(m, n). The Jacobi sweeps of all problems run together;
 * info as for ptn_svd, for the whole batch. The per-sector SVDs of
 * block_sparse_svd (pytenet/block_sparse_util.py:285-307) use it: every
 * sector is zero-padded to the largest shape of its size class, and the
 * padding stays exactly zero through the factorisation. */
size_t ptn_svd_workspace_batched(int64_t batch, int64_t m, int64_t n);
int ptn_svd_batched(int64_t batch, int64_t m, int64_t n, const void* a, void* u, double* s,
                    void* vh, int32_t* info, void* ws, size_t ws_bytes, void* stream);
/* As ptn_svd_batched, with FP32 preconditioning sweeps (matrices with at least
 * 128 columns after the adjoint): the Jacobi iteration starts on a complex64
 * copy on FP32 MFMA, its accumulated rotations are made unitary in FP64
 * (polar factor by Newton-Schulz) and the FP64 sweeps finish from there under
 * the same stopping rule, so the output has the FP64 path's accuracy. Used by
 * the mixed-precision sweeps in place of np.linalg.svd at
 * pytenet/block_sparse_util.py:294. info[3] = number of FP32 sweeps. */
size_t ptn_svd_workspace_precond32(int64_t batch, int64_t m, int64_t n);
int ptn_svd_precond32(int64_t batch, int64_t m, int64_t n, const void* a, void* u, double* s,
                      void* vh, int32_t* info, void* ws, size_t ws_bytes, void* stream);

/* ---- Live launch timing (no counterpart in the reference, which has no
 * device code): between ptn_timing_begin() and ptn_timing_end(), every launch
 * of a complex product and of the three Jacobi kernels of the SVD (Gram,
 * rotation solve, apply) is bracketed by two HIP events on the stream it runs
 * on (launches recorded into a graph are not timed). ptn_timing_end() waits
 * for them and writes, per kernel instance (largest total time first, at most
 * max_out rows), the launch count, the summed device time and the summed
 * algorithmic flops (8 per complex multiply-add) and bytes (each operand
 * once). */
typedef struct {
  char name[160];    /* e.g. "zgemm_kernel<float, c64, c64, c128, 4, 8, true, true>" */
  int64_t launches;
  double total_ms;
  double flop;
  double bytes;
} ptn_kernel_timing;
int ptn_timing_begin(void);
int ptn_timing_end(ptn_kernel_timing* out, int max_out, int* n_out);

#ifdef __cplusplus
}
#endif

#endif /* PTN_H */
