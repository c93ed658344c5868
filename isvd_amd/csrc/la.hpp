// Device linear-algebra building blocks (host launchers).  Every routine enqueues work on
// ctx().stream and takes DEVICE pointers.  Matrices are column-major with explicit leading
// dimensions, exactly as the BLAS/LAPACK calls of the reference, so each stage reads like the
// reference stage it replaces.
#pragma once

#include "common.hpp"

namespace isvd {

// C := alpha * op(A) * op(B) + beta * C   (op = transpose when trans is 'T'/'t').
// Batched when batch > 1, with element strides sA, sB, sC between consecutive problems.
// MFMA (v_mfma_f64_16x16x4f64 / v_mfma_f32_16x16x4f32) tiles; split-K with a deterministic
// two-pass reduction when the output has too few tiles to fill the chip.
template <typename T>
void gemm(char transa, char transb, i64 m, i64 n, i64 k, T alpha, const T *A, i64 lda,
          const T *B, i64 ldb, T beta, T *C, i64 ldc, i64 batch = 1, i64 sA = 0, i64 sB = 0,
          i64 sC = 0);

// The same product on 64 x 64 tiles without split-K: for a short K (<= a few hundred) whose
// output alone fills the device, where a split would only add a reduction pass.
template <typename T>
void gemm_shortk(char transa, char transb, i64 m, i64 n, i64 k, T alpha, const T *A, i64 lda,
                 const T *B, i64 ldb, T beta, T *C, i64 ldc, i64 batch = 1, i64 sA = 0,
                 i64 sB = 0, i64 sC = 0);

// Raw split-K partial sums of op(A) op(B) (64 x 64 tiles): W[s] (m x n, dense column-major)
// is the product over the s-th of about ksplit K chunks.  Returns the number of chunks written;
// the caller sums them (in a fixed order) where it needs the product.
template <typename T>
int gemm_partials(char transa, char transb, i64 m, i64 n, i64 k, const T *A, i64 lda,
                  const T *B, i64 ldb, T *W, int ksplit);

// C^T := alpha * (op(A) * op(B))^T + beta * C^T: the same product stored transposed, i.e. C
// is n x m with leading dimension ldc >= n.  (M B for a symmetric M is (B^T M)^T, which takes
// the wide-tile kernel when B has few columns.)
template <typename T>
void gemm_ct(char transa, char transb, i64 m, i64 n, i64 k, T alpha, const T *A, i64 lda,
             const T *B, i64 ldb, T beta, T *C, i64 ldc);

// C := alpha op(A) op(A)^T (n x n, full storage): op(A) = A (trans 'N', A n x k) or A^T
// (trans 'T', A k x n).  Computes the upper-triangle tiles and mirrors them.
template <typename T>
void syrk_full(char trans, i64 n, i64 k, T alpha, const T *A, i64 lda, T *C, i64 ldc);

// B := alpha * op(A)   (m x n result; mkl_?omatcopy semantics, 'N' or 'T').
template <typename T>
void omatcopy(char trans, i64 m, i64 n, T alpha, const T *A, i64 lda, T *B, i64 ldb,
              i64 batch = 1, i64 sA = 0, i64 sB = 0);

// Fill an m x n block with a constant, and optionally add `diag` to its diagonal.
template <typename T>
void fill(i64 m, i64 n, T value, T *A, i64 lda, T diag = T(0), i64 batch = 1, i64 sA = 0);

// Batched one-sided Jacobi SVD of n x n matrices (n <= 128), LAPACK gesvd conventions:
// s descending (s has stride lds between problems), U (n x n, ldu) and V (n x n, ldv).
// A is overwritten.  Either of U / V may be null.  Returns nothing; sweeps are bounded.
template <typename T>
void svd_small(i64 n, T *A, i64 lda, i64 sA, T *s, i64 lds, T *U, i64 ldu, i64 sU, T *V,
               i64 ldv, i64 sV, i64 batch);

// Symmetric eigendecomposition of n x n matrices (upper triangle referenced), LAPACK syev
// conventions: eigenvalues ascending in w, eigenvectors in Z (n x n, ldz).  Built on
// svd_small of the shifted matrix (A + c I, c >= |lambda|_max), which is positive
// semi-definite, so its SVD is its eigendecomposition.
template <typename T>
void syev_small(i64 n, const T *A, i64 lda, T *w, T *Z, i64 ldz);

// Tall-skinny QR of `batch` m x l column blocks (block b at Y + b*sY, leading dimension ldy,
// i.e. ROW-MAJOR m x l panels of a wider row-major matrix when ldy >= l and the blocks are
// interleaved).  Overwrites each block with the explicit orthonormal factor Q (m x l) and,
// when R != null, writes the l x l upper-triangular factor (col-major, ldr) at R + b*sR.
// Householder TSQR: register-resident panel factorizations over row panels, a binary tree
// of stacked-R factorizations, then top-down explicit formation of Q.
template <typename T>
void tsqr_rowmajor(i64 m, i64 l, T *Y, i64 ldy, i64 sY, i64 batch, T *R, i64 ldr, i64 sR);

// The same factorization (Y overwritten by Q, R = the l x l factor when R != null) by
// CholeskyQR2 on the matrix cores: two passes of Gram, Cholesky and Y R^-1.  Returns false,
// with Y and R untouched, unless the device certifies that the result is as accurate as
// Householder QR (every pivot positive and ||Q1^T Q1 - I||_F <= 1/2 after the first pass); R's
// diagonal is positive.  The l x l Cholesky factorizations and inverses run in FP64.
template <typename T>
bool cholqr2_rowmajor(i64 m, i64 l, T *Y, i64 ldy, i64 sY, i64 batch, T *R, i64 ldr, i64 sR);
// Where CholeskyQR2 replaces a reference factorization (ISVD_CHOLQR): 0 (unset) where the
// caller allows it and the device certifies it; 1 ("0") nowhere; 2 ("1") also in the stage
// entry points (test hook), wherever certified.
int cholqr_mode();

// l x l products (l <= 128): C := alpha op(A) op(B) + beta C + diag I; one short launch.
template <typename T>
void small_gemm(char ta, char tb, i64 m, i64 n, i64 k, T alpha, const T *A, i64 lda, const T *B,
                i64 ldb, T beta, T *C, i64 ldc, T diag = T(0));

// Newton-Schulz step counts for the two square roots of kn_chain from
// rho >= the spectral radius of 4 Z (the iterates are polynomials in the symmetric Z, so the
// residual bound holds eigenvalue by eigenvalue).  The coupled iteration on A = I - R has
// residual R_k = I - W_k Y_k with ||R_{k+1}|| <= ||R_k||^2 (3 + ||R_k||) / 4; stage 1 starts
// from ||R_0|| <= rho, stage 2 from ||I - (I + P)/2|| = max |1 - sqrt(1 - x)| / 2 over
// |x| <= rho, which is <= rho / (2 (1 + sqrt(1 - rho))).  Each stage runs until the bound
// falls below the unit roundoff; false when that takes more than kMaxNs steps (rho too close
// to 1: the caller then uses the eigendecomposition).
constexpr int kMaxNs = 20;
template <typename T>
__host__ __device__ inline bool ns_counts(double rho, int *k1, int *k2) {
  const double eps = (sizeof(T) == 8) ? 0x1p-53 : 0x1p-24;
  if (!(rho < 1.0)) return false;
  bool ok = true;
  double r = rho;
  for (int stage = 0; stage < 2; ++stage) {
    int k = 0;
    while (r > eps && k < kMaxNs) {
      r = r * r * (3.0 + r) / 4.0;
      ++k;
    }
    if (r > eps) ok = false;
    *(stage == 0 ? k1 : k2) = k < 1 ? 1 : k;
    r = rho / (2.0 * (1.0 + sqrt(1.0 - rho)));
  }
  return ok;
}

// One launch for the l x l part of a Kolmogorov-Nagumo iteration (l <= 128), from the partial
// sums of [D | Dg] (l x 2l):
//   D | Dg = sum of the partials / N;  Z = Dg - D^T D;
//   C = f(Z) = (I/2 + (I/4 - Z)^(1/2))^(1/2) and C^-1: the Taylor series of f and 1/f when
//   4 ||Z^2||_F^(1/2) is small enough for a series of degree <= 31 (Paterson-Stockmeyer), else
//   coupled Newton-Schulz square roots with step counts from rho = 4 ||Z^4||_F^(1/4)
//   (ns_counts);  F = C - D C^-1;
//   W3 = [F; C^-1 / N], [F; C^-1], [F; C^-1]  (three 2l x l blocks, ld 2l, block stride 2 l^2).
// scal: [0] ||C - I||_F, [1] ||Z^4||_F (0 on the series path), [2] the rho used, [3] status:
// 0 done, 1 rho out of reach of both (or mode 1): nothing after rho is written and the caller
// takes the eigendecomposition of Z; 2 a grid barrier timed out (an error).
// mode: 0 automatic, 1 eigendecomposition, 2 Newton-Schulz (never the series).
// The workgroups meet at grid barriers (one per dependent product); every wait is bounded.
// KnChainWork holds the intermediates (Z is its first l x l block), the series coefficients
// and the barrier state; kn_chain_init prepares it once per integration.
template <typename T>
struct KnChainWork {
  T *buf = nullptr;                   // kn_chain_work_size(l) entries
  unsigned long long *bar = nullptr;  // kKnChainBarWords words
  T *Z() const { return buf; }
};
constexpr int kKnChainPart = 256;
constexpr int kKnChainCoef = 66;
constexpr int kKnChainBarWords = 8;
inline i64 kn_chain_work_size(i64 l) { return 9 * l * l + kKnChainPart + kKnChainCoef; }
// workgroups of one chain launch: two 16 x 16 tiles of the l x l result each
inline int kn_chain_blocks(i64 l) {
  const i64 t = (l + 15) / 16;
  return (int)((t * t + 1) / 2);
}
// parts: nparts raw partial sums of B^T [B | Bg] (l x 2l each, dense, as gemm_partials writes
// them); their sum / N is written to DDg (D | Dg, ld l) before the chain starts.
template <typename T>
void kn_chain(int l, T ninv, const T *parts, int nparts, T *DDg, const KnChainWork<T> &w,
              T *W3, T *scal, int mode);
template <typename T>
void kn_chain_init(const KnChainWork<T> &w, int l);
void zero_u64(unsigned long long *p, int n);


// In-place inverse (Gauss-Jordan, partial pivoting), n <= 128; *info_dev = 0 or the 1-based
// index of a zero pivot.
template <typename T>
void inv_small(i64 n, T *A, i64 lda, int *info_dev);

// Reductions used by the integration loops (results left on device).
template <typename T>
void dot(i64 n, const T *x, const T *y, T *result);  // result[0] = sum x*y (deterministic)

}  // namespace isvd
