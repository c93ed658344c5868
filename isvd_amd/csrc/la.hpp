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

// l x l products (l <= 128): C := alpha op(A) op(B) + beta C + diag I; one short launch.
template <typename T>
void small_gemm(char ta, char tb, i64 m, i64 n, i64 k, T alpha, const T *A, i64 lda, const T *B,
                i64 ldb, T beta, T *C, i64 ldc, T diag = T(0));

// C = (I/2 + (I/4 - Z)^{1/2})^{1/2} and C^-1 (n x n, ld n) by coupled Newton-Schulz square
// roots; valid while ||4 Z||_2 < 1.  work: 6 n*n entries.  err (device, may be null):
// err[0] = ||C - I||_F.
template <typename T>
void kn_matfun(i64 n, const T *Z, T *C, T *Cinv, T *work, int iters1, int iters2,
               T *err = nullptr);

// out[0] = ||A - c I||_F  (n x n, ld n)
template <typename T>
void frob_shift(i64 n, const T *A, T c, T *out);

// In-place inverse (Gauss-Jordan, partial pivoting), n <= 128; *info_dev = 0 or the 1-based
// index of a zero pivot.
template <typename T>
void inv_small(i64 n, T *A, i64 lda, int *info_dev);

// Reductions used by the integration loops (results left on device).
template <typename T>
void dot(i64 n, const T *x, const T *y, T *result);  // result[0] = sum x*y (deterministic)

}  // namespace isvd
