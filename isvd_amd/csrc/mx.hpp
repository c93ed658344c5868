// BF16 products with FP32 accumulation for the mixed-precision pipeline (the hand-written
// matrix-core kernel of bf16_gemm.hip), plus the precision conversions around them.  FP32 data enter a product as materialized BF16 splits
// x = h0 + h1 (+ h2), each term the round-to-nearest-even BF16 of the remainder: two terms for
// the "split bf16 x3" Gram (hi*hi + hi*lo + lo*hi), three terms (exact to FP32) against exact
// BF16 data for the projection.
#pragma once

#include "common.hpp"

namespace isvd {

// C (m x n, FP32, ldc) := alpha op(A) op(B) + beta C for BF16 A, B (column-major, op 'N' / 'T')
// on the BF16 matrix cores with FP32 accumulation
void gemm_bf16(char ta, char tb, i64 m, i64 n, i64 k, float alpha, const uint16_t *A, i64 lda,
               const uint16_t *B, i64 ldb, float beta, float *C, i64 ldc);

// C := op(A) op(B) as gemm_bf16 (alpha 1, beta 0), except that the elements of C's top
// upper_rows x upper_rows block below its diagonal are not all written: the tiles of that block
// strictly below the diagonal are skipped.  For a symmetric top block, read it as
// C(min(i, j), max(i, j)).
void gemm_bf16_upper(char ta, char tb, i64 m, i64 n, i64 k, i64 upper_rows, const uint16_t *A,
                     i64 lda, const uint16_t *B, i64 ldb, float *C, i64 ldc);

// Raw split-K partial sums of op(A) op(B) for BF16 A, B: W[s] (m x n, dense, ld m) is the
// product over the s-th of at most max_split K chunks (FP32 accumulation).  Returns the number of
// chunks written; the caller sums them.
int gemm_bf16_partials(char ta, char tb, i64 m, i64 n, i64 k, const uint16_t *A, i64 lda,
                       const uint16_t *B, i64 ldb, float *W, int max_split);

// A symmetric n x n FP32 matrix held as its two BF16 terms S = [Mh; Ml] (2n x n, ld 2n, as
// split_bf16 with two parts writes them); it stands for M'' = Mh + Ml, which is exact in FP32.
// bop (2n x 3l BF16) and gw (split_sym_gw_size entries) are work buffers of split_sym_apply.
struct SplitSym {
  i64 n = 0;
  const uint16_t *S = nullptr;
  uint16_t *bop = nullptr;
  float *gw = nullptr;
  int max_split = 8;
};
i64 split_sym_gw_size(i64 n, i64 l, int max_split);
// G (n x l, ldg) := alpha M'' B for an FP32 B (n x l, ldb).  B = b0 + b1 + b2 exactly (its BF16
// terms); the product is Mh b0 + Ml b0 + Mh b1 + Ml b1 + Mh b2 on the BF16 matrix cores with
// FP32 accumulation, one BF16 GEMM [Mh | Ml] [b0 b1 b2; b0 b1 0] whose K-chunk partials and
// column blocks are summed in a fixed order.  The term left out, Ml b2, is below 2^-26 of
// |M''| |B|, under FP32 rounding, so the result is M'' B to FP32 accuracy.
void split_sym_apply(const SplitSym &m, i64 l, const float *B, i64 ldb, float alpha, float *G,
                     i64 ldg);
// B (n x l, ldb) := columns 0 .. l-1 of M''
void split_sym_columns(const SplitSym &m, i64 l, float *B, i64 ldb);

// dst (rows x cols, ldd) := src (rows x cols, lds), converted
template <typename Tout, typename Tin>
void convert(i64 rows, i64 cols, const Tin *src, i64 lds, Tout *dst, i64 ldd);

// BF16 bits (round to nearest even) of an FP32 array
void f32_to_bf16(i64 count, const float *src, uint16_t *dst);

// dst (parts*rows x cols, ldd), rows p*rows .. p*rows + rows - 1 := BF16 term p of src
// (rows x cols, lds): x = h0 + h1 (+ h2), each term the nearest-even BF16 of the remainder
void split_bf16(i64 rows, i64 cols, int parts, const float *src, i64 lds, uint16_t *dst,
                i64 ldd);

// dst (rows x cols, ldd) := the sum of the `parts` row blocks of src (parts*rows x cols)
void sum_blocks(i64 rows, i64 cols, int parts, const float *src, i64 lds, float *dst, i64 ldd);

// M (n x n) := HH + LH + LH^T from C = [H^T H ; L^T H] (2n x n), HH read from its upper
// triangle (HH(min(i, j), max(i, j))), so M is exactly symmetric
void gram_combine(i64 n, const float *C, i64 ldc, float *M, i64 ldm);

}  // namespace isvd
