// BF16 matrix-core products with FP32 accumulation (v_mfma_f32_16x16x32_bf16) for the
// mixed-precision pipeline, plus the precision conversions around them.
//
// An operand is BF16 data used as is, or FP32 data split on the fly into `parts` BF16 terms
// x = h0 + h1 (+ h2) (each term the round-to-nearest-even BF16 of the remainder).  For two
// split operands the product keeps the terms with i + j < max(parts_a, parts_b): with two
// parts each that is hi*hi + hi*lo + lo*hi, the "split bf16 x3" compensated FP32 product; an
// FP32 operand of three parts against exact BF16 data gives h0*b + h1*b + h2*b.
#pragma once

#include "common.hpp"

namespace isvd {

struct MxOperand {
  const void *p = nullptr;
  i64 ld = 0;        // leading dimension (elements)
  i64 stride = 0;    // elements between batched problems
  bool f32 = false;  // FP32 storage (split into `parts` BF16 terms); otherwise BF16 bits
  bool kc = false;   // K-contiguous: element (mn, kk) at p[kk + mn*ld]; else p[mn + kk*ld]
  int parts = 1;
};

// C (m x n, FP32, column-major, ldc) := alpha * opA * opB + beta * C, where
// opA(i, kk) is operand a at (mn = i, kk) and opB(kk, j) is operand b at (mn = j, kk).
// upper: only the tiles that touch the upper triangle are formed, then the strict lower
// triangle is mirrored from the upper one (a symmetric result, m == n).
void gemm_mx(const MxOperand &a, const MxOperand &b, i64 m, i64 n, i64 k, float alpha,
             float beta, float *C, i64 ldc, i64 batch = 1, i64 sC = 0, bool upper = false);

// C (m x n, FP32, ldc) := alpha op(A) op(B) + beta C for BF16 A, B (column-major, op 'N' / 'T')
// through hipBLASLt
void gemm_bf16_lib(char ta, char tb, i64 m, i64 n, i64 k, float alpha, const uint16_t *A,
                   i64 lda, const uint16_t *B, i64 ldb, float beta, float *C, i64 ldc);

// dst (rows x cols, ldd) := src (rows x cols, lds), converted
template <typename Tout, typename Tin>
void convert(i64 rows, i64 cols, const Tin *src, i64 lds, Tout *dst, i64 ldd);

// BF16 bits (round to nearest even) of an FP32 array
void f32_to_bf16(i64 count, const float *src, uint16_t *dst);

}  // namespace isvd
