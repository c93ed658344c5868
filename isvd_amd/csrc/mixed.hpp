// Mixed-precision stages: A and Omega stored as BF16, sketches and bases in FP32, the l x l
// factorizations and the integration iterations in FP64.  Device pointers, P = 1.
#pragma once

#include "stages.hpp"

namespace isvd {

// Omega^T (rows x cols, ld) as BF16: the FP32 Philox/Box-Muller stream of gaussian_omega,
// rounded to nearest even
void gaussian_omega_b(i64 rows, i64 cols, uint16_t *omegat, i64 ld, unsigned long long seed);

// Y^T := Omega^T op(A)^T on BF16 MFMA, FP32 accumulation and output
void stage_sketch_b(const Dims &d, char ordera, const uint16_t *a, i64 lda,
                    const uint16_t *omegat, i64 ldo, float *yst, i64 ldyst);

// Gramian orthogonalization applied twice in FP32 (Gram products and basis updates on FP32
// MFMA; the l x l eigendecompositions in FP64)
// (with cholqr: CholeskyQR2 where certified, as stage_orth_gramian)
void stage_orth_gramian_b(const Dims &d, float *yst, i64 ldyst, bool cholqr = false);

// FP32 tall-skinny QR (stage_orth_tsqr: CholeskyQR2 when cholqr and certified, else the
// Householder tree)
void stage_orth_tsqr_b(const Dims &d, float *yst, i64 ldyst, bool cholqr = false);

// M = Qs^T Qs by split-BF16 x3 products (FP32 accuracy), iterations in FP64 coordinates,
// Qbar^T := X^T Qs^T in FP32
void stage_integrate_kn_b(const Dims &d, const KnArgs &args, const float *yst, i64 ldyst,
                          float *qt, i64 ldqt, i64 *iter, double *err);
void stage_integrate_wy_b(const Dims &d, const WyArgs &args, const float *yst, i64 ldyst,
                          float *qt, i64 ldqt, i64 *iter, double *mu);

// Z^T := Qbar^T op(A) with Qbar split into three BF16 terms against the BF16 A (FP32
// accumulation), then Gramian (tsqr = false) or TSQR postprocessing in FP64
void stage_post_b(const Dims &d, bool tsqr, char ordera, const uint16_t *a, i64 lda,
                  const float *qt, i64 ldqt, float *s, float *ut, i64 ldut, float *vt, i64 ldvt,
                  bool want_u, bool want_v);

}  // namespace isvd
