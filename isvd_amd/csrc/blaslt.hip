// Plain BF16 x BF16 -> FP32 products through hipBLASLt (the library GEMM): the mixed sketch
// Y^T = Omega^T A^T, and the split-BF16 Gram and projection once their FP32 operands are
// materialized as BF16 terms (bf16.hip).  On these shapes the library's MFMA kernels ran 2.5-4x
// faster than a hand-written kernel that split the FP32 operands on the way into LDS (46 ms vs
// 14.7 ms for the sketch at 4096 x 131072 x 16384; 68.8 ms vs ~33 ms for the Gram at
// 6144 x 6144 x 262144).
#include <hipblaslt/hipblaslt.h>

#include "mx.hpp"

namespace isvd {

namespace {

#define LT_CHECK(expr)                                                                  \
  do {                                                                                  \
    hipblasStatus_t s_ = (expr);                                                        \
    if (s_ != HIPBLAS_STATUS_SUCCESS) {                                                 \
      char buf_[256];                                                                   \
      std::snprintf(buf_, sizeof(buf_), "%s failed (status %d)", #expr, (int)s_);      \
      ::isvd::fail(buf_);                                                               \
    }                                                                                   \
  } while (0)

hipblasLtHandle_t lt_handle() {
  static hipblasLtHandle_t h = nullptr;
  if (!h) LT_CHECK(hipblasLtCreate(&h));
  return h;
}

}  // namespace

void gemm_bf16_lib(char ta, char tb, i64 m, i64 n, i64 k, float alpha, const uint16_t *A,
                   i64 lda, const uint16_t *B, i64 ldb, float beta, float *C, i64 ldc) {
  if (m <= 0 || n <= 0) return;
  ISVD_CHECK(k > 0, "gemm_bf16_lib: k must be positive");
  const bool tA = (ta == 'T' || ta == 't'), tB = (tb == 'T' || tb == 't');
  hipblasLtHandle_t h = lt_handle();
  hipblasLtMatmulDesc_t op = nullptr;
  hipblasLtMatrixLayout_t la = nullptr, lb = nullptr, lc = nullptr;
  hipblasLtMatmulPreference_t pref = nullptr;
  LT_CHECK(hipblasLtMatmulDescCreate(&op, HIPBLAS_COMPUTE_32F, HIP_R_32F));
  const hipblasOperation_t opa = tA ? HIPBLAS_OP_T : HIPBLAS_OP_N;
  const hipblasOperation_t opb = tB ? HIPBLAS_OP_T : HIPBLAS_OP_N;
  LT_CHECK(hipblasLtMatmulDescSetAttribute(op, HIPBLASLT_MATMUL_DESC_TRANSA, &opa, sizeof(opa)));
  LT_CHECK(hipblasLtMatmulDescSetAttribute(op, HIPBLASLT_MATMUL_DESC_TRANSB, &opb, sizeof(opb)));
  LT_CHECK(hipblasLtMatrixLayoutCreate(&la, HIP_R_16BF, tA ? k : m, tA ? m : k, lda));
  LT_CHECK(hipblasLtMatrixLayoutCreate(&lb, HIP_R_16BF, tB ? n : k, tB ? k : n, ldb));
  LT_CHECK(hipblasLtMatrixLayoutCreate(&lc, HIP_R_32F, m, n, ldc));
  LT_CHECK(hipblasLtMatmulPreferenceCreate(&pref));
  const uint64_t max_ws = 64ull << 20;
  LT_CHECK(hipblasLtMatmulPreferenceSetAttribute(pref, HIPBLASLT_MATMUL_PREF_MAX_WORKSPACE_BYTES,
                                                 &max_ws, sizeof(max_ws)));
  hipblasLtMatmulHeuristicResult_t res[4];
  int cnt = 0;
  LT_CHECK(hipblasLtMatmulAlgoGetHeuristic(h, op, la, lb, lc, lc, pref, 4, res, &cnt));
  ISVD_CHECK(cnt > 0, "hipBLASLt offers no algorithm for this BF16 product");
  const size_t ws_bytes = res[0].workspaceSize;
  void *ws = ws_bytes ? splitk_workspace(ws_bytes) : nullptr;
  LT_CHECK(hipblasLtMatmul(h, op, &alpha, A, la, B, lb, &beta, C, lc, C, lc, &res[0].algo, ws,
                           ws_bytes, ctx().stream));
  hipblasLtMatmulPreferenceDestroy(pref);
  hipblasLtMatrixLayoutDestroy(lc);
  hipblasLtMatrixLayoutDestroy(lb);
  hipblasLtMatrixLayoutDestroy(la);
  hipblasLtMatmulDescDestroy(op);
}

}  // namespace isvd
