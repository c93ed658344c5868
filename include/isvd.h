/*
 * isvd.h -- public C ABI of the MI355X (gfx950) iSVD library (libisvd_mi355x.so).
 *
 * Drop-in replacement for the stage and driver entry points of emfomy/isvd v1.3.2
 * (isvd_core + isvd_la_* + isvd_gpu_*): same function names, argument order and meaning,
 * same column-major layouts, same selector strings.  Each declaration cites the reference
 * declaration it replaces (paths relative to the reference repository).
 *
 * Differences a caller can observe:
 *   - One process, one GPU (device 0).  isvd_MpiComm is an integer stand-in that is ignored;
 *     isvd_createParam always describes P = 1 (the whole matrix on this GPU).
 *   - Every matrix argument may be a host pointer or a device (HBM) pointer.  Device pointers
 *     are used in place (no PCIe traffic); host pointers are staged through HBM.
 *   - Failures that the reference turns into assert() print to stderr, record a message for
 *     isvd_GetLastError() and return; set ISVD_ASSERT_ABORT=1 to abort like the reference.
 *     Unknown selector strings / characters print and return silently, as the reference does.
 *   - isvd_createParam does not reject m > n (the reference asserts m <= n, param.c:43).
 *   - Omega is drawn on the GPU from `seed` (Philox-4x32-10 + Box-Muller); the stream differs
 *     from MKL's SFMT19937 stream of the reference.  The *Omega extensions below take a
 *     caller-supplied Omega instead.
 *   - Tall-skinny-QR postprocessing returns the mathematically correct U and V for k < l
 *     (the reference swaps the SVD factors of R there; identical when k == l).
 *   - Dimension limit: l = k + p <= 128 for the orthogonalization / integration /
 *     postprocessing kernels.
 */
#ifndef ISVD_MI355X_ISVD_H_
#define ISVD_MI355X_ISVD_H_

#include <stddef.h>
#include <stdint.h>
#include <stdio.h>

#if defined(__cplusplus)
extern "C" {
#endif

/* ---- types: src/include/c/isvd/def.h:45-95 ---- */
#if !defined(ISVD_USE_ILP64)
typedef int32_t isvd_int_t;
#else
typedef int64_t isvd_int_t;
#endif
typedef float isvd_s_val_t;
typedef double isvd_d_val_t;
typedef int mpi_int_t;
typedef int isvd_MpiComm; /* replaces MPI_Comm (def.h:95); ignored, P = 1 */

#define ISVD_COMM_WORLD 0

/* ---- parameters: src/include/c/isvd/core/param.h:27-92 (same fields, same order) ---- */
typedef struct {
  const isvd_MpiComm mpi_comm;
  const mpi_int_t mpi_size;
  const mpi_int_t mpi_root;
  const mpi_int_t mpi_rank;
  const isvd_int_t nrow;
  const isvd_int_t ncol;
  const isvd_int_t nrow_proc;
  const isvd_int_t ncol_proc;
  const isvd_int_t nrow_each;
  const isvd_int_t ncol_each;
  const isvd_int_t nrow_total;
  const isvd_int_t ncol_total;
  const isvd_int_t rowidxbegin;
  const isvd_int_t rowidxend;
  const isvd_int_t colidxbegin;
  const isvd_int_t colidxend;
  const isvd_int_t rank;
  const isvd_int_t over_rank;
  const isvd_int_t dim_sketch;
  const isvd_int_t dim_sketch_total;
  const isvd_int_t num_sketch;
} isvd_Param;

/* replaces isvd_createParam (src/include/c/isvd/core/param.h:100-103, param.c:14-83) */
isvd_Param isvd_createParam(const isvd_int_t nrow, const isvd_int_t ncol, const isvd_int_t rank,
                            const isvd_int_t over_rank, const isvd_int_t num_sketch,
                            const mpi_int_t mpi_root, const isvd_MpiComm mpi_comm);

/* ---- environment: src/include/c/isvd/core/env.h:30-55, src/include/c/isvd/gpu/env.h:19-25 ---- */
/* gpu/env.h:19.  Bytes of HBM for staging a host-resident A (0 = no limit).  As in the
 * reference's GPU stages (gpu/stage/@x@_sketch_gpu.h:63-77, @x@_postprocess_gpu.h:73-82), the
 * sketch and the projection Z^T = Qbar^T A of the s and d families then stage A by blocks of
 * columns (multiples of 64) that fit min(free HBM, limit) beside the buffers resident through
 * the stage (the sketch, the bases, Omega^T, Z^T, the outputs), and fail with an error when not
 * even one column fits.  A device-resident A is used in place and not counted; the b and e
 * families and symmetric postprocessing stage a host-resident A whole and fail when it does not
 * fit.  The integration stages' own workspace is not counted (the reference runs them on the
 * CPU). */
extern size_t isvd_gpu_memory_limit;
void isvd_Init_cpu(int *argcp, char ***argvp, const isvd_MpiComm mpi_comm);
void isvd_Finalize_cpu(void);
void isvd_PrintEnvironment_cpu(const isvd_MpiComm comm);
void isvd_Init_gpu(int *argcp, char ***argvp, const isvd_MpiComm mpi_comm);
void isvd_Finalize_gpu(void);
void isvd_PrintEnvironment_gpu(const isvd_MpiComm comm);

static inline void isvd_Init(int *argcp, char ***argvp, const isvd_MpiComm mpi_comm) {
  isvd_Init_cpu(argcp, argvp, mpi_comm);
  isvd_Init_gpu(argcp, argvp, mpi_comm);
}
static inline void isvd_Finalize(void) {
  isvd_Finalize_cpu();
  isvd_Finalize_gpu();
}
static inline void isvd_PrintEnvironment(const isvd_MpiComm mpi_comm) {
  isvd_PrintEnvironment_cpu(mpi_comm);
  isvd_PrintEnvironment_gpu(mpi_comm);
}

/* ---- host memory: src/include/c/isvd/util/memory.h:19-25 (64-byte aligned host memory) ---- */
isvd_s_val_t *isvd_sMalloc(const size_t num);
isvd_d_val_t *isvd_dMalloc(const size_t num);
void isvd_Free(void *ptr);

/* ---- error state (no reference counterpart: the reference asserts) ---- */
const char *isvd_GetLastError(void); /* "" when the last call succeeded */
int isvd_GetErrorCount(void);        /* failures since load */

/* =========================================================================================
 * Driver: replaces isvd_?Isvd (src/include/c/isvd/core/@x@_driver.h:21-29,
 * src/lib/libisvd/core/driver/@x@_isvd.c:106-282).  Selectors (case-insensitive):
 *   alg_s "GP" | "GP_gpu";  alg_o "TS" | "GR";  alg_i "KN" | "WY" | "HR";
 *   alg_p "TS" | "GR" | "SY" | "TS_gpu" | "GR_gpu" | "SY_gpu".   All run on the GPU.
 * time[4] receives per-stage seconds (GPU event time, stream-ordered).
 * ========================================================================================= */
#define ISVD_DECLARE_DRIVER(x, T)                                                              \
  void isvd_##x##Isvd(const isvd_Param param, const char *alg_s, const char *alg_o,           \
                      const char *alg_i, const char *alg_p, const T *argv[4],                 \
                      const isvd_int_t argc[4], T *retv[4], const isvd_int_t retc[4],         \
                      double time[4], FILE *stream, const char dista, const char ordera,      \
                      const T *a, const isvd_int_t lda, T *s, T *ut, const isvd_int_t ldut,   \
                      T *vt, const isvd_int_t ldvt, const isvd_int_t seed,                    \
                      const mpi_int_t ut_root, const mpi_int_t vt_root);
ISVD_DECLARE_DRIVER(s, isvd_s_val_t)
ISVD_DECLARE_DRIVER(d, isvd_d_val_t)

/* Extension of isvd_?Isvd: Omega (n x N*l, stored row-major, i.e. omegat = Omega^T col-major
 * with leading dimension ldomegat >= N*l) is supplied by the caller instead of drawn from a
 * seed.  Used for parity runs where both sides consume the same host-generated Omega. */
#define ISVD_DECLARE_DRIVER_OMEGA(x, T)                                                        \
  void isvd_##x##IsvdOmega(const isvd_Param param, const char *alg_s, const char *alg_o,      \
                           const char *alg_i, const char *alg_p, const T *argv[4],            \
                           const isvd_int_t argc[4], T *retv[4], const isvd_int_t retc[4],    \
                           double time[4], FILE *stream, const char dista, const char ordera, \
                           const T *a, const isvd_int_t lda, const T *omegat,                 \
                           const isvd_int_t ldomegat, T *s, T *ut, const isvd_int_t ldut,     \
                           T *vt, const isvd_int_t ldvt, const mpi_int_t ut_root,             \
                           const mpi_int_t vt_root);
ISVD_DECLARE_DRIVER_OMEGA(s, isvd_s_val_t)
ISVD_DECLARE_DRIVER_OMEGA(d, isvd_d_val_t)

/* =========================================================================================
 * Stages: replace src/include/c/isvd/core/@x@_stage.h:20-77 (and the GPU twins of
 * src/include/c/isvd/gpu/@x@_stage.h:28-80, which here are the same implementation).
 * ========================================================================================= */
#define ISVD_DECLARE_STAGES(x, T)                                                              \
  /* @x@_stage.h:21-25 / @x@_sketch_gaussian_projection.c:46-84 */                            \
  void isvd_##x##SketchGaussianProjection(                                                     \
      const isvd_Param param, const T *argv, const isvd_int_t argc, T *retv,                  \
      const isvd_int_t retc, const char dista, const char ordera, const T *a,                 \
      const isvd_int_t lda, T *yst, const isvd_int_t ldyst, const isvd_int_t seed,            \
      const mpi_int_t mpi_root);                                                              \
  /* gpu/@x@_stage.h:28-33 */                                                                 \
  void isvd_##x##SketchGaussianProjection_gpu(                                                 \
      const isvd_Param param, const T *argv, const isvd_int_t argc, T *retv,                  \
      const isvd_int_t retc, const char dista, const char ordera, const T *a,                 \
      const isvd_int_t lda, T *yst, const isvd_int_t ldyst, const isvd_int_t seed,            \
      const mpi_int_t mpi_root);                                                              \
  /* extension: caller-supplied Omega (omegat = Omega^T, Nl x n col-major, ld >= Nl) */        \
  void isvd_##x##SketchGaussianProjectionOmega(                                                \
      const isvd_Param param, const char dista, const char ordera, const T *a,                \
      const isvd_int_t lda, const T *omegat, const isvd_int_t ldomegat, T *yst,               \
      const isvd_int_t ldyst);                                                                \
  /* @x@_stage.h:28-31 / @x@_orthogonalize_tall_skinny_qr.c:36-212 */                         \
  void isvd_##x##OrthogonalizeTallSkinnyQr(const isvd_Param param, const T *argv,             \
                                           const isvd_int_t argc, T *retv,                    \
                                           const isvd_int_t retc, T *yst,                     \
                                           const isvd_int_t ldyst);                           \
  /* @x@_stage.h:33-36 / @x@_orthogonalize_gramian.c:33-107 */                                \
  void isvd_##x##OrthogonalizeGramian(const isvd_Param param, const T *argv,                  \
                                      const isvd_int_t argc, T *retv, const isvd_int_t retc,  \
                                      T *yst, const isvd_int_t ldyst);                        \
  /* @x@_stage.h:39-42 / @x@_integrate_kolmogorov_nagumo.c:42-277 */                          \
  void isvd_##x##IntegrateKolmogorovNagumo(const isvd_Param param, const T *argv,             \
                                           const isvd_int_t argc, T *retv,                    \
                                           const isvd_int_t retc, const T *yst,               \
                                           const isvd_int_t ldyst, T *qt,                     \
                                           const isvd_int_t ldqt);                            \
  /* @x@_stage.h:44-47 / @x@_integrate_wen_yin.c:56-386 */                                    \
  void isvd_##x##IntegrateWenYin(const isvd_Param param, const T *argv, const isvd_int_t argc, \
                                 T *retv, const isvd_int_t retc, const T *yst,                \
                                 const isvd_int_t ldyst, T *qt, const isvd_int_t ldqt);       \
  /* @x@_stage.h:49-53 / @x@_integrate_hierarchical_reduction.c:36-141 (yst is destroyed) */   \
  void isvd_##x##IntegrateHierarchicalReduction(const isvd_Param param, const T *argv,         \
                                                const isvd_int_t argc, T *retv,                \
                                                const isvd_int_t retc, T *yst,                 \
                                                const isvd_int_t ldyst, T *qt,                 \
                                                const isvd_int_t ldqt);                        \
  /* @x@_stage.h:71-77 / @x@_postprocess_symmetric.c:64-193 (m == n; s ascending; V = U) */    \
  void isvd_##x##PostprocessSymmetric(                                                         \
      const isvd_Param param, const T *argv, const isvd_int_t argc, T *retv,                  \
      const isvd_int_t retc, const char dista, const char ordera, const T *a,                 \
      const isvd_int_t lda, const T *qt, const isvd_int_t ldqt, T *s, T *ut,                  \
      const isvd_int_t ldut, T *vt, const isvd_int_t ldvt, const mpi_int_t ut_root,          \
      const mpi_int_t vt_root);                                                               \
  /* gpu/@x@_stage.h:74-80 */                                                                 \
  void isvd_##x##PostprocessSymmetric_gpu(                                                     \
      const isvd_Param param, const T *argv, const isvd_int_t argc, T *retv,                  \
      const isvd_int_t retc, const char dista, const char ordera, const T *a,                 \
      const isvd_int_t lda, const T *qt, const isvd_int_t ldqt, T *s, T *ut,                  \
      const isvd_int_t ldut, T *vt, const isvd_int_t ldvt, const mpi_int_t ut_root,          \
      const mpi_int_t vt_root);                                                               \
  /* @x@_stage.h:55-62 / @x@_postprocess_tall_skinny_qr.c:67-309 */                           \
  void isvd_##x##PostprocessTallSkinnyQr(                                                      \
      const isvd_Param param, const T *argv, const isvd_int_t argc, T *retv,                  \
      const isvd_int_t retc, const char dista, const char ordera, const T *a,                 \
      const isvd_int_t lda, const T *qt, const isvd_int_t ldqt, T *s, T *ut,                  \
      const isvd_int_t ldut, T *vt, const isvd_int_t ldvt, const mpi_int_t ut_root,          \
      const mpi_int_t vt_root);                                                               \
  /* @x@_stage.h:64-71 / @x@_postprocess_gramian.c:64-193 */                                  \
  void isvd_##x##PostprocessGramian(                                                           \
      const isvd_Param param, const T *argv, const isvd_int_t argc, T *retv,                  \
      const isvd_int_t retc, const char dista, const char ordera, const T *a,                 \
      const isvd_int_t lda, const T *qt, const isvd_int_t ldqt, T *s, T *ut,                  \
      const isvd_int_t ldut, T *vt, const isvd_int_t ldvt, const mpi_int_t ut_root,          \
      const mpi_int_t vt_root);                                                               \
  /* gpu/@x@_stage.h:42-48 */                                                                 \
  void isvd_##x##PostprocessTallSkinnyQr_gpu(                                                  \
      const isvd_Param param, const T *argv, const isvd_int_t argc, T *retv,                  \
      const isvd_int_t retc, const char dista, const char ordera, const T *a,                 \
      const isvd_int_t lda, const T *qt, const isvd_int_t ldqt, T *s, T *ut,                  \
      const isvd_int_t ldut, T *vt, const isvd_int_t ldvt, const mpi_int_t ut_root,          \
      const mpi_int_t vt_root);                                                               \
  /* gpu/@x@_stage.h:57-63 */                                                                 \
  void isvd_##x##PostprocessGramian_gpu(                                                       \
      const isvd_Param param, const T *argv, const isvd_int_t argc, T *retv,                  \
      const isvd_int_t retc, const char dista, const char ordera, const T *a,                 \
      const isvd_int_t lda, const T *qt, const isvd_int_t ldqt, T *s, T *ut,                  \
      const isvd_int_t ldut, T *vt, const isvd_int_t ldvt, const mpi_int_t ut_root,          \
      const mpi_int_t vt_root);
ISVD_DECLARE_STAGES(s, isvd_s_val_t)
ISVD_DECLARE_STAGES(d, isvd_d_val_t)

/* =========================================================================================
 * Gramian postprocessing that allocates nothing (no counterpart in the reference; for callers
 * that cannot allocate device memory: tight serving loops, harnesses that forbid hipMalloc).
 * Same arguments, checks and results as
 * isvd_dPostprocessGramian, except:
 *   - a, qt, s, ut and vt must be device memory (host arrays are rejected, since staging them
 *     would allocate);
 *   - every temporary (Z^T, the l x l Gram and its singular vectors, split-K partial sums)
 *     lives in a fixed 64 MiB region of device memory per GPU that the library holds until
 *     the process ends; a problem with 8 (l n + 2 l^2) bytes > 64 MiB is rejected with an
 *     error (l = 128: n <= 65280).  When the split-K partial sums of a product do not fit
 *     beside those, the product uses fewer K chunks (results then differ by FP64 rounding);
 *   - calls on one GPU share the region and keep the last call's contents; calls that
 *     overlapped on one GPU would race.  Within a process the library lock serializes them
 *     and each call drains its stream before returning.
 * ========================================================================================= */
void isvd_dPostprocessGramianNoAlloc(const isvd_Param param, const isvd_d_val_t *argv,
                                     const isvd_int_t argc, isvd_d_val_t *retv,
                                     const isvd_int_t retc, const char dista, const char ordera,
                                     const isvd_d_val_t *a, const isvd_int_t lda,
                                     const isvd_d_val_t *qt, const isvd_int_t ldqt,
                                     isvd_d_val_t *s, isvd_d_val_t *ut, const isvd_int_t ldut,
                                     isvd_d_val_t *vt, const isvd_int_t ldvt,
                                     const mpi_int_t ut_root, const mpi_int_t vt_root);

/* =========================================================================================
 * Adjoint of the driver's rank-k result (no counterpart in the reference; for callers that
 * differentiate through the decomposition, e.g. a torch.autograd.Function).
 * Given dL/dS (gs, k values), dL/dU^T (gut, k x m, ld >= k) and dL/dV^T (gvt, k x n,
 * ld >= k) -- any of them NULL for zero -- writes dL/dA (da, same storage order as A,
 * ldda >= m for ordera 'C', >= n for 'R').  U, S and V are recomputed from Qbar^T (qt, l x m,
 * the output of the integration stage) by the postprocessing named by alg_p, as in the
 * forward call.  With all l computed triplets and the complement P = I - U U^T:
 *   dA = U [(F o (J - J^T)) S + S (F o (K - K^T)) + diag(gS)] V^T + E V_k^T + U_k H^T,
 *   J = U^T gU, K = V^T gV, F_ij = 1 / (s_j^2 - s_i^2) for i != j (0 when the gap is below
 *   1e-12 (double) / 1e-6 (single) of max(s_i^2, s_j^2)); gradients padded with zeros to l;
 *   E = X S_k + P A Y, H = A^T X + Y S_k, with x_i = (s_i^2 I - A_c A_c^T)^-1 P gU_i and
 *   y_i = (s_i^2 I - A_c^T A_c)^-1 (I - V V^T) gV_i, A_c = P A (summed as Neumann series).
 * This couples the top-k triplets to every discarded singular pair, so with the exact top-l
 * subspace it is the derivative of the top-k triplets of the dense SVD of A at any rank.  A
 * column whose series does not converge (the discarded part of A reaches s_i) keeps the first
 * term, x_i = P gU_i / s_i^2 (and y_i likewise): the truncated-SVD backward
 *   (I - U U^T) gU S^-1 V^T + U S^-1 gV^T (I - V V^T)
 * for that column, which treats the complement as carrying no singular value.
 * ========================================================================================= */
#define ISVD_DECLARE_BACKWARD(x, T)                                                            \
  void isvd_##x##IsvdBackward(const isvd_Param param, const char *alg_p, const char dista,     \
                              const char ordera, const T *a, const isvd_int_t lda,             \
                              const T *qt, const isvd_int_t ldqt, const T *gs, const T *gut,   \
                              const isvd_int_t ldgut, const T *gvt, const isvd_int_t ldgvt,    \
                              T *da, const isvd_int_t ldda);
ISVD_DECLARE_BACKWARD(s, isvd_s_val_t)
ISVD_DECLARE_BACKWARD(d, isvd_d_val_t)

/* =========================================================================================
 * Mixed precision, the "b" family (no counterpart in the reference; same argument order and
 * meaning as the d family above).  A and Omega are stored as BF16 (isvd_b_val_t holds the
 * bit pattern: the upper half of the FP32 value); every other array is FP32.
 *   sketch       Y = A Omega on BF16 matrix cores with FP32 accumulation;
 *   alg_o "GR"   Gramian orthogonalization applied twice in FP32 (FP64 l x l eigensolves);
 *   alg_o "TS"   FP32 Householder TSQR;
 *   alg_i        M = Qs^T Qs from split-BF16 x3 compensated FP32 products (hi*hi + hi*lo +
 *                lo*hi), the KN / WY iterations in FP64 on l x l and N*l x l matrices;
 *   alg_p        Z^T = Qbar^T A with Qbar split into three BF16 terms (exact against the BF16
 *                A, FP32 accumulation), then the l x l factorizations, U and V in FP64.
 * The seeded forms draw Omega from the FP32 Philox stream of isvd_sIsvd, rounded to BF16.
 * ========================================================================================= */
typedef uint16_t isvd_b_val_t;
void isvd_bIsvd(const isvd_Param param, const char *alg_s, const char *alg_o, const char *alg_i,
                const char *alg_p, const float *argv[4], const isvd_int_t argc[4],
                float *retv[4], const isvd_int_t retc[4], double time[4], FILE *stream,
                const char dista, const char ordera, const isvd_b_val_t *a,
                const isvd_int_t lda, float *s, float *ut, const isvd_int_t ldut, float *vt,
                const isvd_int_t ldvt, const isvd_int_t seed, const mpi_int_t ut_root,
                const mpi_int_t vt_root);
void isvd_bIsvdOmega(const isvd_Param param, const char *alg_s, const char *alg_o,
                     const char *alg_i, const char *alg_p, const float *argv[4],
                     const isvd_int_t argc[4], float *retv[4], const isvd_int_t retc[4],
                     double time[4], FILE *stream, const char dista, const char ordera,
                     const isvd_b_val_t *a, const isvd_int_t lda, const isvd_b_val_t *omegat,
                     const isvd_int_t ldomegat, float *s, float *ut, const isvd_int_t ldut,
                     float *vt, const isvd_int_t ldvt, const mpi_int_t ut_root,
                     const mpi_int_t vt_root);
void isvd_bSketchGaussianProjection(const isvd_Param param, const float *argv,
                                    const isvd_int_t argc, float *retv, const isvd_int_t retc,
                                    const char dista, const char ordera, const isvd_b_val_t *a,
                                    const isvd_int_t lda, float *yst, const isvd_int_t ldyst,
                                    const isvd_int_t seed, const mpi_int_t mpi_root);
void isvd_bSketchGaussianProjectionOmega(const isvd_Param param, const char dista,
                                         const char ordera, const isvd_b_val_t *a,
                                         const isvd_int_t lda, const isvd_b_val_t *omegat,
                                         const isvd_int_t ldomegat, float *yst,
                                         const isvd_int_t ldyst);
void isvd_bOrthogonalizeTallSkinnyQr(const isvd_Param param, const float *argv,
                                     const isvd_int_t argc, float *retv, const isvd_int_t retc,
                                     float *yst, const isvd_int_t ldyst);
void isvd_bOrthogonalizeGramian(const isvd_Param param, const float *argv, const isvd_int_t argc,
                                float *retv, const isvd_int_t retc, float *yst,
                                const isvd_int_t ldyst);
void isvd_bIntegrateKolmogorovNagumo(const isvd_Param param, const float *argv,
                                     const isvd_int_t argc, float *retv, const isvd_int_t retc,
                                     const float *yst, const isvd_int_t ldyst, float *qt,
                                     const isvd_int_t ldqt);
void isvd_bIntegrateWenYin(const isvd_Param param, const float *argv, const isvd_int_t argc,
                           float *retv, const isvd_int_t retc, const float *yst,
                           const isvd_int_t ldyst, float *qt, const isvd_int_t ldqt);
void isvd_bPostprocessTallSkinnyQr(const isvd_Param param, const float *argv,
                                   const isvd_int_t argc, float *retv, const isvd_int_t retc,
                                   const char dista, const char ordera, const isvd_b_val_t *a,
                                   const isvd_int_t lda, const float *qt, const isvd_int_t ldqt,
                                   float *s, float *ut, const isvd_int_t ldut, float *vt,
                                   const isvd_int_t ldvt, const mpi_int_t ut_root,
                                   const mpi_int_t vt_root);
void isvd_bPostprocessGramian(const isvd_Param param, const float *argv, const isvd_int_t argc,
                              float *retv, const isvd_int_t retc, const char dista,
                              const char ordera, const isvd_b_val_t *a, const isvd_int_t lda,
                              const float *qt, const isvd_int_t ldqt, float *s, float *ut,
                              const isvd_int_t ldut, float *vt, const isvd_int_t ldvt,
                              const mpi_int_t ut_root, const mpi_int_t vt_root);

/* =========================================================================================
 * FP8 sketch, the "e" family (no counterpart in the reference).  A is held as OCP e4m3 codes
 * (e4m3fn: bias 7, largest finite value 448, 0x7F / 0xFF NaN) with one FP32 scale per
 * 128 x 128 block:  A[i, j] = e4m3(a8[i * lda8 + j]) * scale[(i / 128) * ldscale + j / 128].
 *   a8      row-major m x n (the n index contiguous), lda8 >= n, lda8 % 16 == 0, 16-byte
 *           aligned; rows are read and written in whole 16-byte chunks, so the buffer must
 *           span (m - 1) * lda8 + roundup(n, 16) bytes (m * lda8 always suffices);
 *   scale   ceil(m / 128) x ceil(n / 128), row-major, ldscale >= ceil(n / 128); every scale
 *           a power of two in [2^-127, 2^127] (applied on the matrix cores as an exact E8M0
 *           exponent; any other value fails the call);
 *   omega   Omega (n x Nl) as unscaled e4m3 codes, column c at omega + c * ldomega (the n
 *           index contiguous), ldomega >= n, ldomega % 16 == 0; spans
 *           (Nl - 1) * ldomega + roundup(n, 16) bytes (Nl * ldomega always suffices), and the
 *           codes of rows n .. roundup(n, 16) - 1 must be zero.
 * The seeded forms draw Omega from the FP32 stream of isvd_sIsvd, rounded to e4m3.
 * isvd_eIsvd sketches from (a8, scale) on FP8 matrix cores (FP32 accumulation) and runs the
 * remaining stages as isvd_bIsvd on the BF16 copy a (orthogonalization, integration, and the
 * FP32 Rayleigh-Ritz postprocessing Z^T = Qbar^T A against a).
 * ========================================================================================= */
typedef uint8_t isvd_e_val_t;
/* a8, scale := e4m3 codes and block scales of the FP32 matrix a (ordera 'C' or 'R'): each
 * scale is the smallest power of two s with max |A_block| <= 448 s (1 for a zero block), each
 * code A / s rounded to nearest even.  Codes of columns n .. roundup(n, 16) - 1 are zero. */
void isvd_eQuantize(const char ordera, const isvd_int_t m, const isvd_int_t n, const float *a,
                    const isvd_int_t lda, isvd_e_val_t *a8, const isvd_int_t lda8, float *scale,
                    const isvd_int_t ldscale);
void isvd_eSketchGaussianProjection(const isvd_Param param, const float *argv,
                                    const isvd_int_t argc, float *retv, const isvd_int_t retc,
                                    const char dista, const isvd_e_val_t *a8,
                                    const isvd_int_t lda8, const float *scale,
                                    const isvd_int_t ldscale, float *yst,
                                    const isvd_int_t ldyst, const isvd_int_t seed,
                                    const mpi_int_t mpi_root);
void isvd_eSketchGaussianProjectionOmega(const isvd_Param param, const char dista,
                                         const isvd_e_val_t *a8, const isvd_int_t lda8,
                                         const float *scale, const isvd_int_t ldscale,
                                         const isvd_e_val_t *omega, const isvd_int_t ldomega,
                                         float *yst, const isvd_int_t ldyst);
void isvd_eIsvd(const isvd_Param param, const char *alg_s, const char *alg_o, const char *alg_i,
                const char *alg_p, const float *argv[4], const isvd_int_t argc[4],
                float *retv[4], const isvd_int_t retc[4], double time[4], FILE *stream,
                const char dista, const isvd_e_val_t *a8, const isvd_int_t lda8,
                const float *scale, const isvd_int_t ldscale, const char ordera,
                const isvd_b_val_t *a, const isvd_int_t lda, float *s, float *ut,
                const isvd_int_t ldut, float *vt, const isvd_int_t ldvt, const isvd_int_t seed,
                const mpi_int_t ut_root, const mpi_int_t vt_root);

#if defined(__cplusplus)
}
#endif

#endif /* ISVD_MI355X_ISVD_H_ */
