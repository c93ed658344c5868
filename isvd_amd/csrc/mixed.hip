// Mixed-precision iSVD stages (no counterpart in the reference; the FP64 stages they mirror are
// cited at each function).
#include "mixed.hpp"

#include "la.hpp"
#include "mx.hpp"

namespace isvd {

namespace {

// U_b[:, j] /= sqrt(s_b[j])  (s: l values per problem, U: l x l per problem)
__global__ void colscale_rsqrt_kernel(int l, const double *s, double *U) {
  const i64 b = blockIdx.y;
  const double *sb = s + b * l;
  double *Ub = U + b * (i64)l * l;
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < l * l; e += gridDim.x * blockDim.x)
    Ub[e] /= sqrt(sb[e / l]);
}

int grid_for(i64 n) {
  i64 b = (n + 255) / 256;
  return (int)(b < 1 ? 1 : (b > 8192 ? 8192 : b));
}

// M := Qs^T Qs (Nl x Nl) from the FP32 bases by split-BF16 x3 products: Qs = H + L (two BF16
// terms), M = H^T H + L^T H + H^T L.  The split is materialized as S = [H^T ; L^T] (2Nl x m)
// and the two products [H^T H ; L^T H] = S H are one plain BF16 GEMM (FP32 accumulation); the
// combination symmetrizes exactly.
void gram_split(const Dims &d, const float *yst, i64 ldyst, float *m32) {
  const i64 Nl = d.Nl, m = d.m;
  DBuf<uint16_t> S(2 * Nl * m);
  split_bf16(Nl, m, 2, yst, ldyst, S.get(), 2 * Nl);
  DBuf<float> C(2 * Nl * Nl);
  // H^T H is symmetric: its tiles below the diagonal are skipped (gram_combine reads its upper
  // triangle)
  gemm_bf16_upper('N', 'T', 2 * Nl, Nl, m, Nl, S.get(), 2 * Nl, S.get(), 2 * Nl, C.get(), 2 * Nl);
  S.release();
  gram_combine(Nl, C.get(), 2 * Nl, m32, Nl);
}

}  // namespace

void gaussian_omega_b(i64 rows, i64 cols, uint16_t *omegat, i64 ld, unsigned long long seed) {
  ISVD_CHECK(ld == rows, "gaussian_omega_b: ld must equal rows");
  DBuf<float> tmp(rows * cols);
  gaussian_omega<float>(rows, cols, tmp.get(), rows, seed);
  f32_to_bf16(rows * cols, tmp.get(), omegat);
}

// sketchBlockRow (src/lib/libisvd/core/stage/@x@_sketch.h:136-139) in BF16: a plain BF16
// product with FP32 accumulation and output
void stage_sketch_b(const Dims &d, char ordera, const uint16_t *a, i64 lda,
                    const uint16_t *omegat, i64 ldo, float *yst, i64 ldyst) {
  // op(A)^T: a column-major A (m x n) is transposed, a row-major A is already A^T (n x m)
  gemm_bf16('N', ordera == 'C' ? 'T' : 'N', d.Nl, d.m, d.n, 1.f, omegat, ldo, a, lda, 0.f,
                yst, ldyst);
}

// @x@_orthogonalize_gramian.c:81-98, twice
void stage_orth_gramian_b(const Dims &d, float *yst, i64 ldyst, bool cholqr) {
  const int mode = cholqr_mode();
  if (((cholqr && mode == 0) || mode == 2) &&
      cholqr2_rowmajor<float>(d.m, d.l, yst, ldyst, d.l, d.N, nullptr, 0, 0))
    return;
  const i64 l = d.l, N = d.N, m = d.m, Nl = d.Nl;
  DBuf<float> w32(l * l * N), u32(l * l * N), other(Nl * m);
  DBuf<double> w64(l * l * N), u64(l * l * N), s64(l * N);
  // pass 1 reads yst and writes `other`, pass 2 reads `other` and writes yst: no copies
  for (int pass = 0; pass < 2; ++pass) {
    const float *src = pass == 0 ? yst : other.get();
    const i64 lds = pass == 0 ? ldyst : Nl;
    float *dst = pass == 0 ? other.get() : yst;
    const i64 ldd = pass == 0 ? Nl : ldyst;
    // W_i := Y_i^T Y_i
    gemm<float>('N', 'T', l, l, m, 1.f, src, lds, src, lds, 0.f, w32.get(), l, N, l, l, l * l);
    convert<double, float>(l, l * N, w32.get(), l, w64.get(), l);
    // W_i = U_i diag(s_i) U_i^T; U_i := U_i diag(s_i)^{-1/2}
    svd_small<double>(l, w64.get(), l, l * l, s64.get(), l, u64.get(), l, l * l, nullptr, 0, 0, N);
    hipLaunchKernelGGL(colscale_rsqrt_kernel, dim3(grid_for(l * l), (unsigned)N), dim3(256), 0,
                       ctx().stream, (int)l, (const double *)s64.get(), u64.get());
    LAUNCH_CHECK();
    convert<float, double>(l, l * N, u64.get(), l, u32.get(), l);
    // Q_i^T := U_i^T Y_i^T
    gemm<float>('T', 'N', l, m, l, 1.f, u32.get(), l, src, lds, 0.f, dst, ldd, N, l * l, l, l);
  }
}

// @x@_orthogonalize_tall_skinny_qr.c:127-143 in FP32
void stage_orth_tsqr_b(const Dims &d, float *yst, i64 ldyst, bool cholqr) {
  stage_orth_tsqr<float>(d, yst, ldyst, cholqr);
}

// @x@_integrate_kolmogorov_nagumo.c:155-253.  The iteration runs in FP32 on the FP32-accurate
// M (split-BF16 x3 products): FP64 arithmetic on it would carry digits M does not have.
void stage_integrate_kn_b(const Dims &d, const KnArgs &args, const float *yst, i64 ldyst,
                          float *qt, i64 ldqt, i64 *iter, double *err) {
  const i64 l = d.l, Nl = d.Nl;
  DBuf<float> x32(Nl * l);
  // M as its two BF16 terms: the FP32 loop iterates with M'' = Mh + Ml and takes its products
  // with M'' on the BF16 matrix cores
  DBuf<uint16_t> S(2 * Nl * Nl), bop(2 * Nl * 3 * l);
  SplitSym ms;
  ms.n = Nl;
  ms.S = S.get();
  ms.bop = bop.get();
  DBuf<float> gw(split_sym_gw_size(Nl, l, ms.max_split));
  ms.gw = gw.get();
  {
    DBuf<float> m32(Nl * Nl);
    gram_split(d, yst, ldyst, m32.get());
    split_bf16(Nl, Nl, 2, m32.get(), Nl, S.get(), 2 * Nl);
  }
  integrate_kn_coords<float>(d, args, nullptr, x32.get(), iter, err, &ms);
  gemm<float>('T', 'N', l, d.m, Nl, 1.f, x32.get(), Nl, yst, ldyst, 0.f, qt, ldqt);
}

// @x@_integrate_wen_yin.c:201-362
void stage_integrate_wy_b(const Dims &d, const WyArgs &args, const float *yst, i64 ldyst,
                          float *qt, i64 ldqt, i64 *iter, double *mu) {
  const i64 l = d.l, Nl = d.Nl;
  DBuf<float> m32(Nl * Nl), x32(Nl * l);
  DBuf<double> m64(Nl * Nl), x64(Nl * l);
  gram_split(d, yst, ldyst, m32.get());
  convert<double, float>(Nl, Nl, m32.get(), Nl, m64.get(), Nl);
  m32.release();
  integrate_wy_coords<double>(d, args, m64.get(), x64.get(), iter, mu);
  convert<float, double>(Nl, l, x64.get(), Nl, x32.get(), Nl);
  gemm<float>('T', 'N', l, d.m, Nl, 1.f, x32.get(), Nl, yst, ldyst, 0.f, qt, ldqt);
}

// projectBlockRow (@x@_postprocess.h:154-157) + @x@_postprocess_gramian.c:135-185 /
// @x@_postprocess_tall_skinny_qr.c:184-296
void stage_post_b(const Dims &d, bool tsqr, char ordera, const uint16_t *a, i64 lda,
                  const float *qt, i64 ldqt, float *s, float *ut, i64 ldut, float *vt, i64 ldvt,
                  bool want_u, bool want_v) {
  const i64 l = d.l, m = d.m, n = d.n, k = d.k;
  DBuf<float> zt32(l * n);
  {
    // Qbar^T = h0 + h1 + h2 (three BF16 terms, exact to FP32) stacked as 3l x m; one plain
    // BF16 GEMM against the BF16 A reads A once, then the three l x n blocks are summed
    DBuf<uint16_t> q3(3 * l * m);
    split_bf16(l, m, 3, qt, ldqt, q3.get(), 3 * l);
    DBuf<float> c3(3 * l * n);
    gemm_bf16('N', ordera == 'C' ? 'N' : 'T', 3 * l, n, m, 1.f, q3.get(), 3 * l, a, lda,
                  0.f, c3.get(), 3 * l);
    sum_blocks(l, n, 3, c3.get(), 3 * l, zt32.get(), l);
  }
  DBuf<double> zt64(l * n), qt64(l * m), s64(l), ut64(want_u ? l * m : 0),
      vt64(want_v ? l * n : 0);
  convert<double, float>(l, n, zt32.get(), l, zt64.get(), l);
  zt32.release();
  convert<double, float>(l, m, qt, ldqt, qt64.get(), l);
  if (tsqr)
    post_tsqr_z<double>(d, qt64.get(), l, zt64.get(), s64.get(), ut64.get(), l, vt64.get(), l,
                        want_u, want_v);
  else
    post_gramian_z<double>(d, qt64.get(), l, zt64.get(), s64.get(), ut64.get(), l, vt64.get(),
                           l, want_u, want_v);
  convert<float, double>(l, 1, s64.get(), l, s, l);
  if (want_u) convert<float, double>(k, m, ut64.get(), l, ut, ldut);
  if (want_v) convert<float, double>(k, n, vt64.get(), l, vt, ldvt);
}

}  // namespace isvd
