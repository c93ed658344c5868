// Adjoint of the rank-k truncated SVD A ~ U_k S_k V_k^T returned by the driver (no counterpart
// in the reference).  Given dL/dS (k), dL/dU (m x k) and dL/dV (n x k), it forms dL/dA (m x n)
// by the SVD backward over all l computed singular triplets (the output gradients padded with
// zeros from k to l):
//   J = U^T gU, K = V^T gV (l x l),  F_ij = 1 / (s_j^2 - s_i^2) (i != j, gap-guarded),
//   Mid = (F o (J - J^T)) S + S (F o (K - K^T)) + diag(gS),
//   dA = U Mid V^T + (I - U U^T) gU S^-1 V^T + U S^-1 gV^T (I - V V^T).
// This is the derivative of the top-k singular triplets when A has rank <= l and the l
// triplets are exact (the complement of span(U) then carries no singular value); otherwise
// the projection terms stand in for the coupling to the discarded triplets.
// U and V are not stored by the forward pass: they are recomputed here from Qbar by the same
// postprocessing (same kernels, same order of operations), so their signs match the forward
// outputs the gradients refer to.
#include "stages.hpp"

#include <cmath>

#include "la.hpp"

namespace isvd {

namespace {

int grid_for(i64 n) {
  i64 b = (n + 255) / 256;
  return (int)(b < 1 ? 1 : (b > 8192 ? 8192 : b));
}

// Mid (l x l) from J, K (l x l, only the first k columns nonzero), s (l) and gS (k)
template <typename T>
__global__ void adjoint_mid_kernel(int l, int k, const T *J, const T *K, const T *s, const T *gs,
                                   T delta, T *Mid) {
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < l * l; e += gridDim.x * blockDim.x) {
    const int i = e % l, j = e / l;
    T v = 0;
    if (i == j) {
      v = (i < k) ? gs[i] : T(0);
    } else {
      const T si = s[i], sj = s[j];
      const T gap = sj * sj - si * si;
      const T scale = (si * si > sj * sj) ? si * si : sj * sj;
      if (fabs(gap) > delta * scale) {
        const T f = T(1) / gap;  // F_ij
        const T skj = J[i + j * l] - J[j + i * l];
        const T skk = K[i + j * l] - K[j + i * l];
        v = f * skj * sj + si * f * skk;
      }
    }
    Mid[e] = v;
  }
}

// rows r0 .. r0+k-1 of P (ld ldp) := diag(1/s) G (k x cols, ld ldg); zero singular values
// contribute nothing
template <typename T>
__global__ void rowscale_inv_kernel(int k, i64 cols, const T *s, const T *G, i64 ldg, T *P,
                                    i64 ldp, T beta) {
  for (i64 e = blockIdx.x * (i64)blockDim.x + threadIdx.x; e < k * cols;
       e += (i64)gridDim.x * blockDim.x) {
    const int i = (int)(e % k);
    const i64 c = e / k;
    const T inv = (s[i] != T(0)) ? T(1) / s[i] : T(0);
    T v = inv * G[i + c * ldg];
    if (beta != T(0)) v += beta * P[i + c * ldp];
    P[i + c * ldp] = v;
  }
}

}  // namespace

template <typename T>
void stage_isvd_backward(const Dims &d, bool tsqr, char ordera, const T *a, i64 lda,
                         const T *qt, i64 ldqt, const T *gs, const T *gut, i64 ldgut,
                         const T *gvt, i64 ldgvt, T *da, i64 ldda) {
  const i64 m = d.m, n = d.n, k = d.k, l = d.l, kl = l + k;
  // ---- U, S, V (all l triplets) recomputed from Qbar by the forward's postprocessing
  Dims dl = d;
  dl.k = l;
  DBuf<T> s(l), pt(kl * m), rt(kl * n);
  {
    DBuf<T> ut_l(l * m), vt_l(l * n);
    if (tsqr)
      stage_post_tsqr<T>(dl, ordera, a, lda, qt, ldqt, s.get(), ut_l.get(), l, vt_l.get(), l,
                         true, true);
    else
      stage_post_gramian<T>(dl, ordera, a, lda, qt, ldqt, s.get(), ut_l.get(), l, vt_l.get(), l,
                            true, true);
    // Pt = [U^T ; S_k^-1 Gu^T]  ((l + k) x m),  Rt = [R1 ; V_k^T]  ((l + k) x n)
    omatcopy<T>('N', l, m, T(1), ut_l.get(), l, pt.get(), kl);
    omatcopy<T>('N', k, n, T(1), vt_l.get(), l, rt.get() + l, kl);
    DBuf<T> j(l * l), kk(l * l), mid(l * l), gu(k * m), gv(k * n);
    // J[:, :k] = U^T gU, K[:, :k] = V^T gV (the remaining columns are zero)
    fill<T>(l, l, T(0), j.get(), l);
    fill<T>(l, l, T(0), kk.get(), l);
    gemm<T>('N', 'T', l, k, m, T(1), ut_l.get(), l, gut, ldgut, T(0), j.get(), l);
    gemm<T>('N', 'T', l, k, n, T(1), vt_l.get(), l, gvt, ldgvt, T(0), kk.get(), l);
    const T delta = (sizeof(T) == 8) ? T(1e-12) : T(1e-6);
    hipLaunchKernelGGL(adjoint_mid_kernel<T>, dim3(grid_for(l * l)), dim3(256), 0, ctx().stream,
                       (int)l, (int)k, (const T *)j.get(), (const T *)kk.get(),
                       (const T *)s.get(), gs, delta, mid.get());
    LAUNCH_CHECK();
    // Gu^T = gU^T - J[:, :k]^T U^T  (k x m);  Gv^T = gV^T - K[:, :k]^T V^T  (k x n)
    omatcopy<T>('N', k, m, T(1), gut, ldgut, gu.get(), k);
    gemm<T>('T', 'N', k, m, l, T(-1), j.get(), l, ut_l.get(), l, T(1), gu.get(), k);
    omatcopy<T>('N', k, n, T(1), gvt, ldgvt, gv.get(), k);
    gemm<T>('T', 'N', k, n, l, T(-1), kk.get(), l, vt_l.get(), l, T(1), gv.get(), k);
    // Pt rows l..l+k-1 := S_k^-1 Gu^T
    hipLaunchKernelGGL(rowscale_inv_kernel<T>, dim3(grid_for(k * m)), dim3(256), 0, ctx().stream,
                       (int)k, m, (const T *)s.get(), (const T *)gu.get(), k, pt.get() + l, kl,
                       T(0));
    LAUNCH_CHECK();
    // R1 = Mid V^T + [S_k^-1 Gv^T ; 0]  (l x n) into Rt rows 0..l-1
    gemm<T>('N', 'N', l, n, l, T(1), mid.get(), l, vt_l.get(), l, T(0), rt.get(), kl);
    hipLaunchKernelGGL(rowscale_inv_kernel<T>, dim3(grid_for(k * n)), dim3(256), 0, ctx().stream,
                       (int)k, n, (const T *)s.get(), (const T *)gv.get(), k, rt.get(), kl, T(1));
    LAUNCH_CHECK();
  }
  // dA = Pt^T Rt  (m x n; for a row-major A the transpose n x m is formed instead)
  if (ordera == 'C')
    gemm<T>('T', 'N', m, n, kl, T(1), pt.get(), kl, rt.get(), kl, T(0), da, ldda);
  else
    gemm<T>('T', 'N', n, m, kl, T(1), rt.get(), kl, pt.get(), kl, T(0), da, ldda);
}

template void stage_isvd_backward<double>(const Dims &, bool, char, const double *, i64,
                                          const double *, i64, const double *, const double *,
                                          i64, const double *, i64, double *, i64);
template void stage_isvd_backward<float>(const Dims &, bool, char, const float *, i64,
                                         const float *, i64, const float *, const float *, i64,
                                         const float *, i64, float *, i64);

}  // namespace isvd
