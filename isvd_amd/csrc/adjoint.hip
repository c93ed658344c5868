// Adjoint of the rank-k truncated SVD A ~ U_k S_k V_k^T returned by the driver (no counterpart
// in the reference).  Given dL/dS (k), dL/dU (m x k) and dL/dV (n x k), it forms dL/dA (m x n)
// from the l singular triplets recomputed from Qbar (the output gradients padded with zeros
// from k to l) and the complement of their span:
//   J = U^T gU, K = V^T gV (l x l),  F_ij = 1 / (s_j^2 - s_i^2) (i != j, gap-guarded),
//   Mid = (F o (J - J^T)) S + S (F o (K - K^T)) + diag(gS),
//   dA = U Mid V^T + E V_k^T + U_k H^T,
//   E = X S_k + P A Y,  H = A^T X + Y S_k,  P = I - U U^T,  A_c = P A,
//   x_i = (s_i^2 I - A_c A_c^T)^-1 P gU_i,  y_i = (s_i^2 I - A_c^T A_c)^-1 (I - V V^T) gV_i.
// E and H carry the coupling of the top-k triplets to every discarded singular pair
// (sum over j > l of u_j (.) / (s_i^2 - s_j^2)), so that with exact top-l triplets this is
// the derivative of the top-k triplets of the dense SVD at any rank.  x_i and y_i come from
// the Neumann series s_i^-2 sum_p (A_c A_c^T / s_i^2)^p P gU_i (likewise for y), which
// converges at the rate (||A_c||_2 / s_i)^2 < 1 when the discarded spectrum lies below s_i;
// each column is summed until a term falls below the unit roundoff of its sum.  When A_c = 0
// (rank(A) <= l) the series stops at its first term and E = P gU S_k^-1,
// H = (I - V V^T) gV S_k^-1: the truncated-SVD backward, which is also what a column keeps
// when its series does not converge (no gap between s_i and the discarded part of A).
// U and V are not stored by the forward pass: they are recomputed here from Qbar by the same
// postprocessing (same kernels, same order of operations), so their signs match the forward
// outputs the gradients refer to.
#include "stages.hpp"

#include <cmath>
#include <cstdlib>
#include <vector>

#include "la.hpp"

namespace isvd {

namespace {

int grid_for(i64 n) {
  i64 b = (n + 255) / 256;
  return (int)(b < 1 ? 1 : (b > 8192 ? 8192 : b));
}

// ISVD_TRACE=1 prints the complement series' progress to stderr
bool trace_adjoint() {
  static const bool on = [] {
    const char *e = std::getenv("ISVD_TRACE");
    return e && e[0] == '1';
  }();
  return on;
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

// Rows of a k-row operand selected by a bit mask (k <= 128: bit i of word i / 64)
struct RowMask {
  unsigned long long w[2];
  __host__ __device__ bool has(int i) const { return (w[i >> 6] >> (i & 63)) & 1ull; }
};
inline RowMask all_rows(i64 k) {
  RowMask r{{0ull, 0ull}};
  for (int i = 0; i < k; ++i) r.w[i >> 6] |= 1ull << (i & 63);
  return r;
}

// P (k x cols, ld ldp) := diag(s^e) G (ld ldg) + beta P for e in {-2, -1, 0, 1} on the rows in
// `rows`; the other rows are left as they are (keep) or zeroed.  A zero singular value
// contributes nothing to the negative powers.
template <typename T>
__global__ void rowscale_kernel(int k, i64 cols, const T *s, int e, const T *G, i64 ldg, T *P,
                                i64 ldp, T beta, RowMask rows, int keep) {
  for (i64 q = blockIdx.x * (i64)blockDim.x + threadIdx.x; q < k * cols;
       q += (i64)gridDim.x * blockDim.x) {
    const int i = (int)(q % k);
    const i64 c = q / k;
    if (!rows.has(i)) {
      if (!keep) P[i + c * ldp] = T(0);
      continue;
    }
    const T si = s[i];
    T f = T(1);
    if (e == 1) f = si;
    else if (e == -1) f = (si != T(0)) ? T(1) / si : T(0);
    else if (e == -2) f = (si != T(0)) ? T(1) / (si * si) : T(0);
    T v = f * G[i + c * ldg];
    if (beta != T(0)) v += beta * P[i + c * ldp];
    P[i + c * ldp] = v;
  }
}

template <typename T>
void rowscale(i64 k, i64 cols, const T *s, int e, const T *G, i64 ldg, T *P, i64 ldp, T beta,
              RowMask rows, bool keep = true) {
  hipLaunchKernelGGL(rowscale_kernel<T>, dim3(grid_for(k * cols)), dim3(256), 0, ctx().stream,
                     (int)k, cols, s, e, G, ldg, P, ldp, beta, rows, keep ? 1 : 0);
  LAUNCH_CHECK();
}

template <typename T>
void rowscale(i64 k, i64 cols, const T *s, int e, const T *G, i64 ldg, T *P, i64 ldp, T beta) {
  rowscale<T>(k, cols, s, e, G, ldg, P, ldp, beta, all_rows(k));
}

// out[i] := sum over columns of P[i, c]^2 for the k <= 128 rows of P (ld ldp), in a fixed order
// (bitwise reproducible): workgroup b sweeps the columns of chunk b with k consecutive lanes per
// column (coalesced) and writes one partial sum per row; a second pass adds the chunks in order.
constexpr int kNormThreads = 256, kNormChunks = 128;

template <typename T>
__global__ __launch_bounds__(kNormThreads) void rownorm2_parts_kernel(int k, i64 cols, i64 chunk,
                                                                      const T *P, i64 ldp,
                                                                      T *part) {
  __shared__ T red[kNormThreads];
  const int t = threadIdx.x, lanes = kNormThreads / k;
  const i64 beg = (i64)blockIdx.x * chunk, end = (beg + chunk < cols) ? beg + chunk : cols;
  T s = 0;
  if (t < k * lanes) {
    const int i = t % k, lane = t / k;
    for (i64 c = beg + lane; c < end; c += lanes) {
      const T v = P[i + c * ldp];
      s += v * v;
    }
  }
  red[t] = s;
  __syncthreads();
  if (t < k) {
    T r = 0;
    for (int q = 0; q < lanes; ++q) r += red[t + q * k];
    part[(i64)blockIdx.x * k + t] = r;
  }
}

template <typename T>
__global__ __launch_bounds__(128) void rownorm2_final_kernel(int k, int nchunks, const T *part,
                                                             T *out) {
  const int i = threadIdx.x;
  if (i >= k) return;
  T r = 0;
  for (int b = 0; b < nchunks; ++b) r += part[(i64)b * k + i];
  out[i] = r;
}

template <typename T>
void rownorm2(int k, i64 cols, const T *P, i64 ldp, T *out, T *part) {
  i64 chunk = (cols + kNormChunks - 1) / kNormChunks;
  if (chunk < 1) chunk = 1;
  const int nchunks = (int)((cols + chunk - 1) / chunk);
  hipLaunchKernelGGL(rownorm2_parts_kernel<T>, dim3((unsigned)(nchunks > 0 ? nchunks : 1)),
                     dim3(kNormThreads), 0, ctx().stream, k, cols, chunk, P, ldp, part);
  LAUNCH_CHECK();
  hipLaunchKernelGGL(rownorm2_final_kernel<T>, dim3(1), dim3(128), 0, ctx().stream, k,
                     nchunks > 0 ? nchunks : 1, (const T *)part, out);
  LAUNCH_CHECK();
}

}  // namespace

template <typename T>
void stage_isvd_backward(const Dims &d, bool tsqr, char ordera, const T *a, i64 lda,
                         const T *qt, i64 ldqt, const T *gs, const T *gut, i64 ldgut,
                         const T *gvt, i64 ldgvt, T *da, i64 ldda) {
  const i64 m = d.m, n = d.n, k = d.k, l = d.l, kl = l + k;
  // products with A for r-row operands stored transposed (r x m and r x n):
  //   xa: W (r x n) := X^T A,  xat: Z (r x m) := W A^T
  auto xa = [&](i64 r, const T *x, i64 ldx, T *w, i64 ldw) {
    if (ordera == 'C') gemm<T>('N', 'N', r, n, m, T(1), x, ldx, a, lda, T(0), w, ldw);
    else gemm<T>('N', 'T', r, n, m, T(1), x, ldx, a, lda, T(0), w, ldw);
  };
  auto xat = [&](i64 r, const T *w, i64 ldw, T *z, i64 ldz) {
    if (ordera == 'C') gemm<T>('N', 'T', r, m, n, T(1), w, ldw, a, lda, T(0), z, ldz);
    else gemm<T>('N', 'N', r, m, n, T(1), w, ldw, a, lda, T(0), z, ldz);
  };
  // ---- U, S, V (all l triplets) recomputed from Qbar by the forward's postprocessing
  Dims dl = d;
  dl.k = l;
  DBuf<T> s(l), pt(kl * m), rt(kl * n), ut_l(l * m), vt_l(l * n);
  if (tsqr)
    stage_post_tsqr<T>(dl, ordera, a, lda, qt, ldqt, s.get(), ut_l.get(), l, vt_l.get(), l, true,
                       true);
  else
    stage_post_gramian<T>(dl, ordera, a, lda, qt, ldqt, s.get(), ut_l.get(), l, vt_l.get(), l,
                          true, true);
  // z (r x m, r <= 2k) := z - (z U) U^T: the component outside span(U)
  DBuf<T> cproj(2 * k * l);
  auto project = [&](i64 r, T *z, i64 ldz) {
    gemm<T>('N', 'T', r, l, m, T(1), z, ldz, ut_l.get(), l, T(0), cproj.get(), r);
    gemm<T>('N', 'N', r, m, l, T(-1), cproj.get(), r, ut_l.get(), l, T(1), z, ldz);
  };
  // Pt = [U^T ; E^T]  ((l + k) x m),  Rt = [Mid V^T + [H^T ; 0] ; V_k^T]  ((l + k) x n)
  omatcopy<T>('N', l, m, T(1), ut_l.get(), l, pt.get(), kl);
  omatcopy<T>('N', k, n, T(1), vt_l.get(), l, rt.get() + l, kl);
  // the series run half a step apart so that each step takes one product with A and one with
  // A^T, each over 2k rows: zz = [X_p ; (P A Y_p)^T] (2k x m) and ww = [(A^T X_p)^T ; Y_p]
  // (2k x n), ld 2k; xt, yt the sums; zt, wt scratch of k rows
  const i64 k2 = 2 * k;
  DBuf<T> xt(k * m), zt(k * m), yt(k * n), wt(k * n), zz(k2 * m), ww(k2 * n);
  {
    DBuf<T> j(l * l), kk(l * l), mid(l * l);
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
    // P gU (as its transpose, into zt) and (I - V V^T) gV (into wt)
    omatcopy<T>('N', k, m, T(1), gut, ldgut, zt.get(), k);
    gemm<T>('T', 'N', k, m, l, T(-1), j.get(), l, ut_l.get(), l, T(1), zt.get(), k);
    omatcopy<T>('N', k, n, T(1), gvt, ldgvt, wt.get(), k);
    gemm<T>('T', 'N', k, n, l, T(-1), kk.get(), l, vt_l.get(), l, T(1), wt.get(), k);
    // R1 = Mid V^T into Rt rows 0..l-1
    gemm<T>('N', 'N', l, n, l, T(1), mid.get(), l, vt_l.get(), l, T(0), rt.get(), kl);
  }
  // ---- the complement: X = sum_p X_p, X_0 = P gU S^-2, X_{p+1} = P A A^T X_p S^-2, and
  // Y = sum_p Y_p, Y_0 = (I - V V^T) gV S^-2, Y_{p+1} = A^T P A Y_p S^-2 (transposed storage).
  // Column i of either series converges at the rate (||A_c||_2 / s_i)^2.  Each column is summed
  // until its term falls below the unit roundoff of its sum (or for kMaxTerms terms); a column
  // whose terms stop shrinking has no gap below s_i (the discarded part of A reaches s_i, e.g.
  // a subspace from a sketch too small for the spectrum) and keeps its first term only: the
  // truncated-SVD backward, which treats the complement as carrying no singular value.
  DBuf<T> x0(k * m), y0(k * n);
  rowscale<T>(k, m, s.get(), -2, zt.get(), k, x0.get(), k, T(0));
  rowscale<T>(k, m, s.get(), 0, x0.get(), k, zz.get(), k2, T(0));
  rowscale<T>(k, m, s.get(), 0, x0.get(), k, xt.get(), k, T(0));
  rowscale<T>(k, n, s.get(), -2, wt.get(), k, y0.get(), k, T(0));
  rowscale<T>(k, n, s.get(), 0, y0.get(), k, yt.get(), k, T(0));
  // the first half step of the Y series: (P A Y_0)^T into the lower rows of zz
  xat(k, y0.get(), k, zz.get() + k, k2);
  project(k, zz.get() + k, k2);
  DBuf<T> nrm(4 * k), nrm_part(kNormChunks * k);
  std::vector<T> h(4 * k);
  const double eps = (sizeof(T) == 8) ? 0x1p-52 : 0x1p-23;
  constexpr int kMaxTerms = 300;
  RowMask ax = all_rows(k), ay = all_rows(k);
  std::vector<double> prev_x(k, INFINITY), prev_y(k, INFINITY);
  auto any = [](const RowMask &r) { return (r.w[0] | r.w[1]) != 0ull; };
  for (int it = 1; any(ax) || any(ay); ++it) {
    // ww := [X_p ; (P A Y_p)^T] A: the X series' first half step and the Y series' second;
    // Y_{p+1} = S^-2 (A^T P A Y_p) in place (rows of finished columns zeroed), Y += Y_{p+1}
    xa(k2, zz.get(), k2, ww.get(), k2);
    rowscale<T>(k, n, s.get(), -2, ww.get() + k, k2, ww.get() + k, k2, T(0), ay, false);
    rowscale<T>(k, n, s.get(), 0, ww.get() + k, k2, yt.get(), k, T(1), ay);
    // zz := P ww A^T: the X series' second half step and the Y series' first (for Y_{p+1});
    // X_{p+1} = S^-2 (P A A^T X_p) in place, X += X_{p+1}
    xat(k2, ww.get(), k2, zz.get(), k2);
    project(k2, zz.get(), k2);
    rowscale<T>(k, m, s.get(), -2, zz.get(), k2, zz.get(), k2, T(0), ax, false);
    rowscale<T>(k, m, s.get(), 0, zz.get(), k2, xt.get(), k, T(1), ax);
    rownorm2<T>((int)k, m, (const T *)zz.get(), k2, nrm.get(), nrm_part.get());
    rownorm2<T>((int)k, m, (const T *)xt.get(), k, nrm.get() + k, nrm_part.get());
    rownorm2<T>((int)k, n, (const T *)(ww.get() + k), k2, nrm.get() + 2 * k, nrm_part.get());
    rownorm2<T>((int)k, n, (const T *)yt.get(), k, nrm.get() + 3 * k, nrm_part.get());
    read_back(h.data(), nrm.get(), sizeof(T) * 4 * k);
    RowMask rx{{0ull, 0ull}}, ry{{0ull, 0ull}};  // columns reverting to their first term
    auto assess = [&](RowMask &act, RowMask &rev, std::vector<double> &prev, const T *term,
                     const T *sum) {
      for (int i = 0; i < (int)k; ++i) {
        if (!act.has(i)) continue;
        const double t = std::sqrt((double)term[i]), sm = std::sqrt((double)sum[i]);
        const bool done = t <= eps * sm || it >= kMaxTerms;
        const bool grows = !(t < prev[i]);
        prev[i] = t;
        if (done || grows) act.w[i >> 6] &= ~(1ull << (i & 63));
        if (grows && !done) rev.w[i >> 6] |= 1ull << (i & 63);
      }
    };
    assess(ax, rx, prev_x, h.data(), h.data() + k);
    assess(ay, ry, prev_y, h.data() + 2 * k, h.data() + 3 * k);
    if (any(rx)) rowscale<T>(k, m, s.get(), 0, x0.get(), k, xt.get(), k, T(0), rx);
    if (any(ry)) rowscale<T>(k, n, s.get(), 0, y0.get(), k, yt.get(), k, T(0), ry);
    if (trace_adjoint())
      std::fprintf(stderr, "isvd adjoint complement term %d: %d / %d columns still summed\n", it,
                   __builtin_popcountll(ax.w[0]) + __builtin_popcountll(ax.w[1]) +
                       __builtin_popcountll(ay.w[0]) + __builtin_popcountll(ay.w[1]),
                   (int)(2 * k));
  }
  // E^T = S X^T + (P A Y)^T into Pt rows l..l+k-1;  H^T = X^T A + S Y^T added to Rt rows 0..k-1
  xat(k, yt.get(), k, zt.get(), k);
  project(k, zt.get(), k);
  rowscale<T>(k, m, s.get(), 0, zt.get(), k, pt.get() + l, kl, T(0));
  rowscale<T>(k, m, s.get(), 1, xt.get(), k, pt.get() + l, kl, T(1));
  xa(k, xt.get(), k, wt.get(), k);
  rowscale<T>(k, n, s.get(), 1, yt.get(), k, wt.get(), k, T(1));
  rowscale<T>(k, n, s.get(), 0, wt.get(), k, rt.get(), kl, T(1));
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
