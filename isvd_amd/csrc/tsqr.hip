// Householder tall-skinny QR (TSQR) for the orthogonalization and postprocessing stages.
//
// The reference factors each sketch with LAPACK gelqf/orglq and, across MPI ranks, combines
// the triangular factors with a butterfly of stacked-R factorizations
// (src/lib/libisvd/core/stage/@x@_orthogonalize_tall_skinny_qr.c:127-200,
//  src/lib/libisvd/core/stage/@x@_postprocess_tall_skinny_qr.c:184-246).  On one GPU the same
// tree runs over row panels of each sketch:
//   level 0   : every row panel (<= 256 rows) is factored by one workgroup with the panel held
//               in registers (column c = tid % LC, rows rg + NRG*i), Householder vectors
//               written back in place, R written to the next level's buffer;
//   level k>0 : two consecutive R factors form a stacked 2l x l panel that is factored the
//               same way, until one R per sketch remains;
//   formation : top-down, each node applies its reflectors to [C; 0] (C = I at the root) and
//               hands the two l x l halves to its children; the leaves produce Q in place.
// Reflectors follow LAPACK dlarfg (beta = -sign(alpha) * ||x||, tau = (beta - alpha) / beta).
#include "la.hpp"

namespace isvd {

namespace {

constexpr int kNT = 512;    // threads per panel (2 waves per SIMD -> 256 VGPRs per lane)
constexpr int kMaxR = 64;   // rows per thread
constexpr int kHMax = 256;  // rows per panel (LDS broadcast buffer size)

struct PanelGeom {
  i64 m;        // rows of the level-0 matrix
  int nleaf;    // leaves per batch entry
  int base_h;   // leaf height = base_h (+1 for the first `extra` leaves)
  int extra;
};

__device__ __forceinline__ void leaf_rows(const PanelGeom &g, int j, i64 &r0, int &h) {
  r0 = (i64)j * g.base_h + (j < g.extra ? j : g.extra);
  h = g.base_h + (j < g.extra ? 1 : 0);
}

// Panel source description for one launch:
//  leaf:     base = Y + b*sY + r0*ldy, ld = ldy, h from the leaf partition
//  internal: base = Rin + (b*nchild + 2j)*l*l, ld = l, h = l * (children count)
struct LevelArgs {
  int l, LC;
  bool leaf;
  PanelGeom geom;
  // panel storage (in/out for factor; V source for apply)
  void *Y; i64 ldy, sY;
  void *Rin; int nchild;    // internal: stacked children factors
  int nnode;                // nodes at this level per batch entry
  void *tau;                // nnode * l per batch entry
  void *Rout;               // factor: R of each node (row-major l x l)
  void *Cin;                // apply: C of each node (row-major l x l)
  void *Cout;               // apply (internal): children C (row-major), same layout as Rin
};

template <typename T>
__device__ __forceinline__ void panel_of(const LevelArgs &a, int b, int j, T *&base, i64 &ld,
                                         int &h) {
  if (a.leaf) {
    i64 r0;
    leaf_rows(a.geom, j, r0, h);
    base = reinterpret_cast<T *>(a.Y) + (i64)b * a.sY + r0 * a.ldy;
    ld = a.ldy;
  } else {
    const int c0 = 2 * j;
    const int nc = (c0 + 1 < a.nchild) ? 2 : 1;
    h = nc * a.l;
    base = reinterpret_cast<T *>(a.Rin) + ((i64)b * a.nchild + c0) * a.l * a.l;
    ld = a.l;
  }
}

template <typename T>
__global__ __launch_bounds__(kNT) void hh_factor_kernel(LevelArgs a) {
  __shared__ T vsh[kHMax];
  __shared__ T wpart[kNT];
  __shared__ T red[kNT];
  __shared__ T alpha_sh;

  const int j_node = blockIdx.x, b = blockIdx.y;
  const int l = a.l, LC = a.LC, NRG = kNT / LC;
  const int tid = threadIdx.x, c = tid % LC, rg = tid / LC;
  T *base;
  i64 ld;
  int h;
  panel_of<T>(a, b, j_node, base, ld, h);

  // rows of this thread: r = rg + NRG * i for i in [0, ihi); "rows >= j" is i >= first_i(j)
  const int ihi = (h > rg) ? (h - rg + NRG - 1) / NRG : 0;
  auto first_i = [&](int j) { return (j <= rg) ? 0 : (j - rg + NRG - 1) / NRG; };
  T x[kMaxR];
#pragma unroll
  for (int i = 0; i < kMaxR; ++i) {
    const int r = rg + NRG * i;
    x[i] = (i < ihi && c < l) ? base[(i64)r * ld + c] : T(0);
  }
  const int kmax = h < l ? h : l;
  T *tau_out = reinterpret_cast<T *>(a.tau) + ((i64)b * a.nnode + j_node) * l;

  for (int j = 0; j < kmax; ++j) {
    const int i_ge = first_i(j), i_gt = first_i(j + 1);
    const bool own_diag = (rg == j % NRG);
    const int i_diag = j / NRG;
    // (1) norm of column j below the diagonal, and the diagonal entry
    if (c == j) {
      T part = 0;
#pragma unroll
      for (int i = 0; i < kMaxR; ++i) {
        if (i >= i_gt && i < ihi) part += x[i] * x[i];
        if (own_diag && i == i_diag) alpha_sh = x[i];
      }
      red[rg] = part;
    }
    __syncthreads();
    T sigma2 = 0;
    for (int q = 0; q < NRG; ++q) sigma2 += red[q];
    const T alpha = alpha_sh;
    T beta, tau, scale;
    if (sigma2 == T(0)) {
      beta = alpha; tau = T(0); scale = T(0);
    } else {
      const T nrm = sqrt(alpha * alpha + sigma2);
      beta = (alpha >= T(0)) ? -nrm : nrm;
      tau = (beta - alpha) / beta;
      scale = T(1) / (alpha - beta);
    }
    // (2) form v in column j, broadcast through LDS
    if (c == j) {
#pragma unroll
      for (int i = 0; i < kMaxR; ++i) {
        const int r = rg + NRG * i;
        if (i >= i_gt && i < ihi) { x[i] *= scale; vsh[r] = x[i]; }
        if (own_diag && i == i_diag) { x[i] = beta; vsh[r] = T(1); }
      }
      if (tid == j) tau_out[j] = tau;  // thread (c == j, rg == 0)
    }
    __syncthreads();
    // (3) w_c = v^T A[:, c] for c > j
    if (c > j && c < l) {
      T part = 0;
#pragma unroll
      for (int i = 0; i < kMaxR; ++i)
        if (i >= i_ge && i < ihi) part += vsh[rg + NRG * i] * x[i];
      wpart[rg * LC + c] = part;
    }
    __syncthreads();
    // (4) rank-1 update of the trailing columns
    if (c > j && c < l && tau != T(0)) {
      T w = 0;
      for (int q = 0; q < NRG; ++q) w += wpart[q * LC + c];
      w *= tau;
#pragma unroll
      for (int i = 0; i < kMaxR; ++i)
        if (i >= i_ge && i < ihi) x[i] -= vsh[rg + NRG * i] * w;
    }
  }
  // write V (below the diagonal) + R (on and above) back in place; R to the next level
#pragma unroll
  for (int i = 0; i < kMaxR; ++i) {
    const int r = rg + NRG * i;
    if (i < ihi && c < l) base[(i64)r * ld + c] = x[i];
  }
  T *R = reinterpret_cast<T *>(a.Rout) + ((i64)b * a.nnode + j_node) * l * l;
#pragma unroll
  for (int i = 0; i < kMaxR; ++i) {
    const int r = rg + NRG * i;
    if (r < l && c < l) R[r * l + c] = (r <= c && r < kmax) ? x[i] : T(0);
  }
  // rows of R beyond the panel height (h < l) are zero
  for (int e = tid; e < l * l; e += kNT) {
    const int r = e / l;
    if (r >= h) R[e] = T(0);
  }
}

// Explicit Q: X := H_0 H_1 ... H_{k-1} [C; 0].  The reflectors are applied last-to-first; the
// vector of step j is streamed from the panel into LDS one step ahead (every lane loads one
// entry), so only X lives in registers.
template <typename T>
__global__ __launch_bounds__(kNT) void hh_apply_kernel(LevelArgs a) {
  __shared__ T vsh[2][kHMax];
  __shared__ T wpart[kNT];

  const int j_node = blockIdx.x, b = blockIdx.y;
  const int l = a.l, LC = a.LC, NRG = kNT / LC;
  const int tid = threadIdx.x, c = tid % LC, rg = tid / LC;
  T *base;
  i64 ld;
  int h;
  panel_of<T>(a, b, j_node, base, ld, h);
  const int kmax = h < l ? h : l;

  T x[kMaxR];
  const T *C = reinterpret_cast<const T *>(a.Cin) + ((i64)b * a.nnode + j_node) * l * l;
#pragma unroll
  for (int i = 0; i < kMaxR; ++i) {
    const int r = rg + NRG * i;
    x[i] = (r < kmax && c < l) ? C[r * l + c] : T(0);
  }
  const T *tau_in = reinterpret_cast<const T *>(a.tau) + ((i64)b * a.nnode + j_node) * l;
  // v_j[r] = 0 (r < j), 1 (r == j), V[r, j] (r > j)
  auto vload = [&](int j, int r) -> T {
    return (r < j) ? T(0) : (r == j ? T(1) : base[(i64)r * ld + j]);
  };
  if (kmax > 0 && tid < h) vsh[(kmax - 1) & 1][tid] = vload(kmax - 1, tid);
  __syncthreads();
  for (int j = kmax - 1; j >= 0; --j) {
    const T *vs = vsh[j & 1];
    T vnext = T(0);
    if (j > 0 && tid < h) vnext = vload(j - 1, tid);  // in flight during this step
    const T tau = tau_in[j];
    if (c < l) {
      T part = 0;
#pragma unroll
      for (int i = 0; i < kMaxR; ++i) {
        const int r = rg + NRG * i;
        if (r >= j && r < h) part += vs[r] * x[i];
      }
      wpart[rg * LC + c] = part;
    }
    __syncthreads();
    if (c < l && tau != T(0)) {
      T w = 0;
      for (int q = 0; q < NRG; ++q) w += wpart[q * LC + c];
      w *= tau;
#pragma unroll
      for (int i = 0; i < kMaxR; ++i) {
        const int r = rg + NRG * i;
        if (r >= j && r < h) x[i] -= vs[r] * w;
      }
    }
    if (j > 0 && tid < h) vsh[(j - 1) & 1][tid] = vnext;
    __syncthreads();
  }
  // output: leaf -> the panel itself; internal -> children C blocks (same layout as Rin)
  T *out;
  i64 ldo;
  if (a.leaf) {
    out = base; ldo = ld;
  } else {
    out = reinterpret_cast<T *>(a.Cout) + ((i64)b * a.nchild + 2 * j_node) * l * l;
    ldo = l;
  }
#pragma unroll
  for (int i = 0; i < kMaxR; ++i) {
    const int r = rg + NRG * i;
    if (r < h && c < l) out[(i64)r * ldo + c] = x[i];
  }
}

template <typename T>
__global__ void eye_rowmajor_kernel(int l, i64 count, T *C) {
  for (i64 e = blockIdx.x * (i64)blockDim.x + threadIdx.x; e < count * l * l;
       e += (i64)gridDim.x * blockDim.x) {
    const i64 w = e % ((i64)l * l);
    C[e] = (w / l == w % l) ? T(1) : T(0);
  }
}

// R (row-major l x l per batch entry) -> col-major output with ldr
template <typename T>
__global__ void copy_r_kernel(int l, const T *Rrm, T *R, i64 ldr, i64 sR) {
  const int b = blockIdx.y;
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < l * l; e += gridDim.x * blockDim.x) {
    const int r = e / l, c = e % l;
    R[(i64)b * sR + r + (i64)c * ldr] = Rrm[(i64)b * l * l + e];
  }
}

}  // namespace

template <typename T>
void tsqr_rowmajor(i64 m, i64 l, T *Y, i64 ldy, i64 sY, i64 batch, T *R, i64 ldr, i64 sR) {
  ISVD_CHECK(l >= 1 && l <= 128, "TSQR supports 1 <= l <= 128 (got %lld)", l);
  ISVD_CHECK(m >= l, "TSQR needs m >= l (m = %lld, l = %lld)", m, l);
  ISVD_CHECK(batch >= 1 && batch <= 65535, "bad batch %lld", batch);
  int LC = 1;
  while (LC < l) LC *= 2;
  const int NRG = kNT / LC;
  i64 hmax = (i64)NRG * kMaxR;  // rows one workgroup holds
  if (hmax > kHMax) hmax = kHMax;
  ISVD_CHECK(hmax >= 2 * l, "internal: panel capacity");
  PanelGeom geom{};
  geom.m = m;
  i64 nleaf = (m + hmax - 1) / hmax;
  ISVD_CHECK(nleaf <= 2147483647LL, "too many panels");
  geom.nleaf = (int)nleaf;
  geom.base_h = (int)(m / nleaf);
  geom.extra = (int)(m % nleaf);

  // level sizes
  std::vector<int> nn;  // nodes per level (level 0 = leaves)
  nn.push_back((int)nleaf);
  while (nn.back() > 1) nn.push_back((nn.back() + 1) / 2);
  const int K = (int)nn.size() - 1;  // number of internal levels

  // buffers: Rbuf[k] for k = 1..K+1 holds R of level k-1 nodes (nn[k-1] entries)
  const i64 ll = l * l;
  std::vector<DBuf<T>> Rbuf(K + 2), Cbuf(K + 2), taus(K + 1);
  for (int k = 1; k <= K + 1; ++k) {
    Rbuf[k].alloc(batch * nn[k - 1] * ll);
    Cbuf[k].alloc(batch * nn[k - 1] * ll);
  }
  for (int k = 0; k <= K; ++k) taus[k].alloc(batch * nn[k] * l);

  LevelArgs a{};
  a.l = (int)l; a.LC = LC; a.geom = geom;
  a.Y = Y; a.ldy = ldy; a.sY = sY;
  // ---- upward pass ----
  for (int k = 0; k <= K; ++k) {
    a.leaf = (k == 0);
    a.nnode = nn[k];
    a.nchild = (k == 0) ? 0 : nn[k - 1];
    a.Rin = (k == 0) ? nullptr : Rbuf[k].get();
    a.tau = taus[k].get();
    a.Rout = Rbuf[k + 1].get();
    hipLaunchKernelGGL(hh_factor_kernel<T>, dim3(nn[k], (unsigned)batch), dim3(kNT), 0,
                       ctx().stream, a);
    LAUNCH_CHECK();
  }
  if (R) {
    hipLaunchKernelGGL(copy_r_kernel<T>, dim3(ceil_div(ll, 256), (unsigned)batch), dim3(256), 0,
                       ctx().stream, (int)l, (const T *)Rbuf[K + 1].get(), R, ldr, sR);
    LAUNCH_CHECK();
  }
  // ---- downward pass (explicit Q) ----
  hipLaunchKernelGGL(eye_rowmajor_kernel<T>, dim3(ceil_div(batch * ll, 256)), dim3(256), 0,
                     ctx().stream, (int)l, batch, Cbuf[K + 1].get());
  LAUNCH_CHECK();
  for (int k = K; k >= 0; --k) {
    a.leaf = (k == 0);
    a.nnode = nn[k];
    a.nchild = (k == 0) ? 0 : nn[k - 1];
    a.Rin = (k == 0) ? nullptr : Rbuf[k].get();
    a.tau = taus[k].get();
    a.Cin = Cbuf[k + 1].get();
    a.Cout = (k == 0) ? nullptr : Cbuf[k].get();
    hipLaunchKernelGGL(hh_apply_kernel<T>, dim3(nn[k], (unsigned)batch), dim3(kNT), 0,
                       ctx().stream, a);
    LAUNCH_CHECK();
  }
}

template void tsqr_rowmajor<double>(i64, i64, double *, i64, i64, i64, double *, i64, i64);
template void tsqr_rowmajor<float>(i64, i64, float *, i64, i64, i64, float *, i64, i64);

}  // namespace isvd
