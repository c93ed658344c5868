// Householder tall-skinny QR (TSQR) for the orthogonalization and postprocessing stages.
//
// The reference factors each sketch with LAPACK gelqf/orglq and, across MPI ranks, combines
// the triangular factors with a butterfly of stacked-R factorizations
// (src/lib/libisvd/core/stage/@x@_orthogonalize_tall_skinny_qr.c:127-200,
//  src/lib/libisvd/core/stage/@x@_postprocess_tall_skinny_qr.c:184-246).  On one GPU the same
// tree runs over row panels of each sketch:
//   level 0   : every row panel (<= 256 rows) is factored by one workgroup with the panel held
//               in registers (each wave owns whole columns), Householder vectors written back
//               in place, R written to the next level's buffer;
//   level k>0 : two consecutive R factors form a stacked 2l x l panel that is factored the
//               same way, until one R per sketch remains;
//   formation : top-down, each node applies its reflectors to [C; 0] (C = I at the root) and
//               hands the two l x l halves to its children; the leaves produce Q in place.
// Reflectors follow LAPACK dlarfg (beta = -sign(alpha) * ||x||, tau = (beta - alpha) / beta).
#include "la.hpp"

#include "dpp.hpp"

#include <utility>

namespace isvd {

namespace {

// Panel layout: 8 waves; wave w owns columns [w*cpw, (w+1)*cpw) of the panel (cpw = LC / 8,
// LC = l rounded up to a power of two >= 32, so cpw is 4, 8 or 16).  Inside a wave,
// lane = 16 * cs + rg: the lane holds rows rg + 16 i (i < 16) of the columns
// w*cpw + 4 cs + t (t < 4), so column j sits in register slot t = j % 4 of its lanes.  Column norms and
// the products v^T A[:, c] are reductions over the 16 lanes that share cs (four shuffle
// levels, issued level by level for all of the lane's columns at once); the Householder
// vector of each step is broadcast through a double-buffered LDS vector, one barrier per step.
constexpr int kWaves = 8;
constexpr int kNT = 64 * kWaves;  // threads per panel
constexpr int kRPL = 16;          // rows per lane
constexpr int kCPL = 4;           // columns per lane
constexpr int kMaxCPW = 16;       // columns per wave (l <= 128)
constexpr int kHMax = 16 * kRPL;  // rows per panel

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
__device__ __forceinline__ T group16_sum1(T v) {
  return row16_sum(v);
}

// The lane's columns c0 .. c0 + 3 (c0 = w * cpw + 4 cs) are consecutive, so a row of them is
// read or written with one 4-element access when the group is whole and the row start is
// 4-element aligned: the accesses of a wave then cover whole cache lines instead of 8 bytes
// every 32 (which multiplied the HBM traffic of these kernels several times over).
template <typename T> struct Quad;
template <> struct Quad<double> { typedef double v __attribute__((ext_vector_type(4))); };
template <> struct Quad<float> { typedef float v __attribute__((ext_vector_type(4))); };

__device__ __forceinline__ int quad_cols(int cpw, int cs, int c0, int l) {
  int n = cpw - 4 * cs;
  if (l - c0 < n) n = l - c0;
  return n < 0 ? 0 : (n > kCPL ? kCPL : n);
}
template <typename T>
__device__ __forceinline__ bool quad_ok(const T *p, i64 ld, int ncol) {
  return ncol == kCPL && (ld % kCPL) == 0 &&
         (reinterpret_cast<uintptr_t>(p) % (kCPL * sizeof(T))) == 0;
}
template <typename T>
__device__ __forceinline__ typename Quad<T>::v get_quad(const T *row, int ncol, bool vec) {
  typename Quad<T>::v q = {T(0), T(0), T(0), T(0)};
  if (vec) {
    q = *reinterpret_cast<const typename Quad<T>::v *>(row);
  } else {
#pragma unroll
    for (int t = 0; t < kCPL; ++t)
      if (t < ncol) q[t] = row[t];
  }
  return q;
}
template <typename T>
__device__ __forceinline__ void put_quad(T *row, int ncol, bool vec, typename Quad<T>::v q) {
  if (vec) {
    *reinterpret_cast<typename Quad<T>::v *>(row) = q;
  } else {
#pragma unroll
    for (int t = 0; t < kCPL; ++t)
      if (t < ncol) row[t] = q[t];
  }
}

template <typename T>
__device__ __forceinline__ void group16_sum(T (&v)[kCPL]) {
#pragma unroll
  for (int t = 0; t < kCPL; ++t) v[t] = group16_sum1(v[t]);
}

template <typename T>
__global__ __launch_bounds__(kNT) void hh_factor_kernel(LevelArgs a) {
  __shared__ T vsh[2][kHMax];
  // the taus stay in LDS until the end: a global store inside the loop would be waited for at
  // every step's barrier
  __shared__ T tau_sh[kHMax];
  const int j_node = blockIdx.x, b = blockIdx.y;
  const int l = a.l, cpw = a.LC / kWaves;
  // the wave index is uniform: keep it, and what depends on it, in scalar registers
  const int tid = threadIdx.x, wave = __builtin_amdgcn_readfirstlane(tid >> 6), lane = tid & 63;
  const int cs = lane >> 4, rg = lane & 15;
  T *base;
  i64 ld;
  int h;
  panel_of<T>(a, b, j_node, base, ld, h);
  for (int e = tid; e < 2 * kHMax; e += kNT) (&vsh[0][0])[e] = T(0);

  // column of slot t for this lane, or -1
  auto colof = [&](int t) {
    const int lc = 4 * cs + t;
    const int c = wave * cpw + lc;
    return (lc < cpw && c < l) ? c : -1;
  };
  const int c0 = wave * cpw + 4 * cs, ncol = quad_cols(cpw, cs, c0, l);
  const bool vb = quad_ok(base, ld, ncol);
  T x[kCPL][kRPL];
#pragma unroll
  for (int i = 0; i < kRPL; ++i) {
    const int r = rg + 16 * i;
    typename Quad<T>::v q = {T(0), T(0), T(0), T(0)};
    if (r < h) q = get_quad(base + (i64)r * ld + c0, ncol, vb);
#pragma unroll
    for (int t = 0; t < kCPL; ++t) x[t][i] = q[t];
  }
  const int kmax = h < l ? h : l;
  T *tau_out = reinterpret_cast<T *>(a.tau) + ((i64)b * a.nnode + j_node) * l;
  __syncthreads();

  // one Householder step; column j is in register slot TJ = j % 4 of the lanes that hold it
  auto hh_step = [&](int j, auto tj_c) {
    constexpr int TJ = decltype(tj_c)::value;
    const int buf = j & 1;
    const int lj = j - wave * cpw;  // column j relative to this wave
    if (lj >= 0 && lj < cpw && cs == (lj >> 2)) {
      T part = 0, av = 0;
#pragma unroll
      for (int i = 0; i < kRPL; ++i) {
        const int r = rg + 16 * i;
        if (r > j) part += x[TJ][i] * x[TJ][i];  // rows >= h hold zeros
        if (r == j) av = x[TJ][i];
      }
      const T sigma2 = group16_sum1(part);
      const T alpha = group16_sum1(av);  // one lane contributes
      T beta, tau, scale;
      if (sigma2 == T(0)) {
        beta = alpha; tau = T(0); scale = T(0);
      } else {
        const T nrm = sqrt(alpha * alpha + sigma2);
        beta = (alpha >= T(0)) ? -nrm : nrm;
        tau = (beta - alpha) / beta;
        scale = T(1) / (alpha - beta);
      }
#pragma unroll
      for (int i = 0; i < kRPL; ++i) {
        const int r = rg + 16 * i;
        if (r > j) x[TJ][i] *= scale;
        else if (r == j) x[TJ][i] = beta;
        vsh[buf][r] = (r > j && r < h) ? x[TJ][i] : (r == j ? T(1) : T(0));
      }
      if (rg == 0) tau_sh[j] = tau;
    }
    __syncthreads();
    const T tau = tau_sh[j];
    // a wave whose columns all lie left of j + 1 has nothing to update
    if (tau != T(0) && wave * cpw + cpw - 1 > j) {
      T v[kRPL];
#pragma unroll
      for (int i = 0; i < kRPL; ++i) v[i] = vsh[buf][rg + 16 * i];
      T part[kCPL];
#pragma unroll
      for (int t = 0; t < kCPL; ++t) {
        part[t] = T(0);
#pragma unroll
        for (int i = 0; i < kRPL; ++i) part[t] += v[i] * x[t][i];
      }
      group16_sum(part);
#pragma unroll
      for (int t = 0; t < kCPL; ++t) {
        if (colof(t) > j) {
          const T w = part[t] * tau;
#pragma unroll
          for (int i = 0; i < kRPL; ++i) x[t][i] -= v[i] * w;
        }
      }
    }
  };
  // cpw is a multiple of 4, so j % 4 is the slot of column j in whichever wave holds it
  for (int j0 = 0; j0 < kmax; j0 += 4) {
    hh_step(j0, std::integral_constant<int, 0>{});
    if (j0 + 1 < kmax) hh_step(j0 + 1, std::integral_constant<int, 1>{});
    if (j0 + 2 < kmax) hh_step(j0 + 2, std::integral_constant<int, 2>{});
    if (j0 + 3 < kmax) hh_step(j0 + 3, std::integral_constant<int, 3>{});
  }
  for (int e = tid; e < kmax; e += kNT) tau_out[e] = tau_sh[e];
  // V (below the diagonal) and R (on and above) back in place; R to the next level
  T *R = reinterpret_cast<T *>(a.Rout) + ((i64)b * a.nnode + j_node) * l * l;
  const bool vr = quad_ok(R, l, ncol);
#pragma unroll
  for (int i = 0; i < kRPL; ++i) {
    const int r = rg + 16 * i;
    typename Quad<T>::v q, qr;
#pragma unroll
    for (int t = 0; t < kCPL; ++t) {
      q[t] = x[t][i];
      qr[t] = (r <= c0 + t && r < kmax) ? x[t][i] : T(0);
    }
    if (r < h) put_quad(base + (i64)r * ld + c0, ncol, vb, q);
    if (r < l) put_quad(R + r * l + c0, ncol, vr, qr);
  }
}

// Internal tree node: QR of the stacked pair [R1; R2] of upper-triangular l x l factors
// (LAPACK's triangle-on-triangle case, tpqrt with l = n).  Reflector j is 1 at top row j,
// nonzero on bottom rows 0..j and zero elsewhere, so step j touches top row j and bottom rows
// 0..j only.  Top row j is therefore still the input row when step j starts: it is read from
// memory one step ahead, and the bottom triangle alone is held in registers (8 rows per lane,
// the panel layout otherwise): half the per-step work of a 2l-row panel.
// Outputs match hh_factor_kernel's: R of the node (row-major), the taus, and the bottom parts
// of the reflectors stored over R2 (the top parts below the diagonal are R1's zeros).
constexpr int kTRPL = 8;  // bottom rows per lane (l <= 128)
template <typename T>
__global__ __launch_bounds__(kNT) void tt_factor_kernel(LevelArgs a) {
  // R1, row j replaced by row j of R at step j; written out at the end (as are the taus), so
  // that no global memory operation is in flight at a step's barrier
  __shared__ T r1s[16 * kTRPL][16 * kTRPL + 1];
  __shared__ T vsh[2][16 * kTRPL];
  __shared__ T tau_sh[16 * kTRPL];
  const int j_node = blockIdx.x, b = blockIdx.y;
  const int l = a.l, cpw = a.LC / kWaves;
  const int tid = threadIdx.x, wave = __builtin_amdgcn_readfirstlane(tid >> 6), lane = tid & 63;
  const int cs = lane >> 4, rg = lane & 15;
  const int ch0 = 2 * j_node;  // first child
  const bool pair = ch0 + 1 < a.nchild;  // a lone last child: R2 = 0
  const T *R1 = reinterpret_cast<const T *>(a.Rin) + ((i64)b * a.nchild + ch0) * l * l;
  T *R2 = reinterpret_cast<T *>(a.Rin) + ((i64)b * a.nchild + ch0 + 1) * l * l;
  T *tau_out = reinterpret_cast<T *>(a.tau) + ((i64)b * a.nnode + j_node) * l;
  T *R = reinterpret_cast<T *>(a.Rout) + ((i64)b * a.nnode + j_node) * l * l;
  auto colof = [&](int t) {
    const int lc = 4 * cs + t;
    const int c = wave * cpw + lc;
    return (lc < cpw && c < l) ? c : -1;
  };
  for (int e = tid; e < l * l; e += kNT) r1s[e / l][e % l] = R1[e];
  const int c0 = wave * cpw + 4 * cs, ncol = quad_cols(cpw, cs, c0, l);
  const bool v2 = quad_ok(R2, l, ncol);
  T x[kCPL][kTRPL];
#pragma unroll
  for (int i = 0; i < kTRPL; ++i) {
    const int r = rg + 16 * i;
    typename Quad<T>::v q = {T(0), T(0), T(0), T(0)};
    if (pair && r < l) q = get_quad(R2 + r * l + c0, ncol, v2);
#pragma unroll
    for (int t = 0; t < kCPL; ++t) x[t][i] = q[t];
  }
  __syncthreads();
  auto tt_step = [&](int j, auto tj_c) __attribute__((always_inline)) {
    constexpr int TJ = decltype(tj_c)::value;
    const int buf = j & 1;
    // top row j on the lane's columns: still R1's row (earlier steps touch other top rows)
    T r1[kCPL];
#pragma unroll
    for (int t = 0; t < kCPL; ++t) {
      const int c = colof(t);
      r1[t] = (c >= 0) ? r1s[j][c] : T(0);
    }
    const int lj = j - wave * cpw;
    if (lj >= 0 && lj < cpw && cs == (lj >> 2)) {
      T part = 0;
#pragma unroll
      for (int i = 0; i < kTRPL; ++i) {
        const int r = rg + 16 * i;
        if (r <= j) part += x[TJ][i] * x[TJ][i];  // rows > j hold zeros
      }
      const T sigma2 = group16_sum1(part);
      const T alpha = r1[TJ];
      T beta, tau, scale;
      if (sigma2 == T(0)) {
        beta = alpha; tau = T(0); scale = T(0);
      } else {
        const T nrm = sqrt(alpha * alpha + sigma2);
        beta = (alpha >= T(0)) ? -nrm : nrm;
        tau = (beta - alpha) / beta;
        scale = T(1) / (alpha - beta);
      }
#pragma unroll
      for (int i = 0; i < kTRPL; ++i) {
        const int r = rg + 16 * i;
        x[TJ][i] = (r <= j) ? x[TJ][i] * scale : T(0);
        vsh[buf][r] = x[TJ][i];
      }
      if (rg == 0) {
        tau_sh[j] = tau;
        r1s[j][j] = beta;  // read by no one before the end
      }
    }
    __syncthreads();
    const T tau = tau_sh[j];
    // a wave whose columns all lie left of j + 1 has nothing to update
    if (tau != T(0) && wave * cpw + cpw - 1 > j) {
      T v[kTRPL];
#pragma unroll
      for (int i = 0; i < kTRPL; ++i) v[i] = vsh[buf][rg + 16 * i];
      T part[kCPL];
#pragma unroll
      for (int t = 0; t < kCPL; ++t) {
        part[t] = T(0);
#pragma unroll
        for (int i = 0; i < kTRPL; ++i) part[t] += v[i] * x[t][i];
      }
      group16_sum(part);
#pragma unroll
      for (int t = 0; t < kCPL; ++t) {
        const int c = colof(t);
        if (c > j) {
          const T w = (r1[t] + part[t]) * tau;
#pragma unroll
          for (int i = 0; i < kTRPL; ++i) x[t][i] -= v[i] * w;
          if (rg == 0) r1s[j][c] = r1[t] - w;  // row j of R
        }
      }
    }
  };
  for (int j0 = 0; j0 < l; j0 += 4) {
    tt_step(j0, std::integral_constant<int, 0>{});
    if (j0 + 1 < l) tt_step(j0 + 1, std::integral_constant<int, 1>{});
    if (j0 + 2 < l) tt_step(j0 + 2, std::integral_constant<int, 2>{});
    if (j0 + 3 < l) tt_step(j0 + 3, std::integral_constant<int, 3>{});
  }
  __syncthreads();
  // R (zeros below the diagonal), the taus, and the bottom parts of the reflectors over R2
  // (rows > j of column j are zero)
  for (int e = tid; e < l * l; e += kNT) {
    const int r = e / l, c = e % l;
    R[e] = (c < r) ? T(0) : r1s[r][c];
  }
  for (int e = tid; e < l; e += kNT) tau_out[e] = tau_sh[e];
  if (pair) {
#pragma unroll
    for (int i = 0; i < kTRPL; ++i) {
      const int r = rg + 16 * i;
      typename Quad<T>::v q;
#pragma unroll
      for (int t = 0; t < kCPL; ++t) q[t] = x[t][i];
      if (r < l) put_quad(R2 + r * l + c0, ncol, v2, q);
    }
  }
}

// Explicit Q: X := H_0 H_1 ... H_{k-1} [C; 0], reflectors applied last-to-first.  Every
// reflector is known before the pass starts, so the waves need no step-by-step agreement: the
// reflectors move through LDS in blocks of kVB (vectors expanded to full panel height, zeros
// above the unit diagonal and below the panel), double-buffered, with the next block's global
// loads in flight in registers while every wave applies the current block to its own columns.
// One workgroup barrier per block.
constexpr int kVB = 16;                       // reflectors per LDS block
constexpr int kVLoad = kVB * kHMax / kNT;     // entries each thread loads per block
template <typename T>
__global__ __launch_bounds__(kNT) void hh_apply_kernel(LevelArgs a) {
  __shared__ T vsh[2][kVB][kHMax + 1];
  const int j_node = blockIdx.x, b = blockIdx.y;
  const int l = a.l, cpw = a.LC / kWaves;
  // the wave index is uniform: keep it, and what depends on it, in scalar registers
  const int tid = threadIdx.x, wave = __builtin_amdgcn_readfirstlane(tid >> 6), lane = tid & 63;
  const int cs = lane >> 4, rg = lane & 15;
  T *base;
  i64 ld;
  int h;
  panel_of<T>(a, b, j_node, base, ld, h);
  const int kmax = h < l ? h : l;
  const T *C = reinterpret_cast<const T *>(a.Cin) + ((i64)b * a.nnode + j_node) * l * l;
  const T *tau_in = reinterpret_cast<const T *>(a.tau) + ((i64)b * a.nnode + j_node) * l;
  auto colof = [&](int t) {
    const int lc = 4 * cs + t;
    const int c = wave * cpw + lc;
    return (lc < cpw && c < l) ? c : -1;
  };
  const int c0 = wave * cpw + 4 * cs, ncol = quad_cols(cpw, cs, c0, l);
  T x[kCPL][kRPL];
  {
    const bool vc = quad_ok(C, l, ncol);
#pragma unroll
    for (int i = 0; i < kRPL; ++i) {
      const int r = rg + 16 * i;
      typename Quad<T>::v q = {T(0), T(0), T(0), T(0)};
      if (r < kmax) q = get_quad(C + r * l + c0, ncol, vc);
#pragma unroll
      for (int t = 0; t < kCPL; ++t) x[t][i] = q[t];
    }
  }
  // entry q of this thread's share of a block: row r = e / kVB, reflector jj = e % kVB
  // (e = tid + kNT q), so that consecutive threads read consecutive columns of one panel row;
  // v_j[r] = 0 (r < j), 1 (r == j), V[r, j] (j < r < h), 0 (r >= h).  LDS rows are padded by
  // one entry so that the transposing stores fall in distinct banks.
  T vreg[kVLoad];
  auto vload = [&](int j0) {
#pragma unroll
    for (int q = 0; q < kVLoad; ++q) {
      const int e = tid + kNT * q, r = e / kVB, jj = e % kVB, j = j0 + jj;
      vreg[q] = (j >= kmax || r < j || r >= h) ? T(0)
                                               : (r == j ? T(1) : base[(i64)r * ld + j]);
    }
  };
  auto vstore = [&](int buf) {
#pragma unroll
    for (int q = 0; q < kVLoad; ++q) {
      const int e = tid + kNT * q;
      vsh[buf][e % kVB][e / kVB] = vreg[q];
    }
  };
  const int nblk = (kmax + kVB - 1) / kVB;
  if (nblk > 0) {
    vload((nblk - 1) * kVB);
    vstore((nblk - 1) & 1);
  }
  for (int blk = nblk - 1; blk >= 0; --blk) {
    const int buf = blk & 1, j0 = blk * kVB;
    __syncthreads();  // block blk is in LDS; every wave is done with block blk + 1's buffer
    if (blk > 0) vload(j0 - kVB);
    const int jl = (j0 + kVB < kmax ? j0 + kVB : kmax) - 1;
    for (int j = jl; j >= j0; --j) {
      const T tau = tau_in[j];
      if (tau == T(0)) continue;
      T v[kRPL];
#pragma unroll
      for (int i = 0; i < kRPL; ++i) v[i] = vsh[buf][j - j0][rg + 16 * i];
      T part[kCPL];
#pragma unroll
      for (int t = 0; t < kCPL; ++t) {
        part[t] = T(0);
#pragma unroll
        for (int i = 0; i < kRPL; ++i) part[t] += v[i] * x[t][i];
      }
      group16_sum(part);
#pragma unroll
      for (int t = 0; t < kCPL; ++t) {
        const T w = part[t] * tau;
#pragma unroll
        for (int i = 0; i < kRPL; ++i) x[t][i] -= v[i] * w;
      }
    }
    if (blk > 0) vstore(buf ^ 1);
  }
  T *out;
  i64 ldo;
  if (a.leaf) {
    out = base; ldo = ld;
  } else {
    out = reinterpret_cast<T *>(a.Cout) + ((i64)b * a.nchild + 2 * j_node) * l * l;
    ldo = l;
  }
  // every wave has read its V entries from the panel before any wave overwrites it with Q
  __syncthreads();
  const bool vo = quad_ok(out, ldo, ncol);
#pragma unroll
  for (int i = 0; i < kRPL; ++i) {
    const int r = rg + 16 * i;
    typename Quad<T>::v q;
#pragma unroll
    for (int t = 0; t < kCPL; ++t) q[t] = x[t][i];
    if (r < h) put_quad(out + (i64)r * ldo + c0, ncol, vo, q);
  }
}

// Explicit Q at an internal node: [Xt; Xb] := H_0 ... H_{l-1} [C; 0] for the stacked pair
// factored by tt_factor_kernel.  H_j touches top row j and bottom rows 0..j only, so top row
// j is still C's row j when H_j is applied (last-to-first): it comes from C through LDS with
// the reflectors, block by block, and leaves as the top child's row j.  The bottom child's
// C is held in registers (8 rows per lane).  One workgroup barrier per block of kVB.
template <typename T>
__global__ __launch_bounds__(kNT) void tt_apply_kernel(LevelArgs a) {
  constexpr int kRows = 16 * kTRPL;
  constexpr int kLoad = kVB * kRows / kNT;  // entries of V and of C each thread loads per block
  __shared__ T vsh[2][kVB][kRows + 1];
  __shared__ T csh[2][kVB][kRows + 1];
  const int j_node = blockIdx.x, b = blockIdx.y;
  const int l = a.l, cpw = a.LC / kWaves;
  const int tid = threadIdx.x, wave = __builtin_amdgcn_readfirstlane(tid >> 6), lane = tid & 63;
  const int cs = lane >> 4, rg = lane & 15;
  const int ch0 = 2 * j_node;  // first child
  const bool pair = ch0 + 1 < a.nchild;
  const T *C = reinterpret_cast<const T *>(a.Cin) + ((i64)b * a.nnode + j_node) * l * l;
  T *Xt = reinterpret_cast<T *>(a.Cout) + ((i64)b * a.nchild + ch0) * l * l;
  if (!pair) {  // a lone child: every tau is zero, Q = I
    for (int e = tid; e < l * l; e += kNT) Xt[e] = C[e];
    return;
  }
  T *Xb = Xt + (i64)l * l;
  const int c0 = wave * cpw + 4 * cs, ncol = quad_cols(cpw, cs, c0, l);
  const bool vq = quad_ok(Xt, l, ncol);  // Xt and Xb: same alignment, leading dimension l
  const T *V2 = reinterpret_cast<const T *>(a.Rin) + ((i64)b * a.nchild + ch0 + 1) * l * l;
  const T *tau_in = reinterpret_cast<const T *>(a.tau) + ((i64)b * a.nnode + j_node) * l;
  auto colof = [&](int t) {
    const int lc = 4 * cs + t;
    const int c = wave * cpw + lc;
    return (lc < cpw && c < l) ? c : -1;
  };
  T x[kCPL][kTRPL];
#pragma unroll
  for (int t = 0; t < kCPL; ++t)
#pragma unroll
    for (int i = 0; i < kTRPL; ++i) x[t][i] = T(0);
  // block j0: reflector parts v_j[r] = V2[r][j] (r <= j, else 0) with consecutive threads on
  // consecutive j of one row, and C rows j (consecutive threads on consecutive columns)
  T vreg[kLoad], creg[kLoad];
  auto bload = [&](int j0) {
#pragma unroll
    for (int q = 0; q < kLoad; ++q) {
      const int e = tid + kNT * q;
      const int r = e / kVB, jj = e % kVB, j = j0 + jj;
      vreg[q] = (j < l && r <= j && r < l) ? V2[r * l + j] : T(0);
      const int jc = j0 + e / kRows, c = e % kRows;
      creg[q] = (jc < l && c < l) ? C[jc * l + c] : T(0);
    }
  };
  auto bstore = [&](int buf) {
#pragma unroll
    for (int q = 0; q < kLoad; ++q) {
      const int e = tid + kNT * q;
      vsh[buf][e % kVB][e / kVB] = vreg[q];
      csh[buf][e / kRows][e % kRows] = creg[q];
    }
  };
  const int nblk = (l + kVB - 1) / kVB;
  bload((nblk - 1) * kVB);
  bstore((nblk - 1) & 1);
  for (int blk = nblk - 1; blk >= 0; --blk) {
    const int buf = blk & 1, j0 = blk * kVB;
    __syncthreads();  // block blk is in LDS; every wave is done with block blk + 1's buffers
    if (blk > 0) bload(j0 - kVB);
    const int jl = (j0 + kVB < l ? j0 + kVB : l) - 1;
    for (int j = jl; j >= j0; --j) {
      const T tau = tau_in[j];
      T ct[kCPL];
#pragma unroll
      for (int t = 0; t < kCPL; ++t) {
        const int c = colof(t);
        ct[t] = (c >= 0) ? csh[buf][j - j0][c] : T(0);
      }
      if (tau != T(0)) {
        T v[kTRPL];
#pragma unroll
        for (int i = 0; i < kTRPL; ++i) v[i] = vsh[buf][j - j0][rg + 16 * i];
        T part[kCPL];
#pragma unroll
        for (int t = 0; t < kCPL; ++t) {
          part[t] = T(0);
#pragma unroll
          for (int i = 0; i < kTRPL; ++i) part[t] += v[i] * x[t][i];
        }
        group16_sum(part);
#pragma unroll
        for (int t = 0; t < kCPL; ++t) {
          const T w = (ct[t] + part[t]) * tau;
          ct[t] -= w;
#pragma unroll
          for (int i = 0; i < kTRPL; ++i) x[t][i] -= v[i] * w;
        }
      }
      if (rg == 0) {
        typename Quad<T>::v q;
#pragma unroll
        for (int t = 0; t < kCPL; ++t) q[t] = ct[t];
        put_quad(Xt + j * l + c0, ncol, vq, q);
      }
    }
    if (blk > 0) bstore(buf ^ 1);
  }
#pragma unroll
  for (int i = 0; i < kTRPL; ++i) {
    const int r = rg + 16 * i;
    typename Quad<T>::v q;
#pragma unroll
    for (int t = 0; t < kCPL; ++t) q[t] = x[t][i];
    if (r < l) put_quad(Xb + r * l + c0, ncol, vq, q);
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
  int LC = 4 * kWaves;  // >= 4 columns per wave: column j is in slot j % 4 of its lanes
  while (LC < l) LC *= 2;
  const i64 hmax = kHMax;  // rows one workgroup holds
  ISVD_CHECK(hmax >= 2 * l && LC / kWaves <= kMaxCPW, "internal: panel capacity");
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
    if (k == 0)
      hipLaunchKernelGGL(hh_factor_kernel<T>, dim3(nn[k], (unsigned)batch), dim3(kNT), 0,
                         ctx().stream, a);
    else
      hipLaunchKernelGGL(tt_factor_kernel<T>, dim3(nn[k], (unsigned)batch), dim3(kNT), 0,
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
    if (k == 0)
      hipLaunchKernelGGL(hh_apply_kernel<T>, dim3(nn[k], (unsigned)batch), dim3(kNT), 0,
                         ctx().stream, a);
    else
      hipLaunchKernelGGL(tt_apply_kernel<T>, dim3(nn[k], (unsigned)batch), dim3(kNT), 0,
                         ctx().stream, a);
    LAUNCH_CHECK();
  }
}

template void tsqr_rowmajor<double>(i64, i64, double *, i64, i64, i64, double *, i64, i64);
template void tsqr_rowmajor<float>(i64, i64, float *, i64, i64, i64, float *, i64, i64);

}  // namespace isvd
