// MFMA GEMM for gfx950 (CDNA4), FP64 and FP32, column-major operands.
//
// One kernel template covers the four transpose combinations.  Tiles are staged through
// LDS in k-major layout ([BK][BM] for op(A), [BK][BN] for op(B)), so the MFMA operand
// fragments (lane -> (row = lane%16, k = lane/16)) are read with unit stride.  Global loads
// walk the contiguous dimension of each operand, so both the "N" and the "T" forms load
// coalesced.  A register-prefetch double buffer overlaps the next tile's global loads with
// the current tile's MFMAs (one barrier per K tile).
// The large FP64 products (128 x 256 tiles) run instead on gemm_dl_kernel, which fills LDS
// straight from global memory three K steps ahead; the register-staged kernel keeps the
// smaller tiles, FP32, and operands that are not 16-byte aligned.
#include "la.hpp"
#include "lds_dma.hpp"

#include <type_traits>
#include <utility>

namespace isvd {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

template <typename T>
struct Mfma;

template <>
struct Mfma<double> {
  using acc_t = d4;
  // v_mfma_f64_16x16x4f64: A lane holds A[lane%16][lane/16], B lane holds B[lane/16][lane%16];
  // D register g of a lane holds D[lane/16 + 4g][lane%16].
  static __device__ __forceinline__ acc_t run(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lane, int g) { return (lane >> 4) + 4 * g; }
};

template <>
struct Mfma<float> {
  using acc_t = f4;
  // v_mfma_f32_16x16x4f32: same operand layout; D register g holds D[4*(lane/16) + g][lane%16].
  static __device__ __forceinline__ acc_t run(float a, float b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lane, int g) { return 4 * (lane >> 4) + g; }
};

template <typename T>
struct GemmArgs {
  i64 m, n, k;
  T alpha, beta;
  const T *A;
  i64 lda;
  const T *B;
  i64 ldb;
  T *C;
  i64 ldc;
  i64 sA, sB, sC;
  int ksplit;    // number of K chunks (grid.y = batch * ksplit)
  i64 kchunk;    // K elements per chunk (multiple of BK)
  T *W;          // split-K workspace: [batch][ksplit][m*n] dense col-major partials
  int group_m;   // raster group height (tiles) for L2 reuse
  int upper;     // 1: only tiles touching the upper triangle (symmetric result, mirrored after)
  int ct;        // 1: the result is stored transposed, C[c + r ldc] (n x m)
};

// Tile (tm, tn) of this workgroup; false when it has none.  Workgroups are dispatched
// round-robin over the 8 XCDs (each with its own L2): remap so that each XCD walks a
// contiguous range of the grouped raster, so that the tiles running together on one XCD share
// operand panels through that XCD's L2.  A symmetric product enumerates only the tiles
// touching the upper triangle, column by column.
template <typename T>
__device__ inline bool tile_coords(const GemmArgs<T> &g, int BM, int BN, i64 &tm, i64 &tn) {
  const i64 tiles_m = (g.m + BM - 1) / BM;
  const i64 tiles_n = (g.n + BN - 1) / BN;
  auto upper_count = [&](i64 q) {
    const i64 c = ((q + 1) * BN + BM - 1) / BM;
    return c < tiles_m ? c : tiles_m;
  };
  i64 total = tiles_m * tiles_n;
  if (g.upper) {
    total = 0;
    for (i64 q = 0; q < tiles_n; ++q) total += upper_count(q);
  }
  i64 pid = blockIdx.x;
  if (pid >= total) return false;
  {
    constexpr i64 kXcd = 8;
    const i64 xcd = pid % kXcd, local = pid / kXcd;
    const i64 per = total / kXcd, rem = total % kXcd;
    pid = (xcd < rem) ? xcd * (per + 1) + local : rem * (per + 1) + (xcd - rem) * per + local;
  }
  if (g.upper) {
    tn = 0;
    while (pid >= upper_count(tn)) pid -= upper_count(tn++);
    tm = pid;
  } else {  // grouped raster: group_m tile rows walked column by column
    const i64 in_group = (i64)g.group_m * tiles_n;
    const i64 group_id = pid / in_group;
    const i64 first_m = group_id * g.group_m;
    const i64 gsz = (tiles_m - first_m) < g.group_m ? (tiles_m - first_m) : g.group_m;
    tm = first_m + (pid % in_group) % gsz;
    tn = (pid % in_group) / gsz;
  }
  return true;
}

// Epilogue of one wave: fragment (a, b) holds rows r0 + 16 a + Mfma::row(lane, q) and column
// c0 + 16 b + lane % 16.  Without split-K, C := alpha acc + beta C (C^T when g.ct); with it,
// the raw partial goes to the workspace slice of chunk ks (dense m x n).
template <typename T, int FM, int FN>
__device__ inline void store_tile(const GemmArgs<T> &g,
                                  const typename Mfma<T>::acc_t (&acc)[FM][FN], i64 r0, i64 c0,
                                  int lane, int batch_id, int ks) {
  const int lr = lane & 15;
  if constexpr (std::is_same_v<T, float>) {
    // FP32 layout: a lane's four values are four consecutive rows of one column (rows
    // 4 (lane / 16) + q of a 16-row fragment), stored as one 16-byte vector when aligned
    const bool split = g.ksplit > 1;
    T *Cb = split ? g.W + ((i64)batch_id * g.ksplit + ks) * g.m * g.n : g.C + (i64)batch_id * g.sC;
    const i64 ldc = split ? g.m : g.ldc;
    if (!g.ct && (ldc & 3) == 0 && (reinterpret_cast<uintptr_t>(Cb) & 15) == 0) {
#pragma unroll
      for (int a = 0; a < FM; ++a)
#pragma unroll
        for (int b = 0; b < FN; ++b) {
          const i64 r = r0 + a * 16 + Mfma<T>::row(lane, 0);
          const i64 c = c0 + b * 16 + lr;
          if (c >= g.n || r >= g.m) continue;
          T *dst = Cb + r + c * ldc;
          f4 v = acc[a][b];
          if (!split) {
            v = v * g.alpha;
            if (g.beta != T(0)) {
              for (int q = 0; q < 4; ++q)
                if (r + q < g.m) v[q] += g.beta * dst[q];
            }
          }
          if (r + 3 < g.m) {
            *reinterpret_cast<f4 *>(dst) = v;
          } else {
            for (int q = 0; q < 4 && r + q < g.m; ++q) dst[q] = v[q];
          }
        }
      return;
    }
  }
  if (g.ksplit == 1) {
    T *Cb = g.C + (i64)batch_id * g.sC;
#pragma unroll
    for (int a = 0; a < FM; ++a)
#pragma unroll
      for (int b = 0; b < FN; ++b)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const i64 r = r0 + a * 16 + Mfma<T>::row(lane, q);
          const i64 c = c0 + b * 16 + lr;
          if (r < g.m && c < g.n) {
            const i64 at = g.ct ? c + r * g.ldc : r + c * g.ldc;
            T v = g.alpha * acc[a][b][q];
            if (g.beta != T(0)) v += g.beta * Cb[at];
            Cb[at] = v;
          }
        }
  } else {
    T *Wb = g.W + ((i64)batch_id * g.ksplit + ks) * g.m * g.n;
#pragma unroll
    for (int a = 0; a < FM; ++a)
#pragma unroll
      for (int b = 0; b < FN; ++b)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const i64 r = r0 + a * 16 + Mfma<T>::row(lane, q);
          const i64 c = c0 + b * 16 + lr;
          if (r < g.m && c < g.n) Wb[r + c * g.m] = acc[a][b][q];
        }
  }
}

template <typename T, bool TA, bool TB, int BM, int BN, int BK, int WGM, int WGN>
__global__ __launch_bounds__(64 * WGM * WGN) void gemm_kernel(GemmArgs<T> g) {
  constexpr int NT = 64 * WGM * WGN;
  constexpr int PAD = (sizeof(T) == 8) ? 2 : 4;
  constexpr int LDA_S = BM + PAD;
  constexpr int LDB_S = BN + PAD;
  constexpr int A_LD = BM * BK / NT;  // elements per thread per tile
  constexpr int B_LD = BK * BN / NT;
  static_assert(A_LD * NT == BM * BK && B_LD * NT == BK * BN, "tile/threads mismatch");
  constexpr int WTM = BM / WGM, WTN = BN / WGN;
  constexpr int FM = WTM / 16, FN = WTN / 16;
  using acc_t = typename Mfma<T>::acc_t;

  __shared__ T As[2][BK * LDA_S];
  __shared__ T Bs[2][BK * LDB_S];

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm = wave % WGM, wn = wave / WGM;

  i64 tm, tn;
  if (!tile_coords(g, BM, BN, tm, tn)) return;
  const i64 m0 = tm * BM, n0 = tn * BN;

  const int batch_id = blockIdx.y / g.ksplit;
  const int ks = blockIdx.y % g.ksplit;
  const i64 kbeg = (i64)ks * g.kchunk;
  i64 kend = kbeg + g.kchunk;
  if (kend > g.k) kend = g.k;

  const T *__restrict__ Ab = g.A + (i64)batch_id * g.sA;
  const T *__restrict__ Bb = g.B + (i64)batch_id * g.sB;

  T ra[A_LD], rb[B_LD];

  auto load_tiles = [&](i64 k0) {
#pragma unroll
    for (int r = 0; r < A_LD; ++r) {
      const int e = tid + NT * r;
      int i, p;
      if (!TA) { i = e % BM; p = e / BM; } else { i = e / BK; p = e % BK; }
      const i64 gi = m0 + i, gp = k0 + p;
      T v = T(0);
      if (gi < g.m && gp < kend) v = TA ? Ab[gp + gi * g.lda] : Ab[gi + gp * g.lda];
      ra[r] = v;
    }
#pragma unroll
    for (int r = 0; r < B_LD; ++r) {
      const int e = tid + NT * r;
      int j, p;
      if (TB) { j = e % BN; p = e / BN; } else { j = e / BK; p = e % BK; }
      const i64 gj = n0 + j, gp = k0 + p;
      T v = T(0);
      if (gj < g.n && gp < kend) v = TB ? Bb[gj + gp * g.ldb] : Bb[gp + gj * g.ldb];
      rb[r] = v;
    }
  };
  auto store_tiles = [&](int buf) {
#pragma unroll
    for (int r = 0; r < A_LD; ++r) {
      const int e = tid + NT * r;
      int i, p;
      if (!TA) { i = e % BM; p = e / BM; } else { i = e / BK; p = e % BK; }
      As[buf][p * LDA_S + i] = ra[r];
    }
#pragma unroll
    for (int r = 0; r < B_LD; ++r) {
      const int e = tid + NT * r;
      int j, p;
      if (TB) { j = e % BN; p = e / BN; } else { j = e / BK; p = e % BK; }
      Bs[buf][p * LDB_S + j] = rb[r];
    }
  };

  acc_t acc[FM][FN];
#pragma unroll
  for (int a = 0; a < FM; ++a)
#pragma unroll
    for (int b = 0; b < FN; ++b) acc[a][b] = acc_t{0, 0, 0, 0};

  const i64 nk = (kend > kbeg) ? (kend - kbeg + BK - 1) / BK : 0;
  if (nk > 0) {
    load_tiles(kbeg);
    store_tiles(0);
    __syncthreads();
  }
  const int lr = lane & 15, lk = lane >> 4;
  for (i64 kt = 0; kt < nk; ++kt) {
    const int buf = (int)(kt & 1);
    if (kt + 1 < nk) load_tiles(kbeg + (kt + 1) * BK);
    const T *as = As[buf];
    const T *bs = Bs[buf];
#pragma unroll
    for (int kk = 0; kk < BK; kk += 4) {
      T fa[FM], fb[FN];
#pragma unroll
      for (int a = 0; a < FM; ++a) fa[a] = as[(kk + lk) * LDA_S + wm * WTM + a * 16 + lr];
#pragma unroll
      for (int b = 0; b < FN; ++b) fb[b] = bs[(kk + lk) * LDB_S + wn * WTN + b * 16 + lr];
#pragma unroll
      for (int a = 0; a < FM; ++a)
#pragma unroll
        for (int b = 0; b < FN; ++b) acc[a][b] = Mfma<T>::run(fa[a], fb[b], acc[a][b]);
    }
    if (kt + 1 < nk) store_tiles(buf ^ 1);
    __syncthreads();
  }

  store_tile<T, FM, FN>(g, acc, m0 + wm * WTM, n0 + wn * WTN, lane, batch_id, ks);
}

// FP64 GEMM with the operands moved from global memory straight into LDS
// (global_load_lds_dwordx4, no register staging), three K steps in flight.
//
// Tile 128 x 256 x 16, eight waves in a 2 x 4 grid of 64 x 64 (4 x 4 fragments of
// v_mfma_f64_16x16x4f64).  Each operand tile has one of two LDS layouts, chosen by which of
// its dimensions is contiguous in global memory:
//   K-contiguous (op(A) = A^T, op(B) = B): one 128-byte row of 16 K values per tile row; the
//     eight 16-byte chunks of row r sit at position chunk ^ ((r >> 1) & 7), so that the
//     16 rows of a fragment read fall in distinct banks.  A fill instruction covers 8 rows:
//     lane l fetches row 8 j + l / 8, chunk (l % 8) ^ swizzle (the LDS position is fixed to
//     l % 8 by the instruction).
//   MN-contiguous (op(A) = A, op(B) = B^T): one K row of the tile per 1 KB (A) or 2 KB (B),
//     padded by 64 bytes so that rows 2 apart start 128 bytes apart in the banks.
// The MFMA consumes K in pairs: lane l / 16 = s supplies K index k8 + 2 s + h to the
// instruction h of a pair (h = 0, 1), so a K-contiguous fragment is one 16-byte LDS read
// serving both instructions, and the two operands agree on the K index of every product.
// Rows past the matrix edge are fetched from the last valid row (or column pair): they feed
// only outputs that are not stored.  Requirements (checked by the host): 16-byte aligned
// operands, even leading dimensions and batch strides, even m (n) when op(A) (op(B)) is
// MN-contiguous.  A partial last K step goes through registers with bounds checks.
namespace dl {
constexpr int BM = 128, BN = 256, BK = 16, NT = 512, S = 3;
constexpr int PITCH_A = BM * 8 + 64, PITCH_B = BN * 8 + 64;
constexpr int FILLS = (BM + BN) * BK * 8 / 1024 / (NT / 64);  // per wave per K step
constexpr int A_FILLS = BM * BK * 8 / 1024;                   // fill instructions for op(A)
static_assert(A_FILLS % (NT / 64) == 0, "op(A) fills must be whole per-wave slots");
template <bool KA, bool KB>
struct Layout {
  static constexpr int A_BYTES = KA ? BM * BK * 8 : BK * PITCH_A;
  static constexpr int B_BYTES = KB ? BN * BK * 8 : BK * PITCH_B;
  static constexpr int STAGE = A_BYTES + B_BYTES;
  static_assert(S * STAGE <= 160 * 1024, "LDS budget");
};
__device__ inline int swz(int r) { return (r >> 1) & 7; }
}  // namespace dl

typedef double d2 __attribute__((ext_vector_type(2)));

template <bool KA, bool KB>
__global__ __launch_bounds__(dl::NT) void gemm_dl_kernel(GemmArgs<double> g) {
  using namespace dl;
  using L = Layout<KA, KB>;
  // one LDS object per stage: the compiler then sees that a fill of one stage and the
  // fragment reads of another do not alias
  __shared__ __align__(16) unsigned char st0[L::STAGE], st1[L::STAGE], st2[L::STAGE];
  auto stage = [&](auto k) -> unsigned char * {
    constexpr int kk = decltype(k)::value;
    if constexpr (kk == 0) return st0;
    else if constexpr (kk == 1) return st1;
    else return st2;
  };
  i64 tm, tn;
  if (!tile_coords(g, BM, BN, tm, tn)) return;
  const i64 m0 = tm * BM, n0 = tn * BN;
  const int batch_id = blockIdx.y / g.ksplit;
  const int ks = blockIdx.y % g.ksplit;
  const i64 kbeg = (i64)ks * g.kchunk;
  const i64 kend = (kbeg + g.kchunk < g.k) ? kbeg + g.kchunk : g.k;
  const double *__restrict__ Ab = g.A + (i64)batch_id * g.sA;
  const double *__restrict__ Bb = g.B + (i64)batch_id * g.sB;
  // the wave index is uniform: keep it, and the LDS offsets that depend on it, scalar
  const int t = threadIdx.x, lane = t & 63, wave = __builtin_amdgcn_readfirstlane(t >> 6);
  const int wm = wave & 1, wn = wave >> 1;
  const int lr = lane & 15, lk = lane >> 4;

  // fill slot u of this wave is fill instruction f = 8 u + wave of a stage: op(A) for
  // f < A_FILLS, op(B) after
  // wave-uniform operand bases at this chunk's first K step; per lane, a 32-bit element
  // offset for each fill slot (the host keeps the leading dimensions below 2^23)
  const double *abase = Ab + (KA ? m0 * g.lda + kbeg : m0 + kbeg * g.lda);
  const double *bbase = Bb + (KB ? n0 * g.ldb + kbeg : n0 + kbeg * g.ldb);
  unsigned off[FILLS], loff[FILLS];
#pragma unroll
  for (int u = 0; u < FILLS; ++u) {
    const int f = 8 * u + wave;
    if (u < A_FILLS / 8) {
      if (KA) {
        const int row = 8 * f + (lane >> 3), c = (lane & 7) ^ swz(row);
        const i64 gi = (m0 + row < g.m) ? row : g.m - 1 - m0;
        off[u] = (unsigned)(gi * g.lda + 2 * c);
        loff[u] = f * 1024;
      } else {
        const i64 i = (m0 + 2 * lane <= g.m - 2) ? 2 * lane : g.m - 2 - m0;
        off[u] = (unsigned)(i + f * g.lda);
        loff[u] = f * PITCH_A;
      }
    } else {
      const int fb = f - A_FILLS;
      if (KB) {
        const int row = 8 * fb + (lane >> 3), c = (lane & 7) ^ swz(row);
        const i64 gj = (n0 + row < g.n) ? row : g.n - 1 - n0;
        off[u] = (unsigned)(gj * g.ldb + 2 * c);
        loff[u] = L::A_BYTES + fb * 1024;
      } else {
        const int q = fb >> 1, h = fb & 1;
        const i64 j =
            (n0 + 128 * h + 2 * lane <= g.n - 2) ? 128 * h + 2 * lane : g.n - 2 - n0;
        off[u] = (unsigned)(j + q * g.ldb);
        loff[u] = L::A_BYTES + q * PITCH_B + h * 1024;
      }
    }
  }
  const i64 dA = KA ? BK : BK * g.lda, dB = KB ? BK : BK * g.ldb;  // per K step
  auto fill = [&](i64 kt, unsigned char *st) {
    const unsigned base = lds_addr(st);
    const double *ak = abase + kt * dA, *bk = bbase + kt * dB;
#pragma unroll
    for (int u = 0; u < FILLS; ++u)
      lds_fill16((u < A_FILLS / 8 ? ak : bk) + off[u],
                 __builtin_amdgcn_readfirstlane(base + loff[u]));
  };
  // LDS byte offset of op(A)[row][k] / op(B)[k][col] in a stage
  auto a_off = [](int row, int k) {
    return KA ? row * 128 + (((k >> 1) ^ swz(row)) << 4) + (k & 1) * 8 : k * PITCH_A + row * 8;
  };
  auto b_off = [](int col, int k) {
    return L::A_BYTES +
           (KB ? col * 128 + (((k >> 1) ^ swz(col)) << 4) + (k & 1) * 8 : k * PITCH_B + col * 8);
  };
  // the partial last K step: registers with bounds checks, zeros past kend
  auto fill_checked = [&](i64 k0, unsigned char *st) {
    for (int e = t; e < BM * BK; e += NT) {
      const int row = KA ? e >> 4 : e % BM, k = KA ? e & 15 : e / BM;
      const i64 gi = m0 + row, gk = k0 + k;
      double v = 0.0;
      if (gi < g.m && gk < kend) v = KA ? Ab[gi * g.lda + gk] : Ab[gi + gk * g.lda];
      *reinterpret_cast<double *>(st + a_off(row, k)) = v;
    }
    for (int e = t; e < BN * BK; e += NT) {
      const int col = KB ? e >> 4 : e % BN, k = KB ? e & 15 : e / BN;
      const i64 gj = n0 + col, gk = k0 + k;
      double v = 0.0;
      if (gj < g.n && gk < kend) v = KB ? Bb[gk + gj * g.ldb] : Bb[gj + gk * g.ldb];
      *reinterpret_cast<double *>(st + b_off(col, k)) = v;
    }
  };

  d4 acc[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) acc[a][b] = d4{0, 0, 0, 0};

  // fragments of one half step (8 K values): lane l / 16 = s holds K index k8 + 2 s + h for
  // instruction h of each pair
  typedef double Frag[4][2];
  auto load = [&](const unsigned char *st, int k8, Frag &fa, Frag &fb) {
    const int k = k8 + 2 * lk;  // this lane's K pair
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      const int r = wm * 64 + 16 * a + lr;
      if (KA) {
        const d2 v = *reinterpret_cast<const d2 *>(st + a_off(r, k));
        fa[a][0] = v.x;
        fa[a][1] = v.y;
      } else {
        fa[a][0] = *reinterpret_cast<const double *>(st + a_off(r, k));
        fa[a][1] = *reinterpret_cast<const double *>(st + a_off(r, k + 1));
      }
    }
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      const int c = wn * 64 + 16 * b + lr;
      if (KB) {
        const d2 v = *reinterpret_cast<const d2 *>(st + b_off(c, k));
        fb[b][0] = v.x;
        fb[b][1] = v.y;
      } else {
        fb[b][0] = *reinterpret_cast<const double *>(st + b_off(c, k));
        fb[b][1] = *reinterpret_cast<const double *>(st + b_off(c, k + 1));
      }
    }
  };
  // the 16 products of instruction h of a half step
  auto mma = [&](const Frag &fa, const Frag &fb, int h) {
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int b = 0; b < 4; ++b)
        acc[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(fa[a][h], fb[b][h], acc[a][b], 0, 0, 0);
  };

  const i64 nkf = (kend - kbeg) / BK;  // full K steps
  const bool tail = ((kend - kbeg) % BK) != 0;
  // prologue: steps 0 .. S - 2 into stages 0 .. S - 2
  if (nkf > 0) fill(0, st0);
  if (nkf > 1) fill(1, st1);
  // Step kt uses stage kt % S and refills stage (kt - 1) % S with step kt + S - 1.  The
  // pipeline is rotated by a quarter step: the last 16 products of step kt - 1 (operands
  // already in registers) are issued after barrier kt, while the first fragments of stage kt
  // are read, so the matrix cores do not wait on LDS after the barrier.
  Frag fa0, fb0, fa1, fb1;
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int h = 0; h < 2; ++h) fa1[a][h] = fb1[a][h] = 0.0;  // nothing deferred before step 0
  auto step = [&](i64 kt, auto k) {
    constexpr int kk = decltype(k)::value;
    // this wave's fills of step kt have landed and its fragment reads of stage kt - 1 are
    // done (another wave refills that stage after the barrier)
    if (kt + S - 2 < nkf) __builtin_amdgcn_s_waitcnt(waitcnt_vm_lgkm0(FILLS * (S - 2)));
    else __builtin_amdgcn_s_waitcnt(waitcnt_vm_lgkm0(0));
    // the deferred operands (instruction 1 of the last half step) are complete here: say so,
    // or the compiler waits for them (and for every LDS read issued after the barrier) in
    // front of the deferred products
    asm volatile("" ::"v"(fa1[0][1]), "v"(fa1[1][1]), "v"(fa1[2][1]), "v"(fa1[3][1]),
                 "v"(fb1[0][1]), "v"(fb1[1][1]), "v"(fb1[2][1]), "v"(fb1[3][1]));
    bare_barrier();
    if (kt + S - 1 < nkf)
      fill(kt + S - 1, stage(std::integral_constant<int, (kk + S - 1) % S>{}));
    load(stage(k), 0, fa0, fb0);
    // keep the reads above the deferred products, which then cover their latency
    __builtin_amdgcn_sched_barrier(0);
    mma(fa1, fb1, 1);
    mma(fa0, fb0, 0);
    mma(fa0, fb0, 1);
    load(stage(k), 8, fa1, fb1);
    mma(fa1, fb1, 0);
  };
  for (i64 kb = 0; kb < nkf; kb += S) {
    step(kb, std::integral_constant<int, 0>{});
    if (kb + 1 < nkf) step(kb + 1, std::integral_constant<int, 1>{});
    if (kb + 2 < nkf) step(kb + 2, std::integral_constant<int, 2>{});
  }
  mma(fa1, fb1, 1);  // the deferred products of the last full step (zeros when none)
  if (tail) {
    __syncthreads();  // every wave is done with the stages
    fill_checked(kbeg + nkf * BK, st0);
    __syncthreads();
    load(st0, 0, fa0, fb0);
    load(st0, 8, fa1, fb1);
    mma(fa0, fb0, 0);
    mma(fa0, fb0, 1);
    mma(fa1, fb1, 0);
    mma(fa1, fb1, 1);
  }
  store_tile<double, 4, 4>(g, acc, m0 + wm * 64, n0 + wn * 64, lane, batch_id, ks);
}

// The same kernel for FP32 operands that are both K-contiguous (op(A) = A^T, op(B) = B): a K
// step of 32 floats fills the same 128-byte LDS rows (eight 16-byte chunks of four K values,
// chunk c of row r at position c ^ swz(r)), so the stages, fills and raster are those of the
// FP64 kernel byte for byte.  v_mfma_f32_16x16x4f32 takes one float per lane with lane l / 16
// = s supplying K index k16 + 4 s + h to instruction h (h = 0..3) of a 16-byte chunk read, so a
// half step (16 K values) is four instructions per fragment pair; the last instruction of each
// step is deferred past the next barrier as in the FP64 kernel.
namespace dl {
constexpr int BK32 = 32;
}  // namespace dl

typedef float f4v __attribute__((ext_vector_type(4)));

__global__ __launch_bounds__(dl::NT) void gemm_dl32_kernel(GemmArgs<float> g) {
  using namespace dl;
  using L = Layout<true, true>;  // 128-byte rows for both operands, as in the FP64 kernel
  static_assert(L::A_BYTES == BM * BK32 * 4 && L::B_BYTES == BN * BK32 * 4, "row layout");
  __shared__ __align__(16) unsigned char st0[L::STAGE], st1[L::STAGE], st2[L::STAGE];
  auto stage = [&](auto k) -> unsigned char * {
    constexpr int kk = decltype(k)::value;
    if constexpr (kk == 0) return st0;
    else if constexpr (kk == 1) return st1;
    else return st2;
  };
  i64 tm, tn;
  if (!tile_coords(g, BM, BN, tm, tn)) return;
  const i64 m0 = tm * BM, n0 = tn * BN;
  const int batch_id = blockIdx.y / g.ksplit;
  const int ks = blockIdx.y % g.ksplit;
  const i64 kbeg = (i64)ks * g.kchunk;
  const i64 kend = (kbeg + g.kchunk < g.k) ? kbeg + g.kchunk : g.k;
  const float *__restrict__ Ab = g.A + (i64)batch_id * g.sA;
  const float *__restrict__ Bb = g.B + (i64)batch_id * g.sB;
  const int t = threadIdx.x, lane = t & 63, wave = __builtin_amdgcn_readfirstlane(t >> 6);
  const int wm = wave & 1, wn = wave >> 1;
  const int lr = lane & 15, lk = lane >> 4;

  const float *abase = Ab + m0 * g.lda + kbeg;
  const float *bbase = Bb + n0 * g.ldb + kbeg;
  unsigned off[FILLS], loff[FILLS];
#pragma unroll
  for (int u = 0; u < FILLS; ++u) {
    const int f = 8 * u + wave;
    if (u < A_FILLS / 8) {
      const int row = 8 * f + (lane >> 3), c = (lane & 7) ^ swz(row);
      const i64 gi = (m0 + row < g.m) ? row : g.m - 1 - m0;
      off[u] = (unsigned)(gi * g.lda + 4 * c);
      loff[u] = f * 1024;
    } else {
      const int fb = f - A_FILLS;
      const int row = 8 * fb + (lane >> 3), c = (lane & 7) ^ swz(row);
      const i64 gj = (n0 + row < g.n) ? row : g.n - 1 - n0;
      off[u] = (unsigned)(gj * g.ldb + 4 * c);
      loff[u] = L::A_BYTES + fb * 1024;
    }
  }
  auto fill = [&](i64 kt, unsigned char *st) {
    const unsigned base = lds_addr(st);
    const float *ak = abase + kt * BK32, *bk = bbase + kt * BK32;
#pragma unroll
    for (int u = 0; u < FILLS; ++u)
      lds_fill16((u < A_FILLS / 8 ? ak : bk) + off[u],
                 __builtin_amdgcn_readfirstlane(base + loff[u]));
  };
  // LDS byte offset of op(A)[row][k] / op(B)[k][col] in a stage
  auto a_off = [](int row, int k) { return row * 128 + (((k >> 2) ^ swz(row)) << 4) + (k & 3) * 4; };
  auto b_off = [](int col, int k) {
    return L::A_BYTES + col * 128 + (((k >> 2) ^ swz(col)) << 4) + (k & 3) * 4;
  };
  f4 acc[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) acc[a][b] = f4{0.f, 0.f, 0.f, 0.f};

  // fragments of one half step (16 K values): lane l / 16 = s holds K index k16 + 4 s + h for
  // instruction h
  typedef float Frag[4][4];
  auto load = [&](const unsigned char *st, int k16, Frag &fa, Frag &fb) {
    const int k = k16 + 4 * lk;  // this lane's K quad
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      const f4v v = *reinterpret_cast<const f4v *>(st + a_off(wm * 64 + 16 * a + lr, k));
      fa[a][0] = v.x; fa[a][1] = v.y; fa[a][2] = v.z; fa[a][3] = v.w;
    }
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      const f4v v = *reinterpret_cast<const f4v *>(st + b_off(wn * 64 + 16 * b + lr, k));
      fb[b][0] = v.x; fb[b][1] = v.y; fb[b][2] = v.z; fb[b][3] = v.w;
    }
  };
  // the 16 products of instruction h of a half step
  auto mma = [&](const Frag &fa, const Frag &fb, int h) {
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int b = 0; b < 4; ++b)
        acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x4f32(fa[a][h], fb[b][h], acc[a][b], 0, 0, 0);
  };

  const i64 nkf = (kend - kbeg) / BK32;  // whole K steps only (dl32_eligible)
  if (nkf > 0) fill(0, st0);
  if (nkf > 1) fill(1, st1);
  Frag fa0, fb0, fa1, fb1;
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int h = 0; h < 4; ++h) fa1[a][h] = fb1[a][h] = 0.f;  // nothing deferred before step 0
  auto step = [&](i64 kt, auto k) {
    constexpr int kk = decltype(k)::value;
    if (kt + S - 2 < nkf) __builtin_amdgcn_s_waitcnt(waitcnt_vm_lgkm0(FILLS * (S - 2)));
    else __builtin_amdgcn_s_waitcnt(waitcnt_vm_lgkm0(0));
    asm volatile("" ::"v"(fa1[0][3]), "v"(fa1[1][3]), "v"(fa1[2][3]), "v"(fa1[3][3]),
                 "v"(fb1[0][3]), "v"(fb1[1][3]), "v"(fb1[2][3]), "v"(fb1[3][3]));
    bare_barrier();
    if (kt + S - 1 < nkf)
      fill(kt + S - 1, stage(std::integral_constant<int, (kk + S - 1) % S>{}));
    load(stage(k), 0, fa0, fb0);
    __builtin_amdgcn_sched_barrier(0);
    mma(fa1, fb1, 3);
    mma(fa0, fb0, 0);
    mma(fa0, fb0, 1);
    mma(fa0, fb0, 2);
    mma(fa0, fb0, 3);
    load(stage(k), 16, fa1, fb1);
    mma(fa1, fb1, 0);
    mma(fa1, fb1, 1);
    mma(fa1, fb1, 2);
  };
  for (i64 kb = 0; kb < nkf; kb += S) {
    step(kb, std::integral_constant<int, 0>{});
    if (kb + 1 < nkf) step(kb + 1, std::integral_constant<int, 1>{});
    if (kb + 2 < nkf) step(kb + 2, std::integral_constant<int, 2>{});
  }
  mma(fa1, fb1, 3);  // the deferred products of the last step (zeros when none)
  store_tile<float, 4, 4>(g, acc, m0 + wm * 64, n0 + wn * 64, lane, batch_id, ks);
}

// Whether gemm_dl32_kernel can take this product: both operands K-contiguous, every fill an
// aligned 16-byte read inside the operands, and every K chunk whole K steps (the kernel has no
// partial-step path; K = l or N l in the products that take it).
inline bool dl32_eligible(const GemmArgs<float> &g, i64 batch) {
  auto aligned = [](const void *p) { return (reinterpret_cast<uintptr_t>(p) & 15) == 0; };
  if (g.k % dl::BK32 != 0 || (g.ksplit > 1 && g.kchunk % dl::BK32 != 0)) return false;
  if (!aligned(g.A) || !aligned(g.B)) return false;
  if ((g.lda & 3) || (g.ldb & 3)) return false;
  if (batch > 1 && ((g.sA & 3) || (g.sB & 3))) return false;
  if (g.lda >= (1 << 23) || g.ldb >= (1 << 23)) return false;
  return true;
}

// Whether gemm_dl_kernel<KA, KB> can take this product (every fill is an aligned 16-byte
// read inside the operands).
inline bool dl_eligible(const GemmArgs<double> &g, bool KA, bool KB, i64 batch) {
  auto aligned = [](const void *p) { return (reinterpret_cast<uintptr_t>(p) & 15) == 0; };
  if (!aligned(g.A) || !aligned(g.B)) return false;
  if ((g.lda & 1) || (g.ldb & 1)) return false;
  if (batch > 1 && ((g.sA & 1) || (g.sB & 1))) return false;
  if (!KA && (g.m & 1)) return false;
  if (!KB && (g.n & 1)) return false;
  // per-lane fill offsets are 32-bit element counts below 256 leading dimensions
  if (g.lda >= (1 << 23) || g.ldb >= (1 << 23)) return false;
  return true;
}

// C = alpha * sum_s W[s] + beta * C  (fixed summation order -> deterministic).
template <typename T>
__global__ void splitk_reduce_kernel(i64 m, i64 n, int ksplit, const T *__restrict__ W,
                                     T alpha, T beta, T *__restrict__ C, i64 ldc, i64 sC,
                                     int ct) {
  const int b = blockIdx.y;
  const i64 mn = m * n;
  const T *Wb = W + (i64)b * ksplit * mn;
  T *Cb = C + (i64)b * sC;
  for (i64 e = blockIdx.x * (i64)blockDim.x + threadIdx.x; e < mn;
       e += (i64)gridDim.x * blockDim.x) {
    T s = T(0);
    for (int q = 0; q < ksplit; ++q) s += Wb[q * mn + e];
    const i64 r = e % m, c = e / m;
    const i64 at = ct ? c + r * ldc : r + c * ldc;
    T v = alpha * s;
    if (beta != T(0)) v += beta * Cb[at];
    Cb[at] = v;
  }
}

// C[i, j] := C[j, i] for i > j (n x n): completes a symmetric product computed on upper tiles
template <typename T>
__global__ void mirror_upper_kernel(i64 n, T *C, i64 ldc) {
  for (i64 e = blockIdx.x * (i64)blockDim.x + threadIdx.x; e < n * n;
       e += (i64)gridDim.x * blockDim.x) {
    const i64 i = e % n, j = e / n;
    if (i > j) C[i + j * ldc] = C[j + i * ldc];
  }
}

template <typename T>
__global__ void scale_kernel(i64 m, i64 n, T beta, T *C, i64 ldc, i64 sC) {
  T *Cb = C + (i64)blockIdx.y * sC;
  for (i64 e = blockIdx.x * (i64)blockDim.x + threadIdx.x; e < m * n;
       e += (i64)gridDim.x * blockDim.x) {
    const i64 r = e % m, c = e / m;
    Cb[r + c * ldc] = (beta == T(0)) ? T(0) : beta * Cb[r + c * ldc];
  }
}

namespace {

constexpr int kTargetBlocks = 1024;  // ~4 workgroups per CU on 256 CUs

template <typename T, bool TA, bool TB, int BM, int BN, int BK, int WGM, int WGN>
void launch_cfg(GemmArgs<T> g, i64 batch) {
  const i64 tiles_m = (g.m + BM - 1) / BM, tiles_n = (g.n + BN - 1) / BN;
  i64 tiles = tiles_m * tiles_n;
  if (g.upper) {  // only tiles touching the upper triangle are launched
    tiles = 0;
    for (i64 q = 0; q < tiles_n; ++q) {
      const i64 c = ((q + 1) * BN + BM - 1) / BM;
      tiles += c < tiles_m ? c : tiles_m;
    }
  }
  const i64 busy = tiles;
  // split-K when the output alone cannot fill the device
  int ksplit = 1;
  if (busy * batch < kTargetBlocks && g.k >= 8 * BK) {
    i64 want = (kTargetBlocks + busy * batch - 1) / (busy * batch);
    i64 maxs = g.k / (4 * BK);
    // partial sums cost ksplit x the output in HBM traffic: keep them within ~64 MB
    const i64 by_traffic = (i64)(64ll << 20) / ((i64)sizeof(T) * g.m * g.n * batch);
    if (maxs > (by_traffic > 4 ? by_traffic : 4)) maxs = by_traffic > 4 ? by_traffic : 4;
    if (want > maxs) want = maxs;
    if (want > 64) want = 64;
    if (want < 1) want = 1;
    ksplit = (int)want;
  }
  // Balance: with a few rounds of workgroups over the CUs a partial last round idles most of
  // them.  For long K, take the split count (up to 16) whose rounds are fullest, as long as the
  // partial sums stay below a quarter of the operand traffic.
  {
    constexpr i64 kCUs = 256;
    const i64 lds_bytes = 2 * (i64)BK * ((BM + 2) + (BN + 2)) * (i64)sizeof(T);
    const i64 per_cu = (160 * 1024) / lds_bytes < 1 ? 1 : (160 * 1024) / lds_bytes;
    const i64 slots = kCUs * per_cu;
    auto fill = [&](i64 ks) {
      const i64 w = busy * batch * ks;
      return (double)w / (double)(((w + slots - 1) / slots) * slots);
    };
    if (busy * batch * ksplit < 8 * slots && g.k >= 256 * BK) {
      const double op_bytes = (double)sizeof(T) * (double)(g.m + g.n) * (double)g.k * batch;
      int best = ksplit;
      for (int ks = ksplit + 1; ks <= 16; ++ks) {
        if ((double)sizeof(T) * ks * (double)g.m * g.n * batch > 0.25 * op_bytes) break;
        if (g.k / ks < 64 * BK) break;
        if (fill(ks) > fill(best) + 0.02) best = ks;
      }
      ksplit = best;
    }
  }
  if (ctx().max_wgs > 0)
    while (ksplit > 1 && busy * batch * ksplit > ctx().max_wgs) --ksplit;
  // the partial sums must fit the space available for them (bounded under a scratch arena)
  {
    const size_t per = sizeof(T) * (size_t)(batch * g.m * g.n);
    const size_t budget = splitk_budget();
    while (ksplit > 1 && per * (size_t)ksplit > budget) --ksplit;
  }
  i64 kchunk = (g.k + ksplit - 1) / ksplit;
  kchunk = (kchunk + BK - 1) / BK * BK;
  ksplit = (int)((g.k + kchunk - 1) / kchunk);
  if (ksplit < 1) ksplit = 1;
  g.ksplit = ksplit;
  g.kchunk = (ksplit == 1) ? (g.k > 0 ? g.k : 1) : kchunk;
  g.group_m = 8;
  if (ksplit > 1)
    g.W = static_cast<T *>(splitk_workspace(sizeof(T) * (size_t)(batch * ksplit * g.m * g.n)));
  ISVD_CHECK(batch * ksplit <= 65535, "grid.y too large (batch %lld x ksplit %d)", batch, ksplit);
  dim3 grid((unsigned)tiles, (unsigned)(batch * ksplit));
  bool done = false;
  // (op(A) = A^T with op(B) = B^T, unused by the stages, stays on the register-staged kernel)
  if constexpr (std::is_same_v<T, double> && BM == dl::BM && BN == dl::BN && !(TA && TB)) {
    if (dl_eligible(g, TA, !TB, batch)) {
      hipLaunchKernelGGL((gemm_dl_kernel<TA, !TB>), grid, dim3(dl::NT), 0, ctx().stream, g);
      done = true;
    }
  }
  if constexpr (std::is_same_v<T, float> && BM == dl::BM && BN == dl::BN && TA && !TB) {
    if (dl32_eligible(g, batch)) {
      hipLaunchKernelGGL(gemm_dl32_kernel, grid, dim3(dl::NT), 0, ctx().stream, g);
      done = true;
    }
  }
  if (!done)
    hipLaunchKernelGGL((gemm_kernel<T, TA, TB, BM, BN, BK, WGM, WGN>), grid,
                       dim3(64 * WGM * WGN), 0, ctx().stream, g);
  LAUNCH_CHECK();
  if (ksplit > 1) {
    const i64 mn = g.m * g.n;
    int blocks = (int)((mn + 255) / 256);
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(splitk_reduce_kernel<T>, dim3(blocks, (unsigned)batch), dim3(256), 0,
                       ctx().stream, g.m, g.n, ksplit, (const T *)g.W, g.alpha, g.beta, g.C,
                       g.ldc, g.sC, g.ct);
    LAUNCH_CHECK();
  }
}

// 64 x 64 tiles with a given number of K chunks and no other choice made: the caller either
// wants the raw partials (ksplit > 1, written to g.W) or a product whose short K needs no split.
// Returns the number of chunks actually used (K is cut into multiples of the K tile).
template <typename T, bool TA, bool TB>
int launch_fixed64(GemmArgs<T> g, i64 batch, int ksplit) {
  constexpr int BM = 64, BN = 64, BK = 16;
  const i64 tiles = ((g.m + BM - 1) / BM) * ((g.n + BN - 1) / BN);
  if (ksplit < 1) ksplit = 1;
  i64 kchunk = (g.k + ksplit - 1) / ksplit;
  kchunk = (kchunk + BK - 1) / BK * BK;
  ksplit = (int)((g.k + kchunk - 1) / kchunk);
  if (ksplit < 1) ksplit = 1;
  g.ksplit = ksplit;
  g.kchunk = (ksplit == 1) ? (g.k > 0 ? g.k : 1) : kchunk;
  g.group_m = 8;
  ISVD_CHECK(batch * ksplit <= 65535, "grid.y too large (batch %lld x ksplit %d)", batch, ksplit);
  hipLaunchKernelGGL((gemm_kernel<T, TA, TB, BM, BN, BK, 2, 2>),
                     dim3((unsigned)tiles, (unsigned)(batch * ksplit)), dim3(256), 0,
                     ctx().stream, g);
  LAUNCH_CHECK();
  return ksplit;
}

template <typename T>
int fixed64(char transa, char transb, GemmArgs<T> g, i64 batch, int ksplit) {
  const bool ta = (transa == 'T' || transa == 't');
  const bool tb = (transb == 'T' || transb == 't');
  if (!ta && !tb) return launch_fixed64<T, false, false>(g, batch, ksplit);
  if (!ta && tb) return launch_fixed64<T, false, true>(g, batch, ksplit);
  if (ta && !tb) return launch_fixed64<T, true, false>(g, batch, ksplit);
  return launch_fixed64<T, true, true>(g, batch, ksplit);
}

template <typename T, bool TA, bool TB>
void launch(GemmArgs<T> g, i64 batch) {
  const bool big = (g.m >= 128 && g.n >= 128);
  if (!big) {
    launch_cfg<T, TA, TB, 64, 64, 16, 2, 2>(g, batch);
    return;
  }
  // 128 x 256 tiles (2 x 4 waves of 64 x 64); 128 x 128 when a 256-wide tile would be half
  // empty.  (256 x 128, 128 x 128 x 32 and 256 x 256 tiles were measured slower on the sketch.)
  if (g.n <= 128) {
    launch_cfg<T, TA, TB, 128, 128, 16, 2, 2>(g, batch);
    return;
  }
  launch_cfg<T, TA, TB, 128, 256, 16, 2, 4>(g, batch);
}

}  // namespace

template <typename T>
static void gemm_impl(char transa, char transb, i64 m, i64 n, i64 k, T alpha, const T *A,
                      i64 lda, const T *B, i64 ldb, T beta, T *C, i64 ldc, i64 batch, i64 sA,
                      i64 sB, i64 sC, int upper, int ct = 0) {
  if (m <= 0 || n <= 0 || batch <= 0) return;
  const bool ta = (transa == 'T' || transa == 't' || transa == 'C' || transa == 'c');
  const bool tb = (transb == 'T' || transb == 't' || transb == 'C' || transb == 'c');
  ISVD_CHECK(ldc >= (ct ? n : m), "ldc (%lld) too small", ldc);
  if (k <= 0 || alpha == T(0)) {
    int blocks = (int)((m * n + 255) / 256);
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(scale_kernel<T>, dim3(blocks, (unsigned)batch), dim3(256), 0,
                       ctx().stream, ct ? n : m, ct ? m : n, beta, C, ldc, sC);
    LAUNCH_CHECK();
    return;
  }
  ISVD_CHECK(lda >= (ta ? k : m), "lda (%lld) too small", lda);
  ISVD_CHECK(ldb >= (tb ? n : k), "ldb (%lld) too small", ldb);
  GemmArgs<T> g{};
  g.m = m; g.n = n; g.k = k; g.alpha = alpha; g.beta = beta;
  g.A = A; g.lda = lda; g.B = B; g.ldb = ldb; g.C = C; g.ldc = ldc;
  g.sA = sA; g.sB = sB; g.sC = sC; g.W = nullptr; g.upper = upper; g.ct = ct;
  if (!ta && !tb) launch<T, false, false>(g, batch);
  else if (!ta && tb) launch<T, false, true>(g, batch);
  else if (ta && !tb) launch<T, true, false>(g, batch);
  else launch<T, true, true>(g, batch);
}

template <typename T>
void gemm(char transa, char transb, i64 m, i64 n, i64 k, T alpha, const T *A, i64 lda,
          const T *B, i64 ldb, T beta, T *C, i64 ldc, i64 batch, i64 sA, i64 sB, i64 sC) {
  gemm_impl<T>(transa, transb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, batch, sA, sB, sC,
               0);
}

template <typename T>
void gemm_ct(char transa, char transb, i64 m, i64 n, i64 k, T alpha, const T *A, i64 lda,
             const T *B, i64 ldb, T beta, T *C, i64 ldc) {
  gemm_impl<T>(transa, transb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, 1, 0, 0, 0, 0, 1);
}

template <typename T>
void gemm_shortk(char transa, char transb, i64 m, i64 n, i64 k, T alpha, const T *A, i64 lda,
                 const T *B, i64 ldb, T beta, T *C, i64 ldc, i64 batch, i64 sA, i64 sB, i64 sC) {
  if (m <= 0 || n <= 0 || batch <= 0) return;
  ISVD_CHECK(k > 0, "gemm_shortk: k = %lld", k);
  GemmArgs<T> g{};
  g.m = m; g.n = n; g.k = k; g.alpha = alpha; g.beta = beta;
  g.A = A; g.lda = lda; g.B = B; g.ldb = ldb; g.C = C; g.ldc = ldc;
  g.sA = sA; g.sB = sB; g.sC = sC;
  (void)fixed64<T>(transa, transb, g, batch, 1);
}

template <typename T>
int gemm_partials(char transa, char transb, i64 m, i64 n, i64 k, const T *A, i64 lda,
                  const T *B, i64 ldb, T *W, int ksplit) {
  ISVD_CHECK(m > 0 && n > 0 && k > 0, "gemm_partials: empty product");
  GemmArgs<T> g{};
  g.m = m; g.n = n; g.k = k; g.alpha = T(1); g.beta = T(0);
  g.A = A; g.lda = lda; g.B = B; g.ldb = ldb;
  g.C = W; g.ldc = m;  // one chunk: the product itself, dense
  g.W = W;
  return fixed64<T>(transa, transb, g, 1, ksplit);
}

template void gemm_shortk<double>(char, char, i64, i64, i64, double, const double *, i64,
                                  const double *, i64, double, double *, i64, i64, i64, i64,
                                  i64);
template void gemm_shortk<float>(char, char, i64, i64, i64, float, const float *, i64,
                                 const float *, i64, float, float *, i64, i64, i64, i64, i64);
template int gemm_partials<double>(char, char, i64, i64, i64, const double *, i64,
                                   const double *, i64, double *, int);
template int gemm_partials<float>(char, char, i64, i64, i64, const float *, i64, const float *,
                                  i64, float *, int);

template <typename T>
void syrk_full(char trans, i64 n, i64 k, T alpha, const T *A, i64 lda, T *C, i64 ldc) {
  // C = alpha op(A) op(A)^T with op(A) = A (trans 'N', n x k) or A^T (trans 'T', A is k x n);
  // only tiles that touch the upper triangle are computed, then mirrored
  const char ta = (trans == 'T' || trans == 't') ? 'T' : 'N';
  const char tb = (ta == 'T') ? 'N' : 'T';
  gemm_impl<T>(ta, tb, n, n, k, alpha, A, lda, A, lda, T(0), C, ldc, 1, 0, 0, 0, 1);
  int blocks = (int)((n * n + 255) / 256);
  if (blocks > 4096) blocks = 4096;
  hipLaunchKernelGGL(mirror_upper_kernel<T>, dim3(blocks), dim3(256), 0, ctx().stream, n, C, ldc);
  LAUNCH_CHECK();
}

template void syrk_full<double>(char, i64, i64, double, const double *, i64, double *, i64);
template void syrk_full<float>(char, i64, i64, float, const float *, i64, float *, i64);

template void gemm_ct<double>(char, char, i64, i64, i64, double, const double *, i64,
                              const double *, i64, double, double *, i64);
template void gemm_ct<float>(char, char, i64, i64, i64, float, const float *, i64,
                             const float *, i64, float, float *, i64);
template void gemm<double>(char, char, i64, i64, i64, double, const double *, i64,
                           const double *, i64, double, double *, i64, i64, i64, i64, i64);
template void gemm<float>(char, char, i64, i64, i64, float, const float *, i64, const float *,
                          i64, float, float *, i64, i64, i64, i64, i64);

}  // namespace isvd
