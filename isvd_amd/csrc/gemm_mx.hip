// BF16 MFMA GEMM family for gfx950 with FP32 accumulation (v_mfma_f32_16x16x32_bf16).
//
// Used for the two fused split products of the mixed pipeline (the plain BF16 sketch goes
// through hipBLASLt): the Gram of the stacked FP32 bases (256 x 128 tiles, 8 waves of 64 x 64)
// and the projection Qbar^T A (128 x 128 tiles, 4 waves of 64 x 64); K step 64.  Both operands
// are staged in LDS as [mn][k] rows of 64 BF16 values (+8 padding), so every MFMA fragment
// (lane -> mn = lane%16, k = 8*(lane/16) .. +7) is one 16-byte LDS read.  Two LDS stages when
// they fit (global loads run two K steps ahead, one barrier per step), one otherwise.
//  - K-contiguous sources are copied 16 bytes at a time.
//  - MN-contiguous sources (the Gram of the stacked bases, Qbar^T of the projection) are
//    loaded as 8 consecutive mn at 4 consecutive k per thread and transposed in registers into
//    8-byte [mn][k..k+3] LDS writes; lanes 0..15 of a wave write one 128-byte LDS row span, so
//    the transpose writes are conflict-free too.
//  - FP32 sources are split into BF16 terms while they are written to LDS.
// Global loads for the next K step are issued before the MFMAs of the current one (register
// prefetch), with the tile raster walked XCD-contiguously as in the FP64/FP32 kernel.
#include "mx.hpp"

#include <mutex>

namespace isvd {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f4v __attribute__((ext_vector_type(4)));

namespace {

constexpr int kBM = 128, kBN = 128;

struct MxArgs {
  i64 m, n, k;
  const void *A;
  i64 lda, sA;
  const void *B;
  i64 ldb, sB;
  float *C;
  i64 ldc, sC;
  float alpha, beta;
  float *W;  // split-K partials: [batch][ksplit][m*n], dense column-major
  int ksplit;
  i64 kchunk;
  int group_m, upper;
  int avec, bvec;  // 16-byte global loads allowed for the operand
};

__device__ __forceinline__ uint32_t bf16_rne(float x) {
  const uint32_t u = __float_as_uint(x);
  return (u + 0x7FFFu + ((u >> 16) & 1u)) >> 16;  // finite inputs
}
__device__ __forceinline__ float bf16_val(uint32_t h) { return __uint_as_float(h << 16); }

template <int P>
__device__ __forceinline__ void split_bf16(float x, uint32_t (&h)[P]) {
  float r = x;
#pragma unroll
  for (int p = 0; p < P; ++p) {
    h[p] = bf16_rne(r);
    r -= bf16_val(h[p]);
  }
}

// One operand tile (R mn x KT k) in flight: global -> registers -> LDS (P parts).
// T = uint16_t (BF16 bits, P = 1) or float (split into P parts).  NT threads; each thread
// moves NL runs of 8 consecutive elements.
template <typename T, bool KC, int P, int R, int KT, int NT>
struct Tile {
  static constexpr int kW = (int)(8 * sizeof(T) / 16);  // 16-byte words per 8 elements
  static constexpr int kLK = KT + 8;                     // LDS row length (BF16 values)
  // K-contiguous: RUNS runs of 8 k; MN-contiguous: UNITS units of 8 mn x 4 k
  static constexpr int RUNS = R * KT / 8, UNITS = R * KT / 32;
  static constexpr int NL = KC ? (RUNS + NT - 1) / NT : 4 * ((UNITS + NT - 1) / NT);
  static_assert(KC ? RUNS % NT == 0 : (UNITS % NT == 0 || UNITS < NT), "tile / thread mismatch");
  uint4 r[NL][kW];

  __device__ __forceinline__ static float elem_f(const uint4 *w, int e) {
    const uint4 q = w[e >> 2];
    const uint32_t v = (e & 3) == 0 ? q.x : (e & 3) == 1 ? q.y : (e & 3) == 2 ? q.z : q.w;
    return __uint_as_float(v);
  }
  __device__ __forceinline__ static uint32_t elem_h(const uint4 *w, int e) {
    const uint4 q = w[0];
    const int d = e >> 1;
    const uint32_t v = d == 0 ? q.x : d == 1 ? q.y : d == 2 ? q.z : q.w;
    return (e & 1) ? (v >> 16) : (v & 0xFFFFu);
  }

  // 8 consecutive elements starting at p (stride 1); zero past `valid`
  __device__ __forceinline__ static void load8(const T *p, int valid, bool vec, uint4 *w) {
    if (vec && valid >= 8) {
      const uint4 *q = reinterpret_cast<const uint4 *>(p);
#pragma unroll
      for (int i = 0; i < kW; ++i) w[i] = q[i];
      return;
    }
    T v[8];
#pragma unroll
    for (int e = 0; e < 8; ++e) v[e] = (e < valid) ? p[e] : T(0);
    const uint4 *q = reinterpret_cast<const uint4 *>(v);
#pragma unroll
    for (int i = 0; i < kW; ++i) w[i] = q[i];
  }

  __device__ __forceinline__ void load(const T *X, i64 ld, i64 mn0, i64 mnmax, i64 k0, i64 kend,
                                       bool vec, int tid) {
    if (KC) {
      // run u: row c / (KT/8), k 8 (c % (KT/8)) .. +7, c = tid + NT u
#pragma unroll
      for (int u = 0; u < NL; ++u) {
        const int c = tid + NT * u;
        const i64 mn = mn0 + c / (KT / 8), kk = k0 + 8 * (c % (KT / 8));
        int valid = (mn < mnmax) ? (int)(kend - kk) : 0;
        if (valid < 0) valid = 0;
        if (valid > 8) valid = 8;
        load8(X + (valid > 0 ? kk + mn * ld : 0), valid, vec, r[u]);
      }
    } else {
      // unit c = tid + NT uu: mn 8 (c / (KT/4)) .. +7 at k 4 (c % (KT/4)) + j, j < 4
#pragma unroll
      for (int uu = 0; uu < NL / 4; ++uu) {
        const int c = tid + NT * uu;
        if (c >= UNITS) break;
        const i64 mn = mn0 + 8 * (c / (KT / 4));
        int valid = (int)(mnmax - mn);
        if (valid < 0) valid = 0;
        if (valid > 8) valid = 8;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const i64 kk = k0 + 4 * (c % (KT / 4)) + j;
          const int v = (kk < kend) ? valid : 0;
          load8(X + (v > 0 ? mn + kk * ld : 0), v, vec, r[4 * uu + j]);
        }
      }
    }
  }

  __device__ __forceinline__ void store(uint16_t *s, int tid) const {
    if (KC) {
#pragma unroll
      for (int u = 0; u < NL; ++u) {
        const int c = tid + NT * u;
        uint16_t *dst = s + (c / (KT / 8)) * kLK + 8 * (c % (KT / 8));
        if (sizeof(T) == 2) {
          *reinterpret_cast<uint4 *>(dst) = r[u][0];
        } else {
          uint32_t h[8][P];
#pragma unroll
          for (int e = 0; e < 8; ++e) split_bf16<P>(elem_f(r[u], e), h[e]);
#pragma unroll
          for (int p = 0; p < P; ++p) {
            uint4 w;
            w.x = h[0][p] | (h[1][p] << 16);
            w.y = h[2][p] | (h[3][p] << 16);
            w.z = h[4][p] | (h[5][p] << 16);
            w.w = h[6][p] | (h[7][p] << 16);
            *reinterpret_cast<uint4 *>(dst + p * R * kLK) = w;
          }
        }
      }
    } else {
      // lanes with consecutive k groups write one contiguous LDS row span (conflict-free)
#pragma unroll
      for (int uu = 0; uu < NL / 4; ++uu) {
        const int c = tid + NT * uu;
        if (c >= UNITS) break;
        const int row0 = 8 * (c / (KT / 4)), kq = 4 * (c % (KT / 4));
        const uint4(*q)[kW] = r + 4 * uu;
#pragma unroll
        for (int e = 0; e < 8; ++e) {
          uint16_t *dst = s + (row0 + e) * kLK + kq;
          if (sizeof(T) == 2) {
            uint2 w;
            w.x = elem_h(q[0], e) | (elem_h(q[1], e) << 16);
            w.y = elem_h(q[2], e) | (elem_h(q[3], e) << 16);
            *reinterpret_cast<uint2 *>(dst) = w;
          } else {
            uint32_t h[4][P];
#pragma unroll
            for (int j = 0; j < 4; ++j) split_bf16<P>(elem_f(q[j], e), h[j]);
#pragma unroll
            for (int p = 0; p < P; ++p) {
              uint2 w;
              w.x = h[0][p] | (h[1][p] << 16);
              w.y = h[2][p] | (h[3][p] << 16);
              *reinterpret_cast<uint2 *>(dst + p * R * kLK) = w;
            }
          }
        }
      }
    }
  }
};

__device__ __forceinline__ bf16x8 frag(const uint16_t *s) {
  return __builtin_bit_cast(bf16x8, *reinterpret_cast<const uint4 *>(s));
}

// Workgroup tile BM x BN, K step KT, WGM x WGN waves (each a (BM/WGM) x (BN/WGN) block of
// 16 x 16 MFMA tiles).
template <int BM_, int BN_, int WGM_, int WGN_, int KT_>
struct Cfg {
  static constexpr int BM = BM_, BN = BN_, WGM = WGM_, WGN = WGN_, KT = KT_;
  static constexpr int NT = 64 * WGM * WGN;
  static constexpr int FM = BM / WGM / 16, FN = BN / WGN / 16;
  // LDS bytes of one stage for PA / PB BF16 parts per operand
  template <int QA, int QB>
  static constexpr int stage_bytes() {
    return (QA * BM + QB * BN) * (KT + 8) * 2;
  }
  template <int QA, int QB>
  static constexpr bool kTwoStages() {
    return 2 * stage_bytes<QA, QB>() <= 160 * 1024;
  }
};
using CfgTall = Cfg<256, 128, 4, 2, 64>;  // 8 waves of 64 x 64
using Cfg128 = Cfg<128, 128, 2, 2, 64>;  // 4 waves of 64 x 64

template <typename TA, bool AK, int PA, typename TB, bool BK, int PB, class C>
__global__ __launch_bounds__(C::NT) void gemm_mx_kernel(MxArgs g) {
  constexpr int PMAX = PA > PB ? PA : PB;
  constexpr int kBM = C::BM, kBN = C::BN, kBK = C::KT, kLK = C::KT + 8;
  constexpr int FM = C::FM, FN = C::FN;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  // one LDS stage: [PA][kBM][kLK] then [PB][kBN][kLK]; two stages when they fit
  constexpr int kStage = PA * kBM * kLK + PB * kBN * kLK;
  constexpr bool kTwo = C::template kTwoStages<PA, PB>();
  uint16_t *const s0 = reinterpret_cast<uint16_t *>(smem_raw);
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave % C::WGM, wn = wave / C::WGM;

  // ---- tile of this workgroup (XCD-contiguous; symmetric products walk the upper tiles)
  const i64 tiles_m = (g.m + kBM - 1) / kBM, tiles_n = (g.n + kBN - 1) / kBN;
  auto upper_count = [&](i64 tn) {
    const i64 c = ((tn + 1) * kBN + kBM - 1) / kBM;
    return c < tiles_m ? c : tiles_m;
  };
  i64 total = tiles_m * tiles_n;
  if (g.upper) {
    total = 0;
    for (i64 q = 0; q < tiles_n; ++q) total += upper_count(q);
  }
  i64 pid = blockIdx.x;
  if (pid >= total) return;
  {
    constexpr i64 kXcd = 8;
    const i64 xcd = pid % kXcd, local = pid / kXcd;
    const i64 per = total / kXcd, rem = total % kXcd;
    pid = (xcd < rem) ? xcd * (per + 1) + local : rem * (per + 1) + (xcd - rem) * per + local;
  }
  i64 tm, tn;
  if (g.upper) {
    tn = 0;
    while (pid >= upper_count(tn)) pid -= upper_count(tn++);
    tm = pid;
  } else {
    const i64 in_group = (i64)g.group_m * tiles_n;
    const i64 first_m = (pid / in_group) * g.group_m;
    const i64 gsz = (tiles_m - first_m) < g.group_m ? (tiles_m - first_m) : g.group_m;
    tm = first_m + (pid % in_group) % gsz;
    tn = (pid % in_group) / gsz;
  }
  const i64 m0 = tm * kBM, n0 = tn * kBN;
  const int batch_id = blockIdx.y / g.ksplit, ks = blockIdx.y % g.ksplit;
  const TA *A = reinterpret_cast<const TA *>(g.A) + batch_id * g.sA;
  const TB *B = reinterpret_cast<const TB *>(g.B) + batch_id * g.sB;
  const i64 kbeg = ks * g.kchunk;
  const i64 kend = (kbeg + g.kchunk < g.k) ? kbeg + g.kchunk : g.k;
  const int nk = (kend > kbeg) ? (int)((kend - kbeg + kBK - 1) / kBK) : 0;

  Tile<TA, AK, PA, kBM, kBK, C::NT> ta;
  Tile<TB, BK, PB, kBN, kBK, C::NT> tb;
  f4v acc[FM][FN];
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) acc[i][j] = f4v{0.f, 0.f, 0.f, 0.f};

  const int fr = lane & 15, fk = 8 * (lane >> 4);
  const int wrow = wm * (kBM / C::WGM), wcol = wn * (kBN / C::WGN);
  auto compute = [&](const uint16_t *sA, const uint16_t *sB) {
#pragma unroll
    for (int s = 0; s < kBK / 32; ++s) {
      bf16x8 bf[PB][FN];
#pragma unroll
      for (int p = 0; p < PB; ++p)
#pragma unroll
        for (int tj = 0; tj < FN; ++tj)
          bf[p][tj] = frag(sB + p * kBN * kLK + (wcol + tj * 16 + fr) * kLK + 32 * s + fk);
#pragma unroll
      for (int ti = 0; ti < FM; ++ti) {
        bf16x8 af[PA];
#pragma unroll
        for (int p = 0; p < PA; ++p)
          af[p] = frag(sA + p * kBM * kLK + (wrow + ti * 16 + fr) * kLK + 32 * s + fk);
#pragma unroll
        for (int tj = 0; tj < FN; ++tj)
#pragma unroll
          for (int ia = 0; ia < PA; ++ia)
#pragma unroll
            for (int ib = 0; ib < PB; ++ib)
              if (ia + ib < PMAX)
                acc[ti][tj] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[ia], bf[ib][tj],
                                                                       acc[ti][tj], 0, 0, 0);
      }
    }
  };
  auto load = [&](int kt) {
    const i64 k0 = kbeg + (i64)kt * kBK;
    ta.load(A, g.lda, m0, g.m, k0, kend, g.avec, tid);
    tb.load(B, g.ldb, n0, g.n, k0, kend, g.bvec, tid);
  };
  auto store = [&](uint16_t *st) {
    ta.store(st, tid);
    tb.store(st + PA * kBM * kLK, tid);
  };
  if (kTwo) {
    // two LDS stages: the global loads of step kt + 2 are issued while step kt computes, and
    // one barrier per step separates the stage being read from the stage being written
    if (nk > 0) {
      load(0);
      store(s0);
    }
    __syncthreads();
    if (nk > 1) load(1);
    for (int kt = 0; kt < nk; ++kt) {
      const uint16_t *cur = s0 + (kt & 1) * kStage;
      compute(cur, cur + PA * kBM * kLK);
      if (kt + 1 < nk) store(s0 + ((kt + 1) & 1) * kStage);
      if (kt + 2 < nk) load(kt + 2);
      __syncthreads();
    }
  } else {
    if (nk > 0) load(0);
    for (int kt = 0; kt < nk; ++kt) {
      store(s0);
      __syncthreads();
      if (kt + 1 < nk) load(kt + 1);
      compute(s0, s0 + PA * kBM * kLK);
      __syncthreads();
    }
  }

  // ---- epilogue: D register q of a lane holds row 4*(lane/16) + q, column lane%16
  float *Cm = g.C + batch_id * g.sC;
  float *W = (g.ksplit > 1) ? g.W + ((i64)batch_id * g.ksplit + ks) * g.m * g.n : nullptr;
#pragma unroll
  for (int ti = 0; ti < FM; ++ti)
#pragma unroll
    for (int tj = 0; tj < FN; ++tj)
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const i64 row = m0 + wrow + ti * 16 + 4 * (lane >> 4) + q;
        const i64 col = n0 + wcol + tj * 16 + fr;
        if (row < g.m && col < g.n) {
          if (W) {
            W[row + col * g.m] = acc[ti][tj][q];
          } else {
            float v = g.alpha * acc[ti][tj][q];
            if (g.beta != 0.f) v += g.beta * Cm[row + col * g.ldc];
            Cm[row + col * g.ldc] = v;
          }
        }
      }
}

__global__ void mx_reduce_kernel(i64 m, i64 n, int ksplit, const float *W, float alpha,
                                 float beta, float *C, i64 ldc, i64 sC) {
  const i64 b = blockIdx.y, mn = m * n;
  const float *Wb = W + b * ksplit * mn;
  float *Cb = C + b * sC;
  for (i64 e = blockIdx.x * (i64)blockDim.x + threadIdx.x; e < mn;
       e += (i64)gridDim.x * blockDim.x) {
    float s = 0.f;
    for (int q = 0; q < ksplit; ++q) s += Wb[q * mn + e];  // fixed order: deterministic
    const i64 r = e % m, c = e / m;
    float v = alpha * s;
    if (beta != 0.f) v += beta * Cb[r + c * ldc];
    Cb[r + c * ldc] = v;
  }
}

__global__ void mx_mirror_kernel(i64 n, float *C, i64 ldc) {
  for (i64 e = blockIdx.x * (i64)blockDim.x + threadIdx.x; e < n * n;
       e += (i64)gridDim.x * blockDim.x) {
    const i64 i = e % n, j = e / n;
    if (i > j) C[i + j * ldc] = C[j + i * ldc];
  }
}

template <typename Tout, typename Tin>
__global__ void convert_kernel(i64 rows, i64 cols, const Tin *src, i64 lds, Tout *dst, i64 ldd) {
  for (i64 e = blockIdx.x * (i64)blockDim.x + threadIdx.x; e < rows * cols;
       e += (i64)gridDim.x * blockDim.x) {
    const i64 i = e % rows, j = e / rows;
    dst[i + j * ldd] = (Tout)src[i + j * lds];
  }
}

__global__ void f32_to_bf16_kernel(i64 count, const float *src, uint16_t *dst) {
  for (i64 e = blockIdx.x * (i64)blockDim.x + threadIdx.x; e < count;
       e += (i64)gridDim.x * blockDim.x)
    dst[e] = (uint16_t)bf16_rne(src[e]);
}

int grid_of(i64 n) {
  i64 b = (n + 255) / 256;
  return (int)(b < 1 ? 1 : (b > 8192 ? 8192 : b));
}

constexpr int kTargetBlocks = 1024;

template <typename TA, bool AK, int PA, typename TB, bool BK, int PB, class C>
void launch_mx(MxArgs g, i64 batch) {
  constexpr int kBM = C::BM, kBN = C::BN, kBK = C::KT;
  auto kern = gemm_mx_kernel<TA, AK, PA, TB, BK, PB, C>;
  const int lds = C::template stage_bytes<PA, PB>() * (C::template kTwoStages<PA, PB>() ? 2 : 1);
  static std::once_flag once;
  std::call_once(once, [&] {
    HIP_CHECK(hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  lds));
  });
  const i64 tiles_m = (g.m + kBM - 1) / kBM, tiles_n = (g.n + kBN - 1) / kBN;
  i64 tiles = tiles_m * tiles_n;
  if (g.upper) {
    tiles = 0;
    for (i64 q = 0; q < tiles_n; ++q) {
      const i64 c = ((q + 1) * kBN + kBM - 1) / kBM;
      tiles += c < tiles_m ? c : tiles_m;
    }
  }
  // split K until the launch has ~4 workgroups of 256 threads per CU
  const i64 target = kTargetBlocks * 256 / C::NT;
  int ksplit = 1;
  if (tiles * batch < target && g.k >= 8 * kBK) {
    i64 want = (target + tiles * batch - 1) / (tiles * batch);
    const i64 maxs = g.k / (4 * kBK);
    if (want > maxs) want = maxs;
    if (want > 64) want = 64;
    if (want < 1) want = 1;
    ksplit = (int)want;
  }
  i64 kchunk = (g.k + ksplit - 1) / ksplit;
  kchunk = (kchunk + kBK - 1) / kBK * kBK;
  ksplit = (int)((g.k + kchunk - 1) / kchunk);
  if (ksplit < 1) ksplit = 1;
  g.ksplit = ksplit;
  g.kchunk = (ksplit == 1) ? (g.k > 0 ? g.k : 1) : kchunk;
  g.group_m = 8;
  if (ksplit > 1)
    g.W = static_cast<float *>(
        splitk_workspace(sizeof(float) * (size_t)(batch * ksplit * g.m * g.n)));
  ISVD_CHECK(batch * ksplit <= 65535, "grid.y too large (batch %lld x ksplit %d)", batch, ksplit);
  ISVD_CHECK(tiles <= 0x7fffffff, "too many tiles (%lld)", tiles);
  hipLaunchKernelGGL(kern, dim3((unsigned)tiles, (unsigned)(batch * ksplit)), dim3(C::NT), lds,
                     ctx().stream, g);
  LAUNCH_CHECK();
  if (ksplit > 1) {
    hipLaunchKernelGGL(mx_reduce_kernel, dim3(grid_of(g.m * g.n), (unsigned)batch), dim3(256), 0,
                       ctx().stream, g.m, g.n, ksplit, (const float *)g.W, g.alpha, g.beta, g.C,
                       g.ldc, g.sC);
    LAUNCH_CHECK();
  }
}

bool aligned16(const void *p, i64 ld, i64 stride, i64 esize) {
  const i64 per = 16 / esize;
  return (reinterpret_cast<uintptr_t>(p) % 16 == 0) && (ld % per == 0) && (stride % per == 0);
}

}  // namespace

void gemm_mx(const MxOperand &a, const MxOperand &b, i64 m, i64 n, i64 k, float alpha,
             float beta, float *C, i64 ldc, i64 batch, i64 sC, bool upper) {
  if (m <= 0 || n <= 0 || batch <= 0) return;
  ISVD_CHECK(k > 0, "gemm_mx: k must be positive");
  ISVD_CHECK(ldc >= m, "gemm_mx: ldc (%lld) < m (%lld)", ldc, m);
  ISVD_CHECK(!upper || m == n, "gemm_mx: a symmetric product must be square");
  ISVD_CHECK(a.ld >= (a.kc ? k : m) && b.ld >= (b.kc ? k : n), "gemm_mx: leading dimension too small");
  MxArgs g{};
  g.m = m; g.n = n; g.k = k;
  g.A = a.p; g.lda = a.ld; g.sA = a.stride;
  g.B = b.p; g.ldb = b.ld; g.sB = b.stride;
  g.C = C; g.ldc = ldc; g.sC = sC;
  g.alpha = alpha; g.beta = beta; g.upper = upper ? 1 : 0;
  g.avec = aligned16(a.p, a.ld, a.stride, a.f32 ? 4 : 2);
  g.bvec = aligned16(b.p, b.ld, b.stride, b.f32 ? 4 : 2);
  using h16 = uint16_t;
  // 256 x 128 tiles (8 waves) for the Nl x Nl Gram; the l x n projection (m = l <= 128) runs
  // 128 x 128 tiles
  const bool wide = (m >= 256 && n >= 256);
  if (a.f32 && !a.kc && a.parts == 2 && b.f32 && !b.kc && b.parts == 2) {
    // Gram of the stacked bases
    if (wide) launch_mx<float, false, 2, float, false, 2, CfgTall>(g, batch);
    else launch_mx<float, false, 2, float, false, 2, Cfg128>(g, batch);
  } else if (a.f32 && !a.kc && a.parts == 3 && !b.f32 && b.kc) {
    launch_mx<float, false, 3, h16, true, 1, Cfg128>(g, batch);   // Qbar^T A, A column-major
  } else if (a.f32 && !a.kc && a.parts == 3 && !b.f32 && !b.kc) {
    launch_mx<float, false, 3, h16, false, 1, Cfg128>(g, batch);  // Qbar^T A, A row-major
  } else {
    fail("gemm_mx: operand combination not built");
  }
  if (upper) {
    hipLaunchKernelGGL(mx_mirror_kernel, dim3(grid_of(n * n)), dim3(256), 0, ctx().stream, n, C,
                       ldc);
    LAUNCH_CHECK();
  }
}

template <typename Tout, typename Tin>
void convert(i64 rows, i64 cols, const Tin *src, i64 lds, Tout *dst, i64 ldd) {
  if (rows <= 0 || cols <= 0) return;
  hipLaunchKernelGGL((convert_kernel<Tout, Tin>), dim3(grid_of(rows * cols)), dim3(256), 0,
                     ctx().stream, rows, cols, src, lds, dst, ldd);
  LAUNCH_CHECK();
}

void f32_to_bf16(i64 count, const float *src, uint16_t *dst) {
  if (count <= 0) return;
  hipLaunchKernelGGL(f32_to_bf16_kernel, dim3(grid_of(count)), dim3(256), 0, ctx().stream, count,
                     src, dst);
  LAUNCH_CHECK();
}

template void convert<double, float>(i64, i64, const float *, i64, double *, i64);
template void convert<float, double>(i64, i64, const double *, i64, float *, i64);

}  // namespace isvd
