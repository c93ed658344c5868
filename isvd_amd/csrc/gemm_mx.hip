// BF16 MFMA GEMM family for gfx950 with FP32 accumulation (v_mfma_f32_16x16x32_bf16).
//
// Workgroup tile 128 x 128, K step 64, 256 threads = 2 x 2 waves, each wave a 64 x 64 block
// of 4 x 4 MFMA tiles.  Both operands are staged in LDS as [mn][k] rows of 64 BF16 values
// (+8 padding), so every MFMA fragment (lane -> mn = lane%16, k = 8*(lane/16) .. +7) is one
// 16-byte LDS read, conflict-free across each 16-lane group (row stride 36 dwords).
//  - K-contiguous sources are copied 16 bytes at a time.
//  - MN-contiguous sources (the sketch's Omega^T and A^T, the Gram of the stacked bases) are
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

constexpr int kBM = 128, kBN = 128, kBK = 64, kLK = kBK + 8;  // LDS row length (BF16 values)

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

// One operand tile (128 mn x 64 k) in flight: global -> registers -> LDS (P parts).
// T = uint16_t (BF16 bits, P = 1) or float (split into P parts).
template <typename T, bool KC, int P>
struct Tile {
  static constexpr int kW = (int)(8 * sizeof(T) / 16);  // 16-byte words per 8 elements
  uint4 r[4][kW];

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
      // unit u: row = c / 8, k = 8 * (c % 8), c = tid + 256 u
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int c = tid + 256 * u;
        const i64 mn = mn0 + (c >> 3), kk = k0 + 8 * (c & 7);
        int valid = (mn < mnmax) ? (int)(kend - kk) : 0;
        if (valid < 0) valid = 0;
        if (valid > 8) valid = 8;
        load8(X + (valid > 0 ? kk + mn * ld : 0), valid, vec, r[u]);
      }
    } else {
      // thread: mn = 8 (tid / 16) .. +7, k = 4 (tid % 16) + j (j < 4)
      const i64 mn = mn0 + 8 * (tid >> 4);
      int valid = (int)(mnmax - mn);
      if (valid < 0) valid = 0;
      if (valid > 8) valid = 8;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const i64 kk = k0 + 4 * (tid & 15) + j;
        const int v = (kk < kend) ? valid : 0;
        load8(X + (v > 0 ? mn + kk * ld : 0), v, vec, r[j]);
      }
    }
  }

  __device__ __forceinline__ void store(uint16_t *s, int tid) const {
    if (KC) {
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int c = tid + 256 * u;
        uint16_t *dst = s + (c >> 3) * kLK + 8 * (c & 7);
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
            *reinterpret_cast<uint4 *>(dst + p * kBM * kLK) = w;
          }
        }
      }
    } else {
      const int row0 = 8 * (tid >> 4), kq = 4 * (tid & 15);
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        uint16_t *dst = s + (row0 + e) * kLK + kq;
        if (sizeof(T) == 2) {
          uint2 w;
          w.x = elem_h(r[0], e) | (elem_h(r[1], e) << 16);
          w.y = elem_h(r[2], e) | (elem_h(r[3], e) << 16);
          *reinterpret_cast<uint2 *>(dst) = w;
        } else {
          uint32_t h[4][P];
#pragma unroll
          for (int j = 0; j < 4; ++j) split_bf16<P>(elem_f(r[j], e), h[j]);
#pragma unroll
          for (int p = 0; p < P; ++p) {
            uint2 w;
            w.x = h[0][p] | (h[1][p] << 16);
            w.y = h[2][p] | (h[3][p] << 16);
            *reinterpret_cast<uint2 *>(dst + p * kBM * kLK) = w;
          }
        }
      }
    }
  }
};

__device__ __forceinline__ bf16x8 frag(const uint16_t *s) {
  return __builtin_bit_cast(bf16x8, *reinterpret_cast<const uint4 *>(s));
}

template <typename TA, bool AK, int PA, typename TB, bool BK, int PB>
__global__ __launch_bounds__(256) void gemm_mx_kernel(MxArgs g) {
  constexpr int PMAX = PA > PB ? PA : PB;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  uint16_t *sA = reinterpret_cast<uint16_t *>(smem_raw);  // [PA][kBM][kLK]
  uint16_t *sB = sA + PA * kBM * kLK;                      // [PB][kBN][kLK]
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave & 1, wn = wave >> 1;

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

  Tile<TA, AK, PA> ta;
  Tile<TB, BK, PB> tb;
  f4v acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = f4v{0.f, 0.f, 0.f, 0.f};

  if (nk > 0) {
    ta.load(A, g.lda, m0, g.m, kbeg, kend, g.avec, tid);
    tb.load(B, g.ldb, n0, g.n, kbeg, kend, g.bvec, tid);
  }
  const int fr = lane & 15, fk = 8 * (lane >> 4);
  for (int kt = 0; kt < nk; ++kt) {
    ta.store(sA, tid);
    tb.store(sB, tid);
    __syncthreads();
    if (kt + 1 < nk) {
      const i64 k0 = kbeg + (i64)(kt + 1) * kBK;
      ta.load(A, g.lda, m0, g.m, k0, kend, g.avec, tid);
      tb.load(B, g.ldb, n0, g.n, k0, kend, g.bvec, tid);
    }
#pragma unroll
    for (int s = 0; s < kBK / 32; ++s) {
      bf16x8 bf[PB][4];
#pragma unroll
      for (int p = 0; p < PB; ++p)
#pragma unroll
        for (int tj = 0; tj < 4; ++tj)
          bf[p][tj] = frag(sB + p * kBN * kLK + (wn * 64 + tj * 16 + fr) * kLK + 32 * s + fk);
#pragma unroll
      for (int ti = 0; ti < 4; ++ti) {
        bf16x8 af[PA];
#pragma unroll
        for (int p = 0; p < PA; ++p)
          af[p] = frag(sA + p * kBM * kLK + (wm * 64 + ti * 16 + fr) * kLK + 32 * s + fk);
#pragma unroll
        for (int tj = 0; tj < 4; ++tj)
#pragma unroll
          for (int ia = 0; ia < PA; ++ia)
#pragma unroll
            for (int ib = 0; ib < PB; ++ib)
              if (ia + ib < PMAX)
                acc[ti][tj] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[ia], bf[ib][tj],
                                                                       acc[ti][tj], 0, 0, 0);
      }
    }
    __syncthreads();
  }

  // ---- epilogue: D register q of a lane holds row 4*(lane/16) + q, column lane%16
  float *C = g.C + batch_id * g.sC;
  float *W = (g.ksplit > 1) ? g.W + ((i64)batch_id * g.ksplit + ks) * g.m * g.n : nullptr;
#pragma unroll
  for (int ti = 0; ti < 4; ++ti)
#pragma unroll
    for (int tj = 0; tj < 4; ++tj)
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const i64 row = m0 + wm * 64 + ti * 16 + 4 * (lane >> 4) + q;
        const i64 col = n0 + wn * 64 + tj * 16 + fr;
        if (row < g.m && col < g.n) {
          if (W) {
            W[row + col * g.m] = acc[ti][tj][q];
          } else {
            float v = g.alpha * acc[ti][tj][q];
            if (g.beta != 0.f) v += g.beta * C[row + col * g.ldc];
            C[row + col * g.ldc] = v;
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

template <typename TA, bool AK, int PA, typename TB, bool BK, int PB>
void launch_mx(MxArgs g, i64 batch) {
  auto kern = gemm_mx_kernel<TA, AK, PA, TB, BK, PB>;
  const int lds = (PA + PB) * kBM * kLK * 2;
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
  int ksplit = 1;
  if (tiles * batch < kTargetBlocks && g.k >= 8 * kBK) {
    i64 want = (kTargetBlocks + tiles * batch - 1) / (tiles * batch);
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
  DBuf<float> work;
  if (ksplit > 1) {
    work.alloc(batch * ksplit * g.m * g.n);
    g.W = work.get();
  }
  ISVD_CHECK(batch * ksplit <= 65535, "grid.y too large (batch %lld x ksplit %d)", batch, ksplit);
  ISVD_CHECK(tiles <= 0x7fffffff, "too many tiles (%lld)", tiles);
  hipLaunchKernelGGL(kern, dim3((unsigned)tiles, (unsigned)(batch * ksplit)), dim3(256), lds,
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
  // the operand combinations of the mixed pipeline
  if (!a.f32 && !a.kc && !b.f32 && !b.kc)
    launch_mx<h16, false, 1, h16, false, 1>(g, batch);  // sketch, A column-major
  else if (!a.f32 && !a.kc && !b.f32 && b.kc)
    launch_mx<h16, false, 1, h16, true, 1>(g, batch);   // sketch, A row-major
  else if (a.f32 && !a.kc && a.parts == 2 && b.f32 && !b.kc && b.parts == 2)
    launch_mx<float, false, 2, float, false, 2>(g, batch);  // Gram of the stacked bases
  else if (a.f32 && !a.kc && a.parts == 3 && !b.f32 && b.kc)
    launch_mx<float, false, 3, h16, true, 1>(g, batch);   // Qbar^T A, A column-major
  else if (a.f32 && !a.kc && a.parts == 3 && !b.f32 && !b.kc)
    launch_mx<float, false, 3, h16, false, 1>(g, batch);  // Qbar^T A, A row-major
  else
    fail("gemm_mx: operand combination not built");
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
