// FP8 (OCP e4m3) sketch on the gfx950 block-scaled matrix cores.
//
// Product: Y^T[r, i] = sum_j Omega[j, r] A[i, j] over j < n, with A[i, j] =
// decode(a8[i, j]) * 2^e(i / 128, j / 128).  One K step is 128 bytes: exactly one
// v_mfma_scale_f32_16x16x128_f8f6f4 per 16 x 16 output fragment, and exactly one scale block
// along K, so the block's exponent rides in the instruction's B-operand scale (E8M0) and the
// accumulation stays in the MFMA's FP32 accumulator with no vector-ALU rescaling.
//
// Operand layout of the instruction (checked on gfx950): lane l holds row l % 16 of the
// 16 x 128 A operand (bytes K = 32 (l / 16) .. +32, ascending) and column l % 16 of the B
// operand (the same K range); D[4 (l / 16) + q][l % 16] is accumulator q of lane l.  Both
// operands here are K-contiguous rows (Omega columns, A rows), so a fragment is two 16-byte
// LDS reads per lane.
#include "fp8.hpp"

#include <mutex>

#include "la.hpp"

namespace isvd {

namespace {

typedef int i32x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

// OCP e4m3 code of x, rounded to nearest even; |x| >= 448 saturates to 448, NaN -> 0x7F
__device__ inline unsigned e4m3_code(float x) {
  const unsigned bits = __float_as_uint(x);
  const unsigned sign = (bits >> 24) & 0x80u;
  const float ax = fabsf(x);
  unsigned code;
  if (!(ax == ax)) {
    code = 0x7Fu;
  } else if (ax >= 448.f) {
    code = 0x7Eu;
  } else if (ax < 0.015625f) {
    // subnormal range: m * 2^-9 with m = 0..7; m = 8 is the smallest normal (code 0x08)
    code = (unsigned)__builtin_rintf(ax * 512.f);
  } else {
    // keep 3 of the 23 FP32 mantissa bits, ties to even, exponent rebiased 127 -> 7
    const unsigned ab = bits & 0x7FFFFFFFu;
    code = ((ab + 0x7FFFFu + ((ab >> 20) & 1u)) >> 20) - (120u << 3);
  }
  return sign | code;
}

// smallest power of two s with amax <= 448 s, within [2^-127, 2^127]; 1 for amax = 0
__device__ inline float block_scale(float amax) {
  if (!(amax > 0.f)) return 1.f;
  const int e0 = ilogbf(amax);  // amax in [2^e0, 2^(e0 + 1))
  int e = (amax <= ldexpf(448.f, e0 - 8)) ? e0 - 8 : e0 - 7;
  e = e < -127 ? -127 : (e > 127 ? 127 : e);
  return ldexpf(1.f, e);
}

__global__ __launch_bounds__(256) void quantize_kernel(char ordera, i64 m, i64 n,
                                                       const float *__restrict__ a, i64 lda,
                                                       uint8_t *__restrict__ a8, i64 lda8,
                                                       float *__restrict__ scale, i64 lds) {
  __shared__ float tile[128][129];
  __shared__ float wmax[4];
  const i64 i0 = (i64)blockIdx.y * 128, j0 = (i64)blockIdx.x * 128;
  const int t = threadIdx.x;
  float amax = 0.f;
  for (int e = t; e < 128 * 128; e += 256) {
    const int fast = e & 127, slow = e >> 7;
    const int ii = ordera == 'R' ? slow : fast, jj = ordera == 'R' ? fast : slow;
    const i64 gi = i0 + ii, gj = j0 + jj;
    float v = 0.f;
    if (gi < m && gj < n) v = ordera == 'R' ? a[gi * lda + gj] : a[gi + gj * lda];
    tile[ii][jj] = v;
    amax = fmaxf(amax, fabsf(v));
  }
  for (int o = 32; o > 0; o >>= 1) amax = fmaxf(amax, __shfl_xor(amax, o));
  if ((t & 63) == 0) wmax[t >> 6] = amax;
  __syncthreads();
  amax = fmaxf(fmaxf(wmax[0], wmax[1]), fmaxf(wmax[2], wmax[3]));
  const float s = block_scale(amax), inv = 1.f / s;
  if (t == 0) scale[(i64)blockIdx.y * lds + blockIdx.x] = s;
  // 128 rows x 8 chunks of 16 codes; columns >= n hold zeros (code 0)
  for (int q = t; q < 128 * 8; q += 256) {
    const int ii = q >> 3, c = q & 7;
    const i64 gi = i0 + ii, gj = j0 + 16 * c;
    if (gi >= m || gj >= n) continue;
    u32x4 w;
    for (int word = 0; word < 4; ++word) {
      unsigned v = 0;
      for (int b = 0; b < 4; ++b) v |= e4m3_code(tile[ii][16 * c + 4 * word + b] * inv) << (8 * b);
      w[word] = v;
    }
    *reinterpret_cast<u32x4 *>(a8 + gi * lda8 + gj) = w;
  }
}

// omega8[j + c * ldo] := e4m3(src[c + j * Nl])  (64 x 64 tiles through LDS)
__global__ __launch_bounds__(256) void omega_e4m3_kernel(i64 n, i64 Nl, const float *__restrict__ src,
                                                         uint8_t *__restrict__ omega8, i64 ldo) {
  __shared__ float tile[64][65];  // [j][c]
  const i64 c0 = (i64)blockIdx.x * 64, j0 = (i64)blockIdx.y * 64;
  const int t = threadIdx.x;
  for (int e = t; e < 64 * 64; e += 256) {
    const int cc = e & 63, jj = e >> 6;
    const i64 c = c0 + cc, j = j0 + jj;
    tile[jj][cc] = (c < Nl && j < n) ? src[c + j * Nl] : 0.f;
  }
  __syncthreads();
  for (int e = t; e < 64 * 64; e += 256) {
    const int jj = e & 63, cc = e >> 6;
    const i64 c = c0 + cc, j = j0 + jj;
    if (c < Nl && j < n) omega8[j + c * ldo] = (uint8_t)e4m3_code(tile[jj][cc]);
  }
}

// E8M0 exponents of the block scales; *bad := 1 when a scale is not a power of two in
// [2^-127, 2^127]
__global__ void scale_e8m0_kernel(i64 nbi, i64 nbj, const float *__restrict__ scale, i64 lds,
                                  int *__restrict__ e8, int *__restrict__ bad) {
  for (i64 e = blockIdx.x * (i64)blockDim.x + threadIdx.x; e < nbi * nbj;
       e += (i64)gridDim.x * blockDim.x) {
    const i64 bi = e / nbj, bj = e % nbj;
    const unsigned b = __float_as_uint(scale[bi * lds + bj]);
    const unsigned ex = b >> 23;  // sign bit included: negative values fail below
    int code = 127;
    if ((b & 0x7FFFFFu) == 0 && ex >= 1 && ex <= 254) code = (int)ex;  // normal 2^(ex - 127)
    else if (b == 0x00400000u) code = 0;                              // 2^-127
    else *bad = 1;
    e8[e] = code;
  }
}

// ---------------------------------------------------------------------------------------------
// the block-scaled GEMM
// ---------------------------------------------------------------------------------------------
template <int BR_, int BI_, int WR_, int WI_>
struct F8Cfg {
  static constexpr int BR = BR_, BI = BI_, WR = WR_, WI = WI_;
  static constexpr int NT = 64 * WR * WI;
  static constexpr int TR = BR / WR, TI = BI / WI;  // wave tile
  static constexpr int FR = TR / 16, FI = TI / 16;  // 16 x 16 fragments per wave
  static constexpr int ROWS = BR + BI;              // LDS rows of 128 bytes per stage
  static constexpr int STAGE = ROWS * 128;
  static constexpr int LOADS = ROWS * 8 / NT;       // 16-byte chunks per thread per stage
  static_assert(TR % 16 == 0 && TI % 16 == 0, "wave tile must be a multiple of 16");
  static_assert(128 % TI == 0, "a wave's A rows must lie in one scale block");
  static_assert((ROWS * 8) % NT == 0, "chunks must divide evenly");
};

using F8Main = F8Cfg<128, 256, 1, 4>;

struct F8Args {
  i64 Nl, m, n;
  const uint8_t *om;  // Omega columns (K contiguous), ld ldo
  i64 ldo;
  const uint8_t *a8;  // A rows (K contiguous), ld lda
  i64 lda;
  const int *e8;      // E8M0 block exponents, nbj per block row
  i64 nbj;
  float *y;           // Y^T column-major, ld ldy
  i64 ldy;
};

// byte offset of 16-byte chunk c of LDS row `row` (chunks XOR-swizzled against bank conflicts
// of the 16-row fragment reads)
__device__ inline int lds_off(int row, int c) { return row * 128 + ((c ^ ((row >> 1) & 7)) << 4); }

template <class C>
__global__ __launch_bounds__(C::NT) void f8_sketch_kernel(F8Args g) {
  extern __shared__ __align__(16) unsigned char smem[];
  const i64 nR = (g.Nl + C::BR - 1) / C::BR, nI = (g.m + C::BI - 1) / C::BI;
  const i64 total = nR * nI;
  i64 pid = blockIdx.x;
  if (pid >= total) return;
  {
    // consecutive tiles on one XCD (blocks are dealt round-robin to the 8 XCDs) share the
    // A rows: the Omega tile index runs fastest
    constexpr i64 kXcd = 8;
    const i64 xcd = pid % kXcd, local = pid / kXcd;
    const i64 per = total / kXcd, rem = total % kXcd;
    pid = (xcd < rem) ? xcd * (per + 1) + local : rem * (per + 1) + (xcd - rem) * per + local;
  }
  const i64 r0 = (pid % nR) * C::BR, i0 = (pid / nR) * C::BI;
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int wr = wave % C::WR, wi = wave / C::WR;
  const i64 nk = (g.n + 127) / 128;

  // ---- global -> register staging of one K step
  u32x4 stage_regs[C::LOADS];
  auto load_global = [&](i64 k0) {
#pragma unroll
    for (int u = 0; u < C::LOADS; ++u) {
      const int idx = t + u * C::NT, row = idx >> 3, c = idx & 7;
      const i64 kk = k0 + 16 * c;
      const uint8_t *src;
      bool ok;
      if (row < C::BR) {
        ok = (r0 + row < g.Nl);
        src = g.om + (r0 + row) * g.ldo + kk;
      } else {
        ok = (i0 + row - C::BR < g.m);
        src = g.a8 + (i0 + row - C::BR) * g.lda + kk;
      }
      u32x4 v = {0u, 0u, 0u, 0u};
      if (ok && kk < g.n) {
        v = *reinterpret_cast<const u32x4 *>(src);
        const i64 rem = g.n - kk;  // bytes of this chunk inside the matrix
        if (rem < 16) {
#pragma unroll
          for (int w = 0; w < 4; ++w) {
            const i64 keep = rem - 4 * w;
            const unsigned mask = keep >= 4 ? 0xFFFFFFFFu : (keep <= 0 ? 0u : ((1u << (8 * keep)) - 1u));
            v[w] &= mask;
          }
        }
      }
      stage_regs[u] = v;
    }
  };
  auto store_lds = [&](int s) {
    unsigned char *base = smem + s * C::STAGE;
#pragma unroll
    for (int u = 0; u < C::LOADS; ++u) {
      const int idx = t + u * C::NT, row = idx >> 3, c = idx & 7;
      *reinterpret_cast<u32x4 *>(base + lds_off(row, c)) = stage_regs[u];
    }
  };

  f32x4 acc[C::FR][C::FI];
#pragma unroll
  for (int a = 0; a < C::FR; ++a)
#pragma unroll
    for (int b = 0; b < C::FI; ++b) acc[a][b] = f32x4{0.f, 0.f, 0.f, 0.f};

  const int frow = lane & 15, fc = 2 * (lane >> 4);
  const i64 bi = (i0 + wi * C::TI) / 128;
  const int *e8row = g.e8 + bi * g.nbj;

  load_global(0);
  store_lds(0);
  __syncthreads();
  for (i64 ks = 0; ks < nk; ++ks) {
    const int s = (int)(ks & 1);
    if (ks + 1 < nk) load_global((ks + 1) * 128);
    const unsigned char *sO = smem + s * C::STAGE;
    const unsigned char *sA = sO + C::BR * 128;
    const int sc = e8row[ks];
    i32x8 bf[C::FI];
#pragma unroll
    for (int b = 0; b < C::FI; ++b) {
      const int row = wi * C::TI + 16 * b + frow;
      const u32x4 lo = *reinterpret_cast<const u32x4 *>(sA + lds_off(row, fc));
      const u32x4 hi = *reinterpret_cast<const u32x4 *>(sA + lds_off(row, fc + 1));
      bf[b] = i32x8{(int)lo[0], (int)lo[1], (int)lo[2], (int)lo[3],
                    (int)hi[0], (int)hi[1], (int)hi[2], (int)hi[3]};
    }
#pragma unroll
    for (int a = 0; a < C::FR; ++a) {
      const int row = wr * C::TR + 16 * a + frow;
      const u32x4 lo = *reinterpret_cast<const u32x4 *>(sO + lds_off(row, fc));
      const u32x4 hi = *reinterpret_cast<const u32x4 *>(sO + lds_off(row, fc + 1));
      const i32x8 af = i32x8{(int)lo[0], (int)lo[1], (int)lo[2], (int)lo[3],
                             (int)hi[0], (int)hi[1], (int)hi[2], (int)hi[3]};
#pragma unroll
      for (int b = 0; b < C::FI; ++b)
        acc[a][b] = __builtin_amdgcn_mfma_scale_f32_16x16x128_f8f6f4(af, bf[b], acc[a][b], 0, 0,
                                                                     0, 127, 0, sc);
    }
    if (ks + 1 < nk) store_lds(s ^ 1);
    __syncthreads();
  }

  // ---- epilogue: lane holds rows r .. r+3 of one column i
  const bool vec = (g.ldy % 4) == 0 && ((reinterpret_cast<uintptr_t>(g.y) & 15) == 0);
#pragma unroll
  for (int a = 0; a < C::FR; ++a) {
    const i64 r = r0 + wr * C::TR + 16 * a + 4 * (lane >> 4);
#pragma unroll
    for (int b = 0; b < C::FI; ++b) {
      const i64 i = i0 + wi * C::TI + 16 * b + frow;
      if (i >= g.m || r >= g.Nl) continue;
      float *dst = g.y + r + i * g.ldy;
      if (vec && r + 3 < g.Nl) {
        *reinterpret_cast<f32x4 *>(dst) = acc[a][b];
      } else {
        for (int q = 0; q < 4 && r + q < g.Nl; ++q) dst[q] = acc[a][b][q];
      }
    }
  }
}

template <class C>
void launch_f8(const F8Args &g) {
  auto kern = f8_sketch_kernel<C>;
  const int lds = 2 * C::STAGE;
  static std::once_flag once;
  std::call_once(once, [&] {
    HIP_CHECK(hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  lds));
  });
  const i64 tiles = ((g.Nl + C::BR - 1) / C::BR) * ((g.m + C::BI - 1) / C::BI);
  ISVD_CHECK(tiles <= 0x7FFFFFFF, "too many tiles");
  hipLaunchKernelGGL(kern, dim3((unsigned)tiles), dim3(C::NT), lds, ctx().stream, g);
  LAUNCH_CHECK();
}

int grid_for(i64 n) {
  i64 b = (n + 255) / 256;
  return (int)(b < 1 ? 1 : (b > 8192 ? 8192 : b));
}

}  // namespace

void quantize_e4m3(char ordera, i64 m, i64 n, const float *a, i64 lda, uint8_t *a8, i64 lda8,
                   float *scale, i64 lds) {
  if (m <= 0 || n <= 0) return;
  ISVD_CHECK(ordera == 'C' || ordera == 'R', "ordera must be 'C' or 'R'");
  ISVD_CHECK(lda >= (ordera == 'C' ? m : n), "lda (%lld) too small", lda);
  ISVD_CHECK(lda8 >= n && lda8 % 16 == 0, "lda8 (%lld) must be >= n and a multiple of 16", lda8);
  ISVD_CHECK((reinterpret_cast<uintptr_t>(a8) & 15) == 0, "a8 must be 16-byte aligned");
  const i64 nbi = (m + 127) / 128, nbj = (n + 127) / 128;
  ISVD_CHECK(lds >= nbj, "ldscale (%lld) must be >= ceil(n / 128) (%lld)", lds, nbj);
  ISVD_CHECK(nbi <= 65535 && nbj <= 0x7FFFFFFF, "matrix too large for the quantizer grid");
  hipLaunchKernelGGL(quantize_kernel, dim3((unsigned)nbj, (unsigned)nbi), dim3(256), 0,
                     ctx().stream, ordera, m, n, a, lda, a8, lda8, scale, lds);
  LAUNCH_CHECK();
}

void gaussian_omega_e(i64 n, i64 Nl, uint8_t *omega8, i64 ldo, unsigned long long seed) {
  ISVD_CHECK(ldo >= n, "ldomega (%lld) must be >= n", ldo);
  DBuf<float> tmp(Nl * n);
  gaussian_omega<float>(Nl, n, tmp.get(), Nl, seed);
  ISVD_CHECK((n + 63) / 64 <= 65535, "n too large for the conversion grid");
  hipLaunchKernelGGL(omega_e4m3_kernel, dim3((unsigned)((Nl + 63) / 64), (unsigned)((n + 63) / 64)),
                     dim3(256), 0, ctx().stream, n, Nl, (const float *)tmp.get(), omega8, ldo);
  LAUNCH_CHECK();
}

void stage_sketch_e(const Dims &d, const uint8_t *a8, i64 lda8, const float *scale, i64 lds,
                    const uint8_t *omega8, i64 ldo, float *yst, i64 ldyst) {
  const i64 m = d.m, n = d.n, Nl = d.Nl;
  ISVD_CHECK(lda8 >= n && lda8 % 16 == 0, "lda8 (%lld) must be >= n and a multiple of 16", lda8);
  ISVD_CHECK(ldo >= n && ldo % 16 == 0, "ldomega (%lld) must be >= n and a multiple of 16", ldo);
  ISVD_CHECK((reinterpret_cast<uintptr_t>(a8) & 15) == 0 &&
                 (reinterpret_cast<uintptr_t>(omega8) & 15) == 0,
             "a8 and omega8 must be 16-byte aligned");
  ISVD_CHECK(ldyst >= Nl, "ldyst (%lld) must be >= Nl", ldyst);
  const i64 nbi = (m + 127) / 128, nbj = (n + 127) / 128;
  ISVD_CHECK(lds >= nbj, "ldscale (%lld) must be >= ceil(n / 128) (%lld)", lds, nbj);
  DBuf<int> e8(nbi * nbj), bad(1);
  HIP_CHECK(hipMemsetAsync(bad.get(), 0, sizeof(int), ctx().stream));
  hipLaunchKernelGGL(scale_e8m0_kernel, dim3(grid_for(nbi * nbj)), dim3(256), 0, ctx().stream,
                     nbi, nbj, scale, lds, e8.get(), bad.get());
  LAUNCH_CHECK();
  int bad_h = 0;
  read_back(&bad_h, bad.get(), sizeof(int));
  ISVD_CHECK(bad_h == 0, "every block scale must be a power of two in [2^-127, 2^127]");
  F8Args g{Nl, m, n, omega8, ldo, a8, lda8, e8.get(), nbj, yst, ldyst};
  launch_f8<F8Main>(g);
}

}  // namespace isvd
