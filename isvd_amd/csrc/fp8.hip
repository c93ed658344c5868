// FP8 (OCP e4m3) sketch on the gfx950 block-scaled matrix cores.
//
// Product: Y^T[r, i] = sum_j Omega[j, r] A[i, j] over j < n, with A[i, j] =
// decode(a8[i, j]) * 2^e(i / 128, j / 128).  The product runs on
// v_mfma_scale_f32_32x32x64_f8f6f4, one K step of 64 bytes per instruction; a 128-wide scale
// block spans two K steps, so the block's exponent rides in the instruction's B-operand scale
// (E8M0) and the accumulation stays in the MFMA's FP32 accumulator with no vector-ALU
// rescaling.  Operand layouts of the instruction (checked on gfx950): lane l holds row l % 32
// of each operand, K bytes 32 (l / 32) .. +32 ascending; accumulator register q of lane l is
// D[8 (q / 4) + 4 (l / 32) + q % 4][l % 32].  Both operands here are K-contiguous rows (Omega
// columns, A rows), so a fragment is two 16-byte LDS reads per lane.
#include "fp8.hpp"

#include <mutex>
#include <utility>

#include "la.hpp"
#include "lds_dma.hpp"

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

typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef __attribute__((address_space(3))) void lds_void_t;

// 16-byte chunk swizzle of a 64-byte LDS row: chunk c of row r sits at position
// c ^ swz(r), with swz(r) = (r / 4) mod 4.  Chosen by measurement on gfx950: for the fragment
// read (lane l: row l % 32, chunk 2 (l / 32) + h, one ds_read_b128) this swizzle, like
// (r / 8) mod 4, counts no LDS bank conflicts and runs as fast as a contiguous read, while
// ((r / 2) ^ (r / 8)) mod 4, (r / 2) mod 4 and ((r / 4) ^ (r / 16)) mod 4 count 4 conflict
// cycles per instruction (1.6x the time) and r mod 4 or no swizzle 12 (3x).
__device__ inline int swz(int r) { return (r >> 2) & 3; }
__device__ inline int lds_off(int row, int c) { return row * 64 + ((c ^ swz(row)) << 4); }

// Block tile 256 (Omega rows) x 256 (A rows), four waves in a 2 x 2 grid, 128 x 128 per wave
// as 4 x 4 fragments of v_mfma_scale_f32_32x32x64_f8f6f4 (lane l: operand row l % 32, K
// bytes 32 (l / 32) .. +32; accumulator q: row 8 (q / 4) + 4 (l / 32) + q % 4, column l % 32).
// One K step is 64 bytes: 512 LDS rows of 64 bytes (32 KB) per stage, S stages in flight.
// Both operands move from HBM straight into LDS with global_load_lds_dwordx4: a wave
// instruction fills 1 KB (16 rows) in lane order, so lane l owns row 16 j + l / 4, position
// l % 4, and fetches the logical chunk (l % 4) ^ swz(row).  Rows past the matrix edge are
// fetched from the last valid row: they only feed outputs that are not stored.  Each wave
// issues 8 fills per K step; step ks has landed once 8 (S - 2) of the wave's later fills remain.
template <int S>
struct F8Cfg {
  static constexpr int BR = 256, BI = 256, NT = 256;
  static constexpr int ROWS = BR + BI;
  static constexpr int STAGE = ROWS * 64;
  static constexpr int FILLS = ROWS / 16 / 4;  // per wave per K step
  static constexpr int LDS = STAGE * S;
};

template <int S>
__global__ __launch_bounds__(256) void f8_sketch_kernel(F8Args g) {
  using C = F8Cfg<S>;
  // one LDS object per stage: the compiler then sees that a fill of one stage and the
  // fragment reads of another do not alias, and waits only on what a read needs
  __shared__ __align__(16) unsigned char st0[C::STAGE], st1[C::STAGE], st2[C::STAGE],
      st3[S > 3 ? C::STAGE : 16], st4[S > 4 ? C::STAGE : 16];
  auto stage = [&](auto k) -> unsigned char * {
    constexpr int kk = decltype(k)::value;
    if constexpr (kk == 0) return st0;
    else if constexpr (kk == 1) return st1;
    else if constexpr (kk == 2) return st2;
    else if constexpr (kk == 3) return st3;
    else return st4;
  };
  const i64 nR = (g.Nl + C::BR - 1) / C::BR, nI = (g.m + C::BI - 1) / C::BI;
  const i64 total = nR * nI;
  i64 pid = blockIdx.x;
  if (pid >= total) return;
  {
    // blocks are dealt round-robin to the 8 XCDs: give each XCD a contiguous range of tiles
    constexpr i64 kXcd = 8;
    const i64 xcd = pid % kXcd, local = pid / kXcd;
    const i64 per = total / kXcd, rem = total % kXcd;
    pid = (xcd < rem) ? xcd * (per + 1) + local : rem * (per + 1) + (xcd - rem) * per + local;
  }
  // grouped raster: tiles run in groups of kGroup A tiles, the A tile index fastest, so the
  // blocks resident on one XCD (32) share 8 A tiles and 4 Omega tiles in its L2 instead of
  // streaming a different Omega tile each
  constexpr i64 kGroup = 8;
  const i64 grp = pid / (kGroup * nR), first_i = grp * kGroup;
  const i64 gsz = (nI - first_i) < kGroup ? (nI - first_i) : kGroup;
  const i64 within = pid % (kGroup * nR);
  const i64 r0 = (within / gsz) * C::BR, i0 = (first_i + within % gsz) * C::BI;
  // the wave index is uniform: keep it (and the scale reads that depend on it) scalar
  const int t = threadIdx.x, lane = t & 63, wave = __builtin_amdgcn_readfirstlane(t >> 6);
  const int wr = wave & 1, wi = wave >> 1;

  // block-local operand bases: the offsets below stay under 2^32
  const uint8_t *om_blk = g.om + r0 * g.ldo;
  const uint8_t *a_blk = g.a8 + i0 * g.lda;
  const int rows_om = (int)((g.Nl - r0) < C::BR ? (g.Nl - r0) : C::BR);
  const int rows_a = (int)((g.m - i0) < C::BI ? (g.m - i0) : C::BI);
  unsigned off[C::FILLS];
#pragma unroll
  for (int u = 0; u < C::FILLS; ++u) {
    const int j = u * 4 + wave;              // fill index within the stage
    const int row = 16 * j + (lane >> 2);    // LDS row
    const int c = (lane & 3) ^ swz(row);     // logical chunk fetched by this lane
    if (row < C::BR) {
      const int rr = row < rows_om ? row : rows_om - 1;
      off[u] = (unsigned)(rr * g.ldo + 16 * c);
    } else {
      const int ii = (row - C::BR) < rows_a ? (row - C::BR) : rows_a - 1;
      off[u] = (unsigned)(ii * g.lda + 16 * c);
    }
  }
  auto fill = [&](i64 k0, unsigned char *base) {
#pragma unroll
    for (int u = 0; u < C::FILLS; ++u) {
      const int j = u * 4 + wave;
      const uint8_t *src = (16 * j < C::BR ? om_blk : a_blk) + k0 + off[u];
      lds_fill16(src, lds_addr(base + j * 1024));
    }
  };
  // the partial last K step: register path with bounds and byte masks, then LDS stores
  auto fill_checked = [&](i64 k0, unsigned char *base) {
    for (int q = t; q < C::ROWS * 4; q += C::NT) {
      const int row = q >> 2, c = q & 3;
      const i64 kk = k0 + 16 * c;
      const uint8_t *src;
      bool ok;
      if (row < C::BR) {
        ok = row < rows_om;
        src = om_blk + (i64)row * g.ldo + kk;
      } else {
        ok = (row - C::BR) < rows_a;
        src = a_blk + (i64)(row - C::BR) * g.lda + kk;
      }
      u32x4 v = {0u, 0u, 0u, 0u};
      if (ok && kk < g.n) {
        v = *reinterpret_cast<const u32x4 *>(src);
        const i64 rem = g.n - kk;
        if (rem < 16) {
          for (int w = 0; w < 4; ++w) {
            const i64 keep = rem - 4 * w;
            v[w] &= keep >= 4 ? 0xFFFFFFFFu : (keep <= 0 ? 0u : ((1u << (8 * keep)) - 1u));
          }
        }
      }
      *reinterpret_cast<u32x4 *>(base + lds_off(row, c)) = v;
    }
  };

  f32x16 acc[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) acc[a][b] = f32x16{};

  // scale row of this wave's 128 A rows (waves wholly past the edge store nothing; clamp)
  const i64 nbi = (g.m + 127) / 128;
  const i64 bi_raw = (i0 + 128 * wi) / 128;
  // read through the constant address space: a scalar load, outside the vector-memory
  // counter that the stage waits count
  const __attribute__((address_space(4))) int *e8row =
      (const __attribute__((address_space(4))) int *)(g.e8 + (bi_raw < nbi ? bi_raw : nbi - 1) * g.nbj);
  const int frow = lane & 31, fc = 2 * (lane >> 5);

  auto compute = [&](const unsigned char *st, i64 ks) {
    const int sc = e8row[ks >> 1];  // scale blocks are 128 = two K steps wide
    i32x8 bf[4];
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      const int row = C::BR + 128 * wi + 32 * b + frow;
      const u32x4 lo = *reinterpret_cast<const u32x4 *>(st + lds_off(row, fc));
      const u32x4 hi = *reinterpret_cast<const u32x4 *>(st + lds_off(row, fc + 1));
      bf[b] = i32x8{(int)lo[0], (int)lo[1], (int)lo[2], (int)lo[3],
                    (int)hi[0], (int)hi[1], (int)hi[2], (int)hi[3]};
    }
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      const int row = 128 * wr + 32 * a + frow;
      const u32x4 lo = *reinterpret_cast<const u32x4 *>(st + lds_off(row, fc));
      const u32x4 hi = *reinterpret_cast<const u32x4 *>(st + lds_off(row, fc + 1));
      const i32x8 af = i32x8{(int)lo[0], (int)lo[1], (int)lo[2], (int)lo[3],
                             (int)hi[0], (int)hi[1], (int)hi[2], (int)hi[3]};
#pragma unroll
      for (int b = 0; b < 4; ++b)
        acc[a][b] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(af, bf[b], acc[a][b], 0, 0,
                                                                    0, 127, 0, sc);
    }
  };

  // ---- full K steps, S - 1 in flight ahead of the one computed
  const i64 nkf = g.n / 64;
  const bool tail = (g.n % 64) != 0;
  // prologue: steps 0 .. S - 2 into stages 0 .. S - 2
  [&]<int... P>(std::integer_sequence<int, P...>) {
    ((P < nkf ? fill(64 * P, stage(std::integral_constant<int, P>{})) : void()), ...);
  }(std::make_integer_sequence<int, S - 1>{});
  // step ks uses stage ks % S and refills stage (ks - 1) % S with step ks + S - 1
  auto step = [&](i64 ks, auto k) {
    constexpr int kk = decltype(k)::value;
    if (ks + S - 2 < nkf) __builtin_amdgcn_s_waitcnt(waitcnt_vm(C::FILLS * (S - 2)));
    else __builtin_amdgcn_s_waitcnt(0);
    // every wave's fills of step ks are in; stage (ks - 1) % S is free.  A bare s_barrier:
    // the waits above are exact, a workgroup fence would wait for every fill in flight
    bare_barrier();
    if (ks + S - 1 < nkf)
      fill(64 * (ks + S - 1), stage(std::integral_constant<int, (kk + S - 1) % S>{}));
    compute(stage(k), ks);
  };
  for (i64 kb = 0; kb < nkf; kb += S) {
    [&]<int... P>(std::integer_sequence<int, P...>) {
      ((kb + P < nkf ? step(kb + P, std::integral_constant<int, P>{}) : void()), ...);
    }(std::make_integer_sequence<int, S>{});
  }
  if (tail) {
    __syncthreads();  // every wave is done with the stages
    fill_checked(64 * nkf, st0);
    __syncthreads();
    compute(st0, nkf);
  }

  // ---- epilogue: accumulator registers 4 q0 .. 4 q0 + 3 are 4 consecutive rows of one column
  const bool vec = (g.ldy % 4) == 0 && ((reinterpret_cast<uintptr_t>(g.y) & 15) == 0);
#pragma unroll
  for (int a = 0; a < 4; ++a) {
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      const i64 i = i0 + 128 * wi + 32 * b + frow;
      if (i >= g.m) continue;
#pragma unroll
      for (int q0 = 0; q0 < 4; ++q0) {
        const i64 r = r0 + 128 * wr + 32 * a + 8 * q0 + 4 * (lane >> 5);
        if (r >= g.Nl) continue;
        float *dst = g.y + r + i * g.ldy;
        const f32x4 v = {acc[a][b][4 * q0], acc[a][b][4 * q0 + 1], acc[a][b][4 * q0 + 2],
                         acc[a][b][4 * q0 + 3]};
        if (vec && r + 3 < g.Nl) {
          *reinterpret_cast<f32x4 *>(dst) = v;
        } else {
          for (int q = 0; q < 4 && r + q < g.Nl; ++q) dst[q] = v[q];
        }
      }
    }
  }
}

template <int S>
void launch_f8(const F8Args &g) {
  using C = F8Cfg<S>;
  const i64 tiles = ((g.Nl + C::BR - 1) / C::BR) * ((g.m + C::BI - 1) / C::BI);
  ISVD_CHECK(tiles <= 0x7FFFFFFF, "too many tiles");
  hipLaunchKernelGGL(f8_sketch_kernel<S>, dim3((unsigned)tiles), dim3(C::NT), 0, ctx().stream,
                     g);
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
  ISVD_CHECK(256 * lda8 < (1ll << 32) && 256 * ldo < (1ll << 32), "n too large (lda8 %lld)", lda8);
  F8Args g{Nl, m, n, omega8, ldo, a8, lda8, e8.get(), nbj, yst, ldyst};
  // three stages: at this tile the time per K step does not change from 3 to 5 stages
  launch_f8<3>(g);
}

}  // namespace isvd
