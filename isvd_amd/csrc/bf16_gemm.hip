// BF16 x BF16 -> FP32 GEMM on the gfx950 matrix cores (v_mfma_f32_32x32x16_bf16): every BF16
// product of the mixed-precision (b) and FP8 (e) families -- the mixed sketch Y^T = Omega^T A^T,
// the split Gram [H^T; L^T] H and the split projection [h0; h1; h2] A.
//
// C (M x N, FP32, column-major) := alpha op(A) op(B) + beta C, op(A) M x K, op(B) K x N, both
// BF16 column-major ('N' / 'T').  Block tile 256 x 256 x 32, four waves in a 2 x 2 grid, each
// 128 x 128 as 4 x 4 fragments of the 32 x 32 x 16 instruction (operand lane l: row l % 32,
// K values 8 (l / 32) .. +8 of the 16; accumulator register q of lane l: row
// 8 (q / 4) + 4 (l / 32) + q % 4, column l % 32).  Both operands move from HBM straight into
// LDS with global_load_lds_dwordx4 (one wave instruction fills 1 KB in lane order), three K
// steps in flight; each wave waits on its own fills with an explicit vector-memory count and a
// bare s_barrier.  With one wave per SIMD the matrix core idles whenever its wave stalls, so
// each K step is scheduled to keep it fed: the 16 products of the step's second K half are
// deferred past the next barrier (their operands already in registers) and cover the LDS
// latency of the next step's first fragments; the step's 8 fills are issued one per two of
// those products instead of back to back after the barrier (where the four waves' fills queue
// for the vector-memory pipe while their matrix cores wait); and the second half's fragments
// are read after the fourth product of the first half.  (Measured at 131072 x 16384 x 4096:
// 16.4 -> 15.1 ms with A row-major, 15.8 -> 14.6 ms column-major; a fourth stage or eight
// waves of 128 x 64 changed nothing.)  Each operand tile takes the LDS layout of its
// contiguous global dimension:
//   K-contiguous (op(A) = A^T, op(B) = B): one 64-byte row (32 K values) per M (N) index, the
//     four 16-byte chunks of row r at position c ^ ((r / 4) mod 4); a fragment is one 16-byte
//     LDS read per lane (chunk 2 h + l / 32 of row l % 32 for the K half h of the step).
//   MN-contiguous (op(A) = A, op(B) = B^T): one 512-byte row (256 M values) per K index, the
//     32 chunks of row k at position c ^ 4 (k mod 4); a fragment is two ds_read_b64_tr_b16
//     (transposing) reads per lane: in each group of 16 lanes, lane w addresses K row
//     k0 + w / 4, M values m0 + 4 (w mod 4) .. +4, and lane i receives M value m0 + i at K
//     rows k0 .. k0 + 3 (layout checked on gfx950).  The swizzle puts the four K rows of a
//     group in distinct banks.
// Rows (K-contiguous) or 16-byte chunks (MN-contiguous) past the edge of the tile are fetched
// from the last valid ones: they only feed outputs that are not stored.  A partial last K step
// goes through registers with bounds checks (zeros past K).  Requirements, met by the host
// (padded copies otherwise): 16-byte aligned operands, leading dimensions that are multiples
// of 8, and an MN extent that is a multiple of 8 for MN-contiguous operands.
#include "mixed.hpp"

#include <utility>

#include "lds_dma.hpp"

namespace isvd {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef short s16x4 __attribute__((ext_vector_type(4)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(3))) s16x4 lds_s16x4;

struct BgArgs {
  i64 M, N, K;      // C is M x N, K the reduction length
  i64 ma, nb;       // rows of op(A) / columns of op(B) that may be fetched (>= M / >= N)
  const uint16_t *A;
  i64 lda;
  const uint16_t *B;
  i64 ldb;
  float *C;
  i64 ldc;
  float alpha, beta;
  int ksplit;       // K chunks (grid.y)
  i64 kchunk;       // K values per chunk (multiple of BK)
  float *W;         // split-K partials: [ksplit][N][M]
  i64 upper_rows;   // tiles inside rows 0 .. upper_rows-1 strictly below the diagonal are skipped
};

namespace bg {
constexpr int BM = 256, BN = 256, BK = 32, NT = 256, S = 3;
constexpr int REGION = 256 * BK * 2;       // bytes of one operand tile per stage (16 KB)
constexpr int STAGE = 2 * REGION;
constexpr int FILLS = STAGE / 1024 / 4;    // fill instructions per wave per stage
__device__ inline int kc_off(int r, int c) { return r * 64 + ((c ^ ((r >> 2) & 3)) << 4); }
__device__ inline int mc_off(int k, int c) { return k * 512 + ((c ^ ((k & 3) << 2)) << 4); }
}  // namespace bg

template <bool KA, bool KB>
__global__ __launch_bounds__(bg::NT) void bf16_gemm_kernel(BgArgs g) {
  using namespace bg;
  __shared__ __align__(16) unsigned char st0[STAGE], st1[STAGE], st2[STAGE];
  auto stage = [&](auto k) -> unsigned char * {
    constexpr int kk = decltype(k)::value;
    if constexpr (kk == 0) return st0;
    else if constexpr (kk == 1) return st1;
    else return st2;
  };
  const i64 tiles_m = (g.M + BM - 1) / BM, tiles_n = (g.N + BN - 1) / BN;
  const i64 total = tiles_m * tiles_n;
  i64 pid = blockIdx.x;
  if (pid >= total) return;
  {
    // workgroups are dealt round-robin to the 8 XCDs: give each XCD a contiguous range
    constexpr i64 kXcd = 8;
    const i64 xcd = pid % kXcd, local = pid / kXcd;
    const i64 per = total / kXcd, rem = total % kXcd;
    pid = (xcd < rem) ? xcd * (per + 1) + local : rem * (per + 1) + (xcd - rem) * per + local;
  }
  // grouped raster: groups of kGroup column tiles (op(B) side), the column index fastest
  constexpr i64 kGroup = 8;
  const i64 grp = pid / (kGroup * tiles_m), first_n = grp * kGroup;
  const i64 gsz = (tiles_n - first_n) < kGroup ? (tiles_n - first_n) : kGroup;
  const i64 within = pid % (kGroup * tiles_m);
  const i64 m0 = (within / gsz) * BM, n0 = (first_n + within % gsz) * BN;
  // the caller mirrors C's symmetric top block (BM == BN: a tile below the diagonal holds only
  // elements below it)
  if (m0 > n0 && m0 + BM <= g.upper_rows) return;
  const int ks = blockIdx.y;
  const i64 kbeg = (i64)ks * g.kchunk;
  const i64 kend = (kbeg + g.kchunk < g.K) ? kbeg + g.kchunk : g.K;

  const int t = threadIdx.x, lane = t & 63, wave = __builtin_amdgcn_readfirstlane(t >> 6);
  const int wr = wave & 1, wc = wave >> 1;

  // operand block bases at this chunk's first K step (wave-uniform)
  const uint16_t *abase = g.A + (KA ? m0 * g.lda + kbeg : m0 + kbeg * g.lda);
  const uint16_t *bbase = g.B + (KB ? n0 * g.ldb + kbeg : n0 + kbeg * g.ldb);
  // fill slot u of this wave is fill instruction f = 4 u + wave of a stage: op(A) for f < 16
  unsigned off[FILLS], loff[FILLS];
#pragma unroll
  for (int u = 0; u < FILLS; ++u) {
    const int f = 4 * u + wave;
    const bool is_a = f < 16;
    const int j = f & 15;
    const bool kc = is_a ? KA : KB;
    const i64 ld = is_a ? g.lda : g.ldb;
    const i64 valid = is_a ? g.ma - m0 : g.nb - n0;  // >= 1 (the tile starts inside)
    if (kc) {
      const int row = 16 * j + (lane >> 2), c = (lane & 3) ^ ((row >> 2) & 3);
      const i64 rs = row < valid ? row : valid - 1;
      off[u] = (unsigned)(rs * ld + 8 * c);
    } else {
      const int k = 2 * j + (lane >> 5), c = (lane & 31) ^ ((k & 3) << 2);
      const i64 ms = (8 * c + 8 <= valid) ? 8 * c : valid - 8;
      off[u] = (unsigned)(k * ld + ms);
    }
    loff[u] = (is_a ? 0 : REGION) + j * 1024;
  }
  const i64 dA = KA ? BK : BK * g.lda, dB = KB ? BK : BK * g.ldb;  // per K step
  auto fill = [&](i64 kt, unsigned char *st) {
    const unsigned base = lds_addr(st);
    const uint16_t *ak = abase + kt * dA, *bk = bbase + kt * dB;
#pragma unroll
    for (int u = 0; u < FILLS; ++u)
      lds_fill16((4 * u + wave < 16 ? ak : bk) + off[u],
                 __builtin_amdgcn_readfirstlane(base + loff[u]));
  };
  // the partial last K step: registers with bounds checks, zeros past kend
  auto fill_checked = [&](i64 k0, unsigned char *st) {
    for (int e = t; e < 2 * 256 * BK; e += NT) {
      const bool is_a = e < 256 * BK;
      const int q = e & (256 * BK - 1);
      const bool kc = is_a ? KA : KB;
      const int mn = kc ? q >> 5 : q & 255, k = kc ? q & 31 : q >> 8;
      const i64 gmn = (is_a ? m0 : n0) + mn, gk = k0 + k;
      const i64 lim = is_a ? g.ma : g.nb;
      const uint16_t *src = is_a ? g.A : g.B;
      const i64 ld = is_a ? g.lda : g.ldb;
      uint16_t v = 0;
      if (gmn < lim && gk < kend) v = kc ? src[gmn * ld + gk] : src[gmn + gk * ld];
      const int boff = kc ? kc_off(mn, k >> 3) + 2 * (k & 7) : mc_off(k, mn >> 3) + 2 * (mn & 7);
      *reinterpret_cast<uint16_t *>(st + (is_a ? 0 : REGION) + boff) = v;
    }
  };

  f32x16 acc[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) acc[a][b] = f32x16{};

  // fragment of rows r0 .. r0 + 31 of one operand region for the K half h of a stage
  auto frag = [&](const unsigned char *reg, bool kc, int r0, int h) -> bf16x8 {
    if (kc) {
      return *reinterpret_cast<const bf16x8 *>(reg + kc_off(r0 + (lane & 31), 2 * h + (lane >> 5)));
    }
    const int w = lane & 15;
    const int mn = r0 + 16 * ((lane >> 4) & 1) + 4 * (w & 3);
    const int k = 16 * h + 8 * (lane >> 5) + (w >> 2);
    const int b0 = mc_off(k, mn >> 3) + 8 * ((mn >> 2) & 1);
    const int b1 = mc_off(k + 4, mn >> 3) + 8 * ((mn >> 2) & 1);
    const s16x4 lo = __builtin_amdgcn_ds_read_tr16_b64_v4i16((lds_s16x4 *)(reg + b0));
    const s16x4 hi = __builtin_amdgcn_ds_read_tr16_b64_v4i16((lds_s16x4 *)(reg + b1));
    const short v[8] = {lo[0], lo[1], lo[2], lo[3], hi[0], hi[1], hi[2], hi[3]};
    bf16x8 r;
    __builtin_memcpy(&r, v, 16);
    return r;
  };
  typedef bf16x8 Frag[4];
  auto load = [&](const unsigned char *st, int h, Frag &fa, Frag &fb) {
#pragma unroll
    for (int a = 0; a < 4; ++a) fa[a] = frag(st, KA, 128 * wr + 32 * a, h);
#pragma unroll
    for (int b = 0; b < 4; ++b) fb[b] = frag(st + REGION, KB, 128 * wc + 32 * b, h);
  };
  auto mma = [&](const Frag &fa, const Frag &fb) {
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int b = 0; b < 4; ++b)
        acc[a][b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fa[a], fb[b], acc[a][b], 0, 0, 0);
  };

  const i64 nkf = (kend - kbeg) / BK;  // full K steps
  const bool tail = ((kend - kbeg) % BK) != 0;
  if (nkf > 0) fill(0, st0);
  if (nkf > 1) fill(1, st1);
  Frag fa0, fb0, fa1, fb1;  // fa1 / fb1: the deferred second half of the previous step
#pragma unroll
  for (int a = 0; a < 4; ++a) fa1[a] = fb1[a] = bf16x8{};  // nothing deferred before step 0
  // step kt uses stage kt % S and refills stage (kt - 1) % S with step kt + S - 1
  auto step = [&](i64 kt, auto k) {
    constexpr int kk = decltype(k)::value;
    // this wave's fills of step kt have landed and its fragment reads of stage (kt - 1) % S
    // are complete (another wave refills that stage after the barrier)
    if (kt + S - 2 < nkf) __builtin_amdgcn_s_waitcnt(waitcnt_vm_lgkm0(FILLS * (S - 2)));
    else __builtin_amdgcn_s_waitcnt(waitcnt_vm_lgkm0(0));
    // the deferred operands are complete here: say so, or the compiler waits for them (and for
    // every LDS read issued after the barrier) in front of the deferred products
    asm volatile("" ::"v"(fa1[0]), "v"(fa1[1]), "v"(fa1[2]), "v"(fa1[3]), "v"(fb1[0]),
                 "v"(fb1[1]), "v"(fb1[2]), "v"(fb1[3]));
    bare_barrier();
    load(stage(k), 0, fa0, fb0);
    __builtin_amdgcn_sched_barrier(0);
    const bool refill = kt + S - 1 < nkf;
    const uint16_t *ak = abase + (kt + S - 1) * dA, *bk = bbase + (kt + S - 1) * dB;
    const unsigned fbase = lds_addr(stage(std::integral_constant<int, (kk + S - 1) % S>{}));
    // the deferred products, with the fills of step kt + S - 1 between them
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      acc[q >> 2][q & 3] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fa1[q >> 2], fb1[q & 3],
                                                                   acc[q >> 2][q & 3], 0, 0, 0);
      if ((q & 1) && refill) {
        const int u = q >> 1;
        __builtin_amdgcn_sched_barrier(0);
        lds_fill16((4 * u + wave < 16 ? ak : bk) + off[u],
                   __builtin_amdgcn_readfirstlane(fbase + loff[u]));
        __builtin_amdgcn_sched_barrier(0);
      }
    }
    // the first half of step kt; the second half's fragments are read after its fourth product
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      acc[q >> 2][q & 3] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fa0[q >> 2], fb0[q & 3],
                                                                   acc[q >> 2][q & 3], 0, 0, 0);
      if (q == 3) {
        __builtin_amdgcn_sched_barrier(0);
        load(stage(k), 1, fa1, fb1);
        __builtin_amdgcn_sched_barrier(0);
      }
    }
  };
  for (i64 kb = 0; kb < nkf; kb += S) {
    step(kb, std::integral_constant<int, 0>{});
    if (kb + 1 < nkf) step(kb + 1, std::integral_constant<int, 1>{});
    if (kb + 2 < nkf) step(kb + 2, std::integral_constant<int, 2>{});
  }
  mma(fa1, fb1);  // the deferred half of the last full step (zeros when there is none)
  if (tail) {
    __syncthreads();  // every wave is done with the stages
    fill_checked(kbeg + nkf * BK, st0);
    __syncthreads();
    load(st0, 0, fa0, fb0);
    mma(fa0, fb0);
    load(st0, 1, fa0, fb0);
    mma(fa0, fb0);
  }

  // epilogue: accumulator registers 4 q0 .. 4 q0 + 3 are four consecutive rows of one column
  const bool split = g.ksplit > 1;
  float *Cb = split ? g.W + (i64)ks * g.M * g.N : g.C;
  const i64 ldc = split ? g.M : g.ldc;
  const bool vec = (ldc % 4) == 0 && ((reinterpret_cast<uintptr_t>(Cb) & 15) == 0);
#pragma unroll
  for (int a = 0; a < 4; ++a) {
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      const i64 col = n0 + 128 * wc + 32 * b + (lane & 31);
      if (col >= g.N) continue;
#pragma unroll
      for (int q0 = 0; q0 < 4; ++q0) {
        const i64 r = m0 + 128 * wr + 32 * a + 8 * q0 + 4 * (lane >> 5);
        if (r >= g.M) continue;
        float *dst = Cb + r + col * ldc;
        f32x4 v = {acc[a][b][4 * q0], acc[a][b][4 * q0 + 1], acc[a][b][4 * q0 + 2],
                   acc[a][b][4 * q0 + 3]};
        if (!split) {
          v = v * g.alpha;
          if (g.beta != 0.f) {
            for (int q = 0; q < 4; ++q)
              if (r + q < g.M) v[q] += g.beta * dst[q];
          }
        }
        if (vec && r + 3 < g.M) {
          *reinterpret_cast<f32x4 *>(dst) = v;
        } else {
          for (int q = 0; q < 4 && r + q < g.M; ++q) dst[q] = v[q];
        }
      }
    }
  }
}

// C := alpha * sum over chunks of W + beta * C (fixed order: deterministic)
__global__ void bf16_splitk_reduce_kernel(i64 M, i64 N, int ksplit, const float *__restrict__ W,
                                          float alpha, float beta, float *__restrict__ C,
                                          i64 ldc) {
  const i64 mn = M * N;
  for (i64 e = blockIdx.x * (i64)blockDim.x + threadIdx.x; e < mn;
       e += (i64)gridDim.x * blockDim.x) {
    float s = 0.f;
    for (int q = 0; q < ksplit; ++q) s += W[q * mn + e];
    const i64 r = e % M, c = e / M;
    float v = alpha * s;
    if (beta != 0.f) v += beta * C[r + c * ldc];
    C[r + c * ldc] = v;
  }
}

// dst (rows_pad x cols, ld rows_pad) := src (rows x cols, lds), zero rows rows .. rows_pad - 1
__global__ void bf16_pad_kernel(i64 rows, i64 rows_pad, i64 cols, const uint16_t *__restrict__ src,
                                i64 lds, uint16_t *__restrict__ dst) {
  for (i64 e = blockIdx.x * (i64)blockDim.x + threadIdx.x; e < rows_pad * cols;
       e += (i64)gridDim.x * blockDim.x) {
    const i64 i = e % rows_pad, j = e / rows_pad;
    dst[e] = i < rows ? src[i + j * lds] : (uint16_t)0;
  }
}

bool aligned16(const void *p) { return (reinterpret_cast<uintptr_t>(p) & 15) == 0; }

// An operand as the kernel needs it: column-major `rows x cols` (rows the contiguous extent),
// 16-byte aligned, ld % 8 == 0, and rows % 8 == 0 when `mn_contig`; a padded copy otherwise.
struct Operand {
  const uint16_t *p;
  i64 ld;
  i64 rows_fetch;  // rows that may be fetched (the padded extent of an MN-contiguous operand)
  DBuf<uint16_t> copy;
  Operand(const uint16_t *src, i64 ld_, i64 rows, i64 cols, bool mn_contig)
      : p(src), ld(ld_), rows_fetch(rows) {
    const bool ok = aligned16(src) && ld_ % 8 == 0 && (!mn_contig || rows % 8 == 0);
    if (ok) return;
    const i64 rp = (rows + 7) / 8 * 8;
    copy.alloc(rp * cols);
    i64 blocks = (rp * cols + 255) / 256;
    if (blocks > 8192) blocks = 8192;
    hipLaunchKernelGGL(bf16_pad_kernel, dim3((unsigned)blocks), dim3(256), 0, ctx().stream, rows,
                       rp, cols, src, ld_, copy.get());
    LAUNCH_CHECK();
    p = copy.get();
    ld = rp;
    rows_fetch = rp;
  }
};

}  // namespace

namespace {

// Shared set-up of gemm_bf16 and gemm_bf16_partials: checks, operand copies where the kernel's
// alignment rules need them, the split-K choice (at most max_split chunks) and the launch.
// With W != nullptr the kernel writes raw partial sums there ([ksplit][n][m], also when ksplit
// is 1) and nothing is reduced; otherwise C := alpha op(A) op(B) + beta C.  Returns ksplit.
int run_bf16(char ta, char tb, i64 m, i64 n, i64 k, float alpha, const uint16_t *A, i64 lda,
             const uint16_t *B, i64 ldb, float beta, float *C, i64 ldc, float *W, int max_split,
             i64 upper_rows = 0) {
  ISVD_CHECK(k > 0, "gemm_bf16: k must be positive");
  const bool KA = (ta == 'T' || ta == 't');  // op(A) = A^T: K contiguous
  const bool KB = !(tb == 'T' || tb == 't');  // op(B) = B: K contiguous
  ISVD_CHECK(lda >= (KA ? k : m) && ldb >= (KB ? k : n) && (W || ldc >= m),
             "gemm_bf16: leading dimension too small (lda %lld ldb %lld ldc %lld)", lda, ldb, ldc);
  // stored extents: op(A) = A (m x k) or A^T (A is k x m); op(B) = B (k x n) or B^T (B n x k)
  Operand oa(A, lda, KA ? k : m, KA ? m : k, !KA);
  Operand ob(B, ldb, KB ? k : n, KB ? n : k, !KB);
  BgArgs g{};
  g.M = m; g.N = n; g.K = k;
  g.ma = KA ? m : oa.rows_fetch;
  g.nb = KB ? n : ob.rows_fetch;
  g.A = oa.p; g.lda = oa.ld; g.B = ob.p; g.ldb = ob.ld;
  g.C = C; g.ldc = ldc; g.alpha = alpha; g.beta = beta;
  g.upper_rows = upper_rows;
  // per-lane fill offsets are 32-bit
  ISVD_CHECK(256 * g.lda < (1ll << 32) && 256 * g.ldb < (1ll << 32),
             "gemm_bf16: leading dimensions too large (%lld, %lld)", g.lda, g.ldb);
  const i64 tiles = ((m + bg::BM - 1) / bg::BM) * ((n + bg::BN - 1) / bg::BN);
  ISVD_CHECK(tiles <= 0x7FFFFFFF, "gemm_bf16: too many tiles");
  // one workgroup per CU (96 KB of LDS, 256 accumulators per lane): split K when the tiles
  // fill less than one round of the 256 CUs, keeping at least 16 K steps per chunk and the
  // partial sums within the split-K budget
  int ksplit = 1;
  constexpr i64 kCUs = 256;
  if (tiles < kCUs && k >= 32 * bg::BK) {
    i64 want = kCUs / tiles;
    const i64 maxs = k / (16 * bg::BK);
    if (want > maxs) want = maxs;
    if (want > max_split) want = max_split;
    const size_t per = sizeof(float) * (size_t)(m * n);
    while (!W && want > 1 && per * (size_t)want > splitk_budget()) --want;
    if (ctx().max_wgs > 0)
      while (want > 1 && tiles * want > ctx().max_wgs) --want;
    ksplit = (int)(want < 1 ? 1 : want);
  }
  i64 kchunk = (k + ksplit - 1) / ksplit;
  kchunk = (kchunk + bg::BK - 1) / bg::BK * bg::BK;
  ksplit = (int)((k + kchunk - 1) / kchunk);
  g.ksplit = ksplit;
  g.kchunk = kchunk;
  if (W) {
    // raw partials: chunk s at W + s m n (dense, ld m); a single chunk goes there unscaled
    if (ksplit > 1) {
      g.W = W;
    } else {
      g.C = W; g.ldc = m; g.alpha = 1.f; g.beta = 0.f;
    }
  } else if (ksplit > 1) {
    g.W = static_cast<float *>(splitk_workspace(sizeof(float) * (size_t)(ksplit * m * n)));
  }
  dim3 grid((unsigned)tiles, (unsigned)ksplit);
  if (KA && KB) hipLaunchKernelGGL((bf16_gemm_kernel<true, true>), grid, dim3(bg::NT), 0, ctx().stream, g);
  else if (KA) hipLaunchKernelGGL((bf16_gemm_kernel<true, false>), grid, dim3(bg::NT), 0, ctx().stream, g);
  else if (KB) hipLaunchKernelGGL((bf16_gemm_kernel<false, true>), grid, dim3(bg::NT), 0, ctx().stream, g);
  else hipLaunchKernelGGL((bf16_gemm_kernel<false, false>), grid, dim3(bg::NT), 0, ctx().stream, g);
  LAUNCH_CHECK();
  if (!W && ksplit > 1) {
    i64 blocks = (m * n + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(bf16_splitk_reduce_kernel, dim3((unsigned)blocks), dim3(256), 0,
                       ctx().stream, m, n, ksplit, (const float *)g.W, alpha, beta, C, ldc);
    LAUNCH_CHECK();
  }
  return ksplit;
}

}  // namespace

void gemm_bf16(char ta, char tb, i64 m, i64 n, i64 k, float alpha, const uint16_t *A, i64 lda,
               const uint16_t *B, i64 ldb, float beta, float *C, i64 ldc) {
  if (m <= 0 || n <= 0) return;
  (void)run_bf16(ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, nullptr, 16);
}

void gemm_bf16_upper(char ta, char tb, i64 m, i64 n, i64 k, i64 upper_rows, const uint16_t *A,
                     i64 lda, const uint16_t *B, i64 ldb, float *C, i64 ldc) {
  if (m <= 0 || n <= 0) return;
  (void)run_bf16(ta, tb, m, n, k, 1.f, A, lda, B, ldb, 0.f, C, ldc, nullptr, 16, upper_rows);
}

int gemm_bf16_partials(char ta, char tb, i64 m, i64 n, i64 k, const uint16_t *A, i64 lda,
                       const uint16_t *B, i64 ldb, float *W, int max_split) {
  ISVD_CHECK(m > 0 && n > 0 && W && max_split >= 1, "gemm_bf16_partials: bad arguments");
  return run_bf16(ta, tb, m, n, k, 1.f, A, lda, B, ldb, 0.f, nullptr, m, W, max_split);
}

}  // namespace isvd
