// MFMA GEMM for gfx950 (CDNA4), FP64 and FP32, column-major operands.
//
// One kernel template covers the four transpose combinations.  Tiles are staged through
// LDS in k-major layout ([BK][BM] for op(A), [BK][BN] for op(B)), so the MFMA operand
// fragments (lane -> (row = lane%16, k = lane/16)) are read with unit stride.  Global loads
// walk the contiguous dimension of each operand, so both the "N" and the "T" forms load
// coalesced.  A register-prefetch double buffer overlaps the next tile's global loads with
// the current tile's MFMAs (one barrier per K tile).
#include "la.hpp"

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
};

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

  // ---- tile coordinates ----
  // Workgroups are dispatched round-robin over the 8 XCDs (each with its own L2).  Remap so
  // that each XCD walks a contiguous range of the grouped raster: the tiles running together
  // on one XCD then share operand panels through that XCD's L2.
  const i64 tiles_m = (g.m + BM - 1) / BM;
  const i64 tiles_n = (g.n + BN - 1) / BN;
  // a symmetric product enumerates only the tiles touching the upper triangle
  auto upper_count = [&](i64 tn) {
    const i64 c = ((tn + 1) * BN + BM - 1) / BM;
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
  if (g.upper) {  // column by column over the upper tiles
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

  // ---- epilogue ----
  if (g.ksplit == 1) {
    T *Cb = g.C + (i64)batch_id * g.sC;
#pragma unroll
    for (int a = 0; a < FM; ++a)
#pragma unroll
      for (int b = 0; b < FN; ++b)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const i64 r = m0 + wm * WTM + a * 16 + Mfma<T>::row(lane, q);
          const i64 c = n0 + wn * WTN + b * 16 + lr;
          if (r < g.m && c < g.n) {
            T v = g.alpha * acc[a][b][q];
            if (g.beta != T(0)) v += g.beta * Cb[r + c * g.ldc];
            Cb[r + c * g.ldc] = v;
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
          const i64 r = m0 + wm * WTM + a * 16 + Mfma<T>::row(lane, q);
          const i64 c = n0 + wn * WTN + b * 16 + lr;
          if (r < g.m && c < g.n) Wb[r + c * g.m] = acc[a][b][q];
        }
  }
}

// C = alpha * sum_s W[s] + beta * C  (fixed summation order -> deterministic).
template <typename T>
__global__ void splitk_reduce_kernel(i64 m, i64 n, int ksplit, const T *__restrict__ W,
                                     T alpha, T beta, T *__restrict__ C, i64 ldc, i64 sC) {
  const int b = blockIdx.y;
  const i64 mn = m * n;
  const T *Wb = W + (i64)b * ksplit * mn;
  T *Cb = C + (i64)b * sC;
  for (i64 e = blockIdx.x * (i64)blockDim.x + threadIdx.x; e < mn;
       e += (i64)gridDim.x * blockDim.x) {
    T s = T(0);
    for (int q = 0; q < ksplit; ++q) s += Wb[q * mn + e];
    const i64 r = e % m, c = e / m;
    T v = alpha * s;
    if (beta != T(0)) v += beta * Cb[r + c * ldc];
    Cb[r + c * ldc] = v;
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
  hipLaunchKernelGGL((gemm_kernel<T, TA, TB, BM, BN, BK, WGM, WGN>), grid, dim3(64 * WGM * WGN),
                     0, ctx().stream, g);
  LAUNCH_CHECK();
  if (ksplit > 1) {
    const i64 mn = g.m * g.n;
    int blocks = (int)((mn + 255) / 256);
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(splitk_reduce_kernel<T>, dim3(blocks, (unsigned)batch), dim3(256), 0,
                       ctx().stream, g.m, g.n, ksplit, (const T *)g.W, g.alpha, g.beta, g.C,
                       g.ldc, g.sC);
    LAUNCH_CHECK();
  }
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
                      i64 sB, i64 sC, int upper) {
  if (m <= 0 || n <= 0 || batch <= 0) return;
  const bool ta = (transa == 'T' || transa == 't' || transa == 'C' || transa == 'c');
  const bool tb = (transb == 'T' || transb == 't' || transb == 'C' || transb == 'c');
  ISVD_CHECK(ldc >= m, "ldc (%lld) < m (%lld)", ldc, m);
  if (k <= 0 || alpha == T(0)) {
    int blocks = (int)((m * n + 255) / 256);
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(scale_kernel<T>, dim3(blocks, (unsigned)batch), dim3(256), 0,
                       ctx().stream, m, n, beta, C, ldc, sC);
    LAUNCH_CHECK();
    return;
  }
  ISVD_CHECK(lda >= (ta ? k : m), "lda (%lld) too small", lda);
  ISVD_CHECK(ldb >= (tb ? n : k), "ldb (%lld) too small", ldb);
  GemmArgs<T> g{};
  g.m = m; g.n = n; g.k = k; g.alpha = alpha; g.beta = beta;
  g.A = A; g.lda = lda; g.B = B; g.ldb = ldb; g.C = C; g.ldc = ldc;
  g.sA = sA; g.sB = sB; g.sC = sC; g.W = nullptr; g.upper = upper;
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

template void gemm<double>(char, char, i64, i64, i64, double, const double *, i64,
                           const double *, i64, double, double *, i64, i64, i64, i64, i64);
template void gemm<float>(char, char, i64, i64, i64, float, const float *, i64, const float *,
                          i64, float, float *, i64, i64, i64, i64, i64);

}  // namespace isvd
