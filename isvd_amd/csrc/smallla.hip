// l x l (l <= 128) building blocks of the integration loops:
//  - small_gemm: C := alpha op(A) op(B) + beta C + diag I for up to two independent problems
//    per launch, one wave per 16 x 16 output tile (v_mfma_f64_16x16x4f64 / _f32_16x16x4f32),
//    operands read straight from L2 (no LDS, no workspace), so a 128^3 product is one short
//    launch;
//  - kn_chain: the l x l part of a Kolmogorov-Nagumo iteration in one launch (C and C^-1 of
//    (I/2 + (I/4 - Z)^{1/2})^{1/2} by a Taylor series or coupled Newton-Schulz square roots);
//  - inv_small: Gauss-Jordan inverse with partial pivoting, matrix held in LDS.
#include "dpp.hpp"
#include "la.hpp"

#include <cmath>
#include <vector>

namespace isvd {

typedef double d4s __attribute__((ext_vector_type(4)));
typedef float f4s __attribute__((ext_vector_type(4)));

namespace {

template <typename T> struct SM;
template <> struct SM<double> {
  using acc = d4s;
  static __device__ __forceinline__ acc mfma(double a, double b, acc c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lane, int g) { return (lane >> 4) + 4 * g; }
};
template <> struct SM<float> {
  using acc = f4s;
  static __device__ __forceinline__ acc mfma(float a, float b, acc c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lane, int g) { return 4 * (lane >> 4) + g; }
};

constexpr int kSmallMaxK = 128;

template <typename T>
struct SmallProblem {
  const T *A;
  const T *B;
  T *C;
  int lda, ldb, ldc;
  bool ta, tb;
  T alpha, beta, diag;
};

// grid: (tiles_m * tiles_n, nprob), block: 64 (one wave)
template <typename T>
__global__ __launch_bounds__(64) void small_gemm_kernel(int m, int n, int k, SmallProblem<T> p0,
                                                        SmallProblem<T> p1) {
  const SmallProblem<T> &p = (blockIdx.y == 0) ? p0 : p1;
  const int tiles_m = (m + 15) / 16;
  const int tm = blockIdx.x % tiles_m, tn = blockIdx.x / tiles_m;
  const int lane = threadIdx.x, lr = lane & 15, lk = lane >> 4;
  const int row = tm * 16 + lr, col = tn * 16 + lr;
  typename SM<T>::acc acc = {0, 0, 0, 0};
  // all fragments are loaded before the MFMA chain so the L2 round trips overlap (k <= 128)
  T fa[kSmallMaxK / 4], fb[kSmallMaxK / 4];
#pragma unroll
  for (int s = 0; s < kSmallMaxK / 4; ++s) {
    const int kk = 4 * s + lk;
    T a = 0, b = 0;
    if (row < m && kk < k) a = p.ta ? p.A[kk + (i64)row * p.lda] : p.A[row + (i64)kk * p.lda];
    if (col < n && kk < k) b = p.tb ? p.B[col + (i64)kk * p.ldb] : p.B[kk + (i64)col * p.ldb];
    fa[s] = a;
    fb[s] = b;
  }
#pragma unroll
  for (int s = 0; s < kSmallMaxK / 4; ++s)
    if (4 * s < k) acc = SM<T>::mfma(fa[s], fb[s], acc);
#pragma unroll
  for (int g = 0; g < 4; ++g) {
    const int r = tm * 16 + SM<T>::row(lane, g), c = tn * 16 + lr;
    if (r < m && c < n) {
      T v = p.alpha * acc[g] + ((r == c) ? p.diag : T(0));
      if (p.beta != T(0)) v += p.beta * p.C[r + (i64)c * p.ldc];
      p.C[r + (i64)c * p.ldc] = v;
    }
  }
}

template <typename T>
void small_gemm2(int m, int n, int k, const SmallProblem<T> &p0, const SmallProblem<T> *p1) {
  ISVD_CHECK(k >= 0 && k <= kSmallMaxK, "small_gemm: k = %d > %d", k, kSmallMaxK);
  const int tiles = ((m + 15) / 16) * ((n + 15) / 16);
  hipLaunchKernelGGL(small_gemm_kernel<T>, dim3(tiles, p1 ? 2 : 1), dim3(64), 0, ctx().stream, m,
                     n, k, p0, p1 ? *p1 : p0);
  LAUNCH_CHECK();
}

// Gauss-Jordan in-place inverse with partial pivoting, n <= 128, matrix in LDS (row-major,
// padded rows).  Row interchanges as getrf; the column interchanges of the computed inverse
// are undone at the end as getri does.
constexpr int kGjThreads = 1024;

template <typename T>
__global__ __launch_bounds__(kGjThreads) void gj_inverse_lds_kernel(int n, T *A, i64 lda,
                                                                    int *info) {
  extern __shared__ unsigned char smem_raw[];
  const int ld = n + 1;
  T *S = reinterpret_cast<T *>(smem_raw);  // n x ld, row-major
  __shared__ T rowj[128], colj[128];
  __shared__ int ipiv[128], pos[128], colidx[128];
  __shared__ int piv_sh, bad_sh;
  const int tid = threadIdx.x;
  for (int e = tid; e < n * n; e += kGjThreads) {
    const int r = e % n, c = e / n;
    S[r * ld + c] = A[r + (i64)c * lda];
  }
  if (tid == 0) bad_sh = 0;
  __syncthreads();
  for (int j = 0; j < n; ++j) {
    if (tid < 64) {  // wave 0: pivot = first index of max |S[r][j]|, r >= j (idamax)
      T best = -1;
      int bi = n;
      for (int r = j + tid; r < n; r += 64) {
        const T v = fabs(S[r * ld + j]);
        if (v > best) { best = v; bi = r; }
      }
      // (largest, then smallest index) is associative and commutative: within each 16-lane
      // row over DPP (pairs, quads, half rows, rows), then the four rows' results by readlane
      auto take = [&](T ob, int oi) {
        if (ob > best || (ob == best && oi < bi)) { best = ob; bi = oi; }
      };
      take(dpp_mov<0xB1>(best), dpp_mov<0xB1>(bi));
      take(dpp_mov<0x4E>(best), dpp_mov<0x4E>(bi));
      take(dpp_mov<0x141>(best), dpp_mov<0x141>(bi));
      take(dpp_mov<0x140>(best), dpp_mov<0x140>(bi));
      {
        T rb[3];
        int ri[3];
        const long long bits = __double_as_longlong((double)best);
        const int lo = (int)(bits & 0xffffffffll), hi = (int)(bits >> 32);
#pragma unroll
        for (int q = 1; q < 4; ++q) {
          const unsigned l32 = (unsigned)__builtin_amdgcn_readlane(lo, 16 * q);
          const long long h32 = __builtin_amdgcn_readlane(hi, 16 * q);
          rb[q - 1] = (T)__longlong_as_double((h32 << 32) | l32);
          ri[q - 1] = __builtin_amdgcn_readlane(bi, 16 * q);
        }
        if (tid == 0)
          for (int q = 0; q < 3; ++q) take(rb[q], ri[q]);
      }
      if (tid == 0) {
        piv_sh = bi;
        ipiv[j] = bi;
        if (!(best > T(0))) bad_sh = j + 1;
      }
    }
    __syncthreads();
    if (bad_sh) break;
    const int p = piv_sh;
    if (p != j) {
      for (int c = tid; c < n; c += kGjThreads) {
        const T t = S[j * ld + c];
        S[j * ld + c] = S[p * ld + c];
        S[p * ld + c] = t;
      }
      __syncthreads();
    }
    for (int c = tid; c < n; c += kGjThreads) rowj[c] = S[j * ld + c];
    for (int r = tid; r < n; r += kGjThreads) colj[r] = S[r * ld + j];
    __syncthreads();
    const T ipv = T(1) / rowj[j];
    // a fixed column per thread, rows strided by the number of row groups (n <= 128)
    {
      constexpr int kGroups = kGjThreads / 128;
      const int c = tid & 127;
      if (c < n) {
        const T rc = rowj[c];
        for (int r = tid >> 7; r < n; r += kGroups) {
          T v;
          if (r == j)
            v = (c == j) ? ipv : rc * ipv;
          else
            v = (c == j) ? -colj[r] * ipv : S[r * ld + c] - colj[r] * ipv * rc;
          S[r * ld + c] = v;
        }
      }
    }
    __syncthreads();
  }
  if (bad_sh) {
    if (tid == 0) *info = bad_sh;
    return;
  }
  if (tid == 0) {
    int *idx = colidx;  // in LDS: a private array would live in scratch memory
    for (int q = 0; q < n; ++q) idx[q] = q;
    for (int q = n - 1; q >= 0; --q) {
      const int t = idx[q];
      idx[q] = idx[ipiv[q]];
      idx[ipiv[q]] = t;
    }
    for (int q = 0; q < n; ++q) pos[idx[q]] = q;
    *info = 0;
  }
  __syncthreads();
  for (int e = tid; e < n * n; e += kGjThreads) {
    const int r = e % n, c = e / n;
    A[r + (i64)pos[c] * lda] = S[r * ld + c];
  }
}

int grid_of(i64 n) {
  i64 b = (n + 255) / 256;
  return (int)(b < 1 ? 1 : (b > 4096 ? 4096 : b));
}

}  // namespace

template <typename T>
void small_gemm(char ta, char tb, i64 m, i64 n, i64 k, T alpha, const T *A, i64 lda, const T *B,
                i64 ldb, T beta, T *C, i64 ldc, T diag) {
  SmallProblem<T> p{A, B, C, (int)lda, (int)ldb, (int)ldc, ta == 'T', tb == 'T', alpha, beta, diag};
  small_gemm2<T>((int)m, (int)n, (int)k, p, nullptr);
}

namespace {

// ---------------------------------------------------------------------------------------------
// kn_chain: the whole l x l chain of a KN iteration in one launch.  One workgroup of 8 waves per
// two 16 x 16 output tiles (ceil(ceil(l/16)^2 / 2) workgroups).  A round computes one or two
// products, K split over the waves of each (tile, product) job, and ends at a grid barrier.
// Every result is a deterministic function of the inputs: tiles, K slices and the order of
// every sum are fixed by l and the partial count alone.
//
// The workgroups sit on different XCDs, each with its own L2, so the intermediates pass between
// them through agent-scope atomic loads and stores, which are coherent across the device by
// themselves, and the barrier is an agent-scope arrival counter: a workgroup's stores have
// completed (vmcnt 0) before it arrives, and its loads after the barrier are issued after it saw
// the count reach the target.  No L2 is written back or invalidated (what agent-scope fences
// cost: several microseconds per barrier).  The counter is monotonic: a barrier waits for the
// next multiple of the workgroup count (the host zeroes it; every launch completes whole
// barriers).
//
// C = f(Z) and C^-1 = g(Z), f(x) = (1/2 + (1/4 - x)^(1/2))^(1/2), g = 1/f, for the symmetric
// Z with spectrum in [0, rho/4]:
//   small rho: the Taylor series of f and g at 0 (they converge for |4x| < 1), truncated where
//     the tail bound sum_{k>d} |c_k| (rho/4)^k falls below a quarter of the unit roundoff and
//     evaluated by Paterson-Stockmeyer in Z^4 (blocks of 4 powers): about 6-8 rounds;
//   otherwise: two coupled Newton-Schulz square roots (Higham, Functions of Matrices, 2008,
//     eq. 6.35) with step counts from ns_counts, 2 rounds a step;
//   rho out of reach of either: the caller takes the eigendecomposition.
// ---------------------------------------------------------------------------------------------
constexpr int kChainWaves = 8;
constexpr int kChainThreads = 64 * kChainWaves;
constexpr unsigned kChainSpinLimit = 1u << 22;  // ~4 s of polling
constexpr int kSeriesBlocks = 8;                // Paterson-Stockmeyer blocks of 4 powers
constexpr int kSeriesTerms = 4 * kSeriesBlocks + 1;

enum { kBarCount = 0, kBarAbort = 1 };

template <typename T>
__device__ __forceinline__ T cld(const T *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T>
__device__ __forceinline__ void cst(T *p, T v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// extra work of a product's epilogue, per output entry v at (r, c) (bit flags)
enum ChainEpi {
  kEpiStart1 = 1,   // (Z2 = Z Z) also Y1 := I - 2 Z - 8 Z2, W1 := I + 2 Z (out2, out3)
  kEpiStart2 = 2,   // (P P)      also W1 := (5 I - P) / 4 (out2; P = Cin)
  kEpiBlock = 4,    // v += sum_j coef[j] Z^j (j < 4; Z^0 = I) before anything else below
  kEpiTop = 8,      // (Z^3)      out2 := sum_j cf[j] Z^j, out3 := sum_j cg[j] Z^j (blocks)
  kEpiNormSq = 16,  // part[tile] := sum of v^2
  kEpiErrSq = 32,   // part[tile] := sum of (v - delta)^2
  kEpiW3Low = 64,   // W3 lower halves := [ninv v | v | v]
  kEpiW3Up = 128,   // W3 upper halves := [v | v | v]
  kEpiTopLast = 256,  // (with kEpiTop, one block) the top blocks are C and C^-1: part[tile +
                      // kPartErr] := sum of (C - delta)^2, W3 lower halves := [ninv C^-1 | ..]
};
// part[] holds ||Z2||_F partials read after the Z2 round; the error partials of a round that
// runs while some workgroups may still read those go to part[kPartErr ..]
constexpr int kPartErr = 64;

template <typename T>
struct ChainProd {
  const T *A;
  const T *B;
  const T *Cin;  // cin_coef * Cin is added when not null (ld l)
  T *out;        // null: the product itself is not stored
  T *out2, *out3;
  bool ta;
  T alpha, diag, cin_coef;
  int epi;
  const T *coef;   // kEpiBlock: 4 coefficients; kEpiTop: cf block (cg = coef + kSeriesTerms)
};

template <typename T>
struct ChainArgs {
  int l, tiles_m, nparts, mode;
  T ninv;
  const T *parts;  // nparts raw partial sums of B^T [B | Bg] (l x 2l each, dense)
  T *D, *Dg;       // out: D | Dg (ld l)
  T *Z, *Z2, *Z3, *Z4, *Y[2], *W[2], *E, *part;
  T *W3, *scal;
  const T *coef;   // Taylor coefficients of f then g (kSeriesTerms each)
  unsigned long long *bar;
};

// Grid barrier.  False when the wait ran past kChainSpinLimit polls or another workgroup gave up
// (the abort flag is then set).
__device__ bool chain_barrier(unsigned long long *bar) {
  asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
  __syncthreads();
  __shared__ int ok_sh;
  if (threadIdx.x == 0) {
    int ok = 1;
    const unsigned long long nwg = gridDim.x;
    const unsigned long long old =
        __hip_atomic_fetch_add(bar + kBarCount, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const unsigned long long target = (old / nwg + 1) * nwg;
    unsigned spins = 0;
    while (__hip_atomic_load(bar + kBarCount, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) <
           target) {
      if (++spins > kChainSpinLimit) {
        __hip_atomic_store(bar + kBarAbort, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        ok = 0;
        break;
      }
      __builtin_amdgcn_s_sleep(1);
    }
    if (__hip_atomic_load(bar + kBarAbort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) ok = 0;
    ok_sh = ok;
  }
  __syncthreads();
  return ok_sh != 0;
}

// Sum of n <= 64 values in a fixed order (lane i loads entry i; butterfly over the wave),
// returned to every thread of the workgroup.
template <typename T>
__device__ T chain_sum(const T *v, int n, T *sh) {
  if (threadIdx.x < 64) {
    T x = threadIdx.x < n ? cld(v + threadIdx.x) : T(0);
    for (int off = 1; off < 64; off <<= 1) x += __shfl_xor(x, off, 64);
    if (threadIdx.x == 0) *sh = x;
  }
  __syncthreads();
  const T r = *sh;
  __syncthreads();
  return r;
}

// sum_j c[j] Z^j (j < 4) at entry e, Z^0 = I
template <typename T>
__device__ __forceinline__ T block_at(const T *c, T dl, T z, T z2, T z3) {
  return c[0] * dl + c[1] * z + c[2] * z2 + c[3] * z3;
}

// One round over this workgroup's two tiles (2 wg, 2 wg + 1): product p0 and, with Dual, p1.
// Jobs (tile, product) are spread over the 8 waves: 2 jobs of 4 waves or 4 jobs of 2 waves, each
// splitting K over its waves (partials summed in LDS in a fixed order).  The operand loads of a
// wave are unconditional (indices clamped into the matrix, out-of-range values zeroed after) so
// that they are all in flight together.
template <typename T, bool Dual>
__device__ void chain_round(const ChainArgs<T> &a, const ChainProd<T> p0, const ChainProd<T> p1,
                            typename SM<T>::acc (*red)[64]) {
  constexpr int jobs = Dual ? 4 : 2, wpj = kChainWaves / jobs;
  constexpr int NS = (kSmallMaxK / 4 + wpj - 1) / wpj;  // K steps per wave at most
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
  const int job = wave / wpj, q = wave % wpj;
  const ChainProd<T> p = (Dual && (job >> 1)) ? p1 : p0;
  const int l = a.l, ntile = a.tiles_m * a.tiles_m;
  const int tile = 2 * blockIdx.x + (job & 1);
  const bool live = tile < ntile;
  const int tm = tile % a.tiles_m, tn = tile / a.tiles_m;
  const int lr = lane & 15, lk = lane >> 4;
  const int row = tm * 16 + lr, col = tn * 16 + lr;
  typename SM<T>::acc acc = {0, 0, 0, 0};
  // the epilogue's other operands (C_in, Z, Z^2, Z^3 at this lane's four entries), requested
  // with the product's operands: they were written in earlier rounds
  const int epi = p.epi;
  const bool need_z = epi & (kEpiStart1 | kEpiBlock | kEpiTop);
  const bool need_z2 = epi & (kEpiBlock | kEpiTop);
  const bool need_z3 = epi & kEpiBlock;
  T pcin[4] = {0, 0, 0, 0}, pz[4] = {0, 0, 0, 0}, pz2[4] = {0, 0, 0, 0}, pz3[4] = {0, 0, 0, 0};
  if (live) {
    if (q == 0) {
#pragma unroll
      for (int g = 0; g < 4; ++g) {
        int r = tm * 16 + SM<T>::row(lane, g), c = tn * 16 + lr;
        r = r < l ? r : l - 1;
        c = c < l ? c : l - 1;
        const i64 e = r + (i64)c * l;
        if (p.Cin) pcin[g] = cld(p.Cin + e);
        if (need_z) pz[g] = cld(a.Z + e);
        if (need_z2) pz2[g] = cld(a.Z2 + e);
        if (need_z3) pz3[g] = cld(a.Z3 + e);
      }
    }
    const int ksteps = (l + 3) / 4, spw = (ksteps + wpj - 1) / wpj;
    const int s0 = q * spw;
    const int rr = row < l ? row : l - 1, cc = col < l ? col : l - 1;
    const i64 sk = p.ta ? 1 : l, sr = p.ta ? l : 1;  // strides of op(A) along K and rows
    const T *pa = p.A + rr * sr;
    const T *pb = p.B + (i64)cc * l;
    T fa[NS], fb[NS];
#pragma unroll
    for (int s = 0; s < NS; ++s) {
      const int kk = 4 * (s0 + s) + lk;
      const int kc = kk < l ? kk : l - 1;
      fa[s] = cld(pa + kc * sk);
      fb[s] = cld(pb + kc);
    }
#pragma unroll
    for (int s = 0; s < NS; ++s) {
      const int kk = 4 * (s0 + s) + lk;
      const bool in = s < spw && kk < l;
      fa[s] = (in && row < l) ? fa[s] : T(0);
      fb[s] = (in && col < l) ? fb[s] : T(0);
    }
#pragma unroll
    for (int s = 0; s < NS; ++s)
      if (s < spw) acc = SM<T>::mfma(fa[s], fb[s], acc);
  }
  red[wave][lane] = acc;
  __syncthreads();
  if (live && q == 0) {
    typename SM<T>::acc sum = red[wave][lane];
#pragma unroll
    for (int w = 1; w < wpj; ++w) {
      const typename SM<T>::acc o = red[wave + w][lane];
      for (int g = 0; g < 4; ++g) sum[g] += o[g];
    }
    T sq = 0;
    const i64 ll = (i64)l * l;
#pragma unroll
    for (int g = 0; g < 4; ++g) {
      const int r = tm * 16 + SM<T>::row(lane, g), c = tn * 16 + lr;
      if (r < l && c < l) {
        const i64 e = r + (i64)c * l;
        const T dl = (r == c) ? T(1) : T(0);
        T v = p.alpha * sum[g] + p.diag * dl;
        if (p.Cin) v += p.cin_coef * pcin[g];
        if (epi & kEpiBlock) v += block_at(p.coef, dl, pz[g], pz2[g], pz3[g]);
        if (p.out) cst(p.out + e, v);
        if (epi & kEpiStart1) {
          const T z = pz[g];
          cst(p.out2 + e, dl - T(2) * z - T(8) * v);
          cst(p.out3 + e, dl + T(2) * z);
        }
        if (epi & kEpiStart2) cst(p.out2 + e, T(0.25) * (T(5) * dl - pcin[g]));
        if (epi & kEpiTop) {
          const T cv = block_at(p.coef, dl, pz[g], pz2[g], v);
          const T gv = block_at(p.coef + kSeriesTerms, dl, pz[g], pz2[g], v);
          cst(p.out2 + e, cv);
          cst(p.out3 + e, gv);
          if (epi & kEpiTopLast) {  // read only after the kernel: plain stores
            sq += (cv - dl) * (cv - dl);
            T *w3 = a.W3 + l + r + (i64)c * 2 * l;
            w3[0] = a.ninv * gv;
            w3[2 * ll] = gv;
            w3[4 * ll] = gv;
          }
        }
        if (epi & kEpiNormSq) sq += v * v;
        if (epi & kEpiErrSq) sq += (v - dl) * (v - dl);
        if (epi & (kEpiW3Low | kEpiW3Up)) {  // read only after the kernel: plain stores
          T *w3 = a.W3 + ((epi & kEpiW3Low) ? l : 0) + r + (i64)c * 2 * l;
          w3[0] = (epi & kEpiW3Low) ? a.ninv * v : v;
          w3[2 * ll] = v;
          w3[4 * ll] = v;
        }
      }
    }
    if (epi & (kEpiNormSq | kEpiErrSq | kEpiTopLast)) {
      for (int off = 32; off > 0; off >>= 1) sq += __shfl_xor(sq, off, 64);
      if (lane == 0) cst(a.part + tile + ((epi & kEpiTopLast) ? kPartErr : 0), sq);
    }
  }
  __syncthreads();  // red is reused by the next round
}

template <typename T>
__device__ __forceinline__ void round1(const ChainArgs<T> &a, const ChainProd<T> &p,
                                       typename SM<T>::acc (*red)[64]) {
  chain_round<T, false>(a, p, p, red);
}
template <typename T>
__device__ __forceinline__ void round2(const ChainArgs<T> &a, const ChainProd<T> &p0,
                                       const ChainProd<T> &p1, typename SM<T>::acc (*red)[64]) {
  chain_round<T, true>(a, p0, p1, red);
}

// Paterson-Stockmeyer blocks (of 4 powers) needed for the Taylor series of f and g to reach
// a quarter of the unit roundoff for 4 ||Z||_2 <= rho; 0 when more than kSeriesBlocks would be.
// The tail past degree 4t - 1 is bounded by the tabulated terms up to kSeriesTerms - 1 plus the
// last one again (for rho < 1/2 the terms decay at least geometrically with ratio < 1/2).
template <typename T>
__device__ int series_blocks(const T *coef, double rho) {
  const double eps = (sizeof(T) == 8) ? 0x1p-53 : 0x1p-24;
  if (!(rho < 0.5)) return 0;
  const double x = rho / 4.0;
  double tf[kSeriesTerms], tg[kSeriesTerms];
  double xp = 1.0;
  for (int k = 0; k < kSeriesTerms; ++k) {
    tf[k] = fabs((double)coef[k]) * xp;
    tg[k] = fabs((double)coef[kSeriesTerms + k]) * xp;
    xp *= x;
  }
  // tails from degree k on (suffix sums); the smallest t whose tail from 4t on is small enough
  double rf = tf[kSeriesTerms - 1], rg = tg[kSeriesTerms - 1];
  int best = 0;
  for (int k = kSeriesTerms - 1; k >= 4; --k) {
    rf += tf[k];
    rg += tg[k];
    if ((k & 3) == 0 && k / 4 <= kSeriesBlocks && rf <= 0.25 * eps && rg <= 0.25 * eps)
      best = k / 4;
  }
  return best;
}

template <typename T>
__device__ void chain_body(const ChainArgs<T> &a, typename SM<T>::acc (*red)[64], T *sh) {
  const int l = a.l, ntile = a.tiles_m * a.tiles_m;
  const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
  auto prod = [&](const T *A, const T *B, T *out, bool ta, T alpha, T diag) {
    return ChainProd<T>{A, B, nullptr, out, nullptr, nullptr, ta, alpha, diag, T(0), 0,
                        nullptr};
  };
  // Z := Dg - D^T D  (@x@_integrate_kolmogorov_nagumo.c:187-190)
  {
    ChainProd<T> p = prod(a.D, a.D, a.Z, true, T(-1), T(0));
    p.Cin = a.Dg;
    p.cin_coef = T(1);
    round1<T>(a, p, red);
  }
  if (!chain_barrier(a.bar)) goto timeout;
  // Z2 := Z Z and ||Z2||_F.  Stage 1 of Newton-Schulz (P = (I - 4Z)^(1/2)) starts from
  // Y0 = I - 4Z, W0 = I, E0 = 3I - Y0; its first step Y1 = Y0 E0 / 2 = I - 2Z - 8Z^2,
  // W1 = E0 / 2 = I + 2Z comes with Z2.
  {
    ChainProd<T> p = prod(a.Z, a.Z, a.Z2, false, T(1), T(0));
    p.epi = kEpiStart1 | kEpiNormSq;
    p.out2 = a.Y[1];
    p.out3 = a.W[1];
    round1<T>(a, p, red);
  }
  if (!chain_barrier(a.bar)) goto timeout;
  {
    // 4 ||Z||_2 <= 4 ||Z^2||_F^(1/2)
    const double rho2 = 4.0 * sqrt(sqrt((double)chain_sum<T>(a.part, ntile, sh)));
    const int t = a.mode == 0 ? series_blocks<T>(a.coef, rho2) : 0;
    if (t > 0) {
      // Paterson-Stockmeyer: C = sum_i Bf_i (Z^4)^i with Bf_i = sum_j cf[4i + j] Z^j (j < 4),
      // likewise C^-1; Horner from the top block.  Z3 := Z Z2 (top blocks from it), Z4 := Z2 Z2.
      if (lead) {
        a.scal[1] = T(0);
        a.scal[2] = (T)rho2;
      }
      int cur = 0;
      {
        ChainProd<T> p3 = prod(a.Z, a.Z2, a.Z3, false, T(1), T(0));
        p3.epi = kEpiTop;
        p3.coef = a.coef + 4 * (t - 1);
        p3.out2 = a.Y[cur];
        p3.out3 = a.W[cur];
        if (t == 1) {
          // C and C^-1 are the top blocks: ||C - I||_F partials and the C^-1 halves of W3 come
          // with them, and Z^4 is not needed
          p3.epi |= kEpiTopLast;
          round1<T>(a, p3, red);
        } else {
          const ChainProd<T> p4 = prod(a.Z2, a.Z2, a.Z4, false, T(1), T(0));
          round2<T>(a, p3, p4, red);
        }
      }
      if (!chain_barrier(a.bar)) goto timeout;
      for (int i = t - 2; i >= 0; --i) {
        // Pf := Pf Z4 + Bf_i, Pg := Pg Z4 + Bg_i
        ChainProd<T> pf = prod(a.Y[cur], a.Z4, a.Y[cur ^ 1], false, T(1), T(0));
        ChainProd<T> pg = prod(a.W[cur], a.Z4, a.W[cur ^ 1], false, T(1), T(0));
        pf.epi = kEpiBlock | (i == 0 ? kEpiErrSq : 0);
        pg.epi = kEpiBlock | (i == 0 ? kEpiW3Low : 0);
        pf.coef = a.coef + 4 * i;
        pg.coef = a.coef + kSeriesTerms + 4 * i;
        round2<T>(a, pf, pg, red);
        if (!chain_barrier(a.bar)) goto timeout;
        cur ^= 1;
      }
      // F := C - D C^-1 (:218) into the F halves of W3
      {
        ChainProd<T> pf = prod(a.D, a.W[cur], nullptr, false, T(-1), T(0));
        pf.Cin = a.Y[cur];
        pf.cin_coef = T(1);
        pf.epi = kEpiW3Up;
        round1<T>(a, pf, red);
      }
      const T err2 = chain_sum<T>(a.part + (t == 1 ? kPartErr : 0), ntile, sh);
      if (lead) {
        a.scal[0] = sqrt(err2);
        a.scal[3] =
            __hip_atomic_load(a.bar + kBarAbort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                ? T(2)
                : T(0);
      }
      return;
    }
  }
  // Newton-Schulz.  Z4 := Z2 Z2 and ||Z4||_F: rho = 4 ||Z4||_F^(1/4) (tighter than ||Z2||_F)
  {
    ChainProd<T> p = prod(a.Z2, a.Z2, nullptr, false, T(1), T(0));
    p.epi = kEpiNormSq;
    round1<T>(a, p, red);
  }
  if (!chain_barrier(a.bar)) goto timeout;
  {
    const T z4 = sqrt(chain_sum<T>(a.part, ntile, sh));
    const double rho = 4.0 * sqrt(sqrt((double)z4));
    int k1 = 0, k2 = 0;
    const bool ok = ns_counts<T>(rho, &k1, &k2) && a.mode != 1;
    if (lead) {
      a.scal[1] = z4;
      a.scal[2] = (T)rho;
      if (!ok) a.scal[3] = T(1);
    }
    if (!ok) return;  // every workgroup takes the same decision
    int cur = 1;
    // one coupled Newton-Schulz step on (Y[cur], W[cur]): E := 3I - W Y; Y' := Y E / 2,
    // W' := E W / 2.  The last step of stage 2 also leaves ||C - I||_F partials and C^-1 in W3.
    auto step = [&](bool last) -> bool {
      const ChainProd<T> pe = prod(a.W[cur], a.Y[cur], a.E, false, T(-1), T(3));
      round1<T>(a, pe, red);
      if (!chain_barrier(a.bar)) return false;
      ChainProd<T> py = prod(a.Y[cur], a.E, a.Y[cur ^ 1], false, T(0.5), T(0));
      ChainProd<T> pw = prod(a.E, a.W[cur], a.W[cur ^ 1], false, T(0.5), T(0));
      if (last) {
        py.epi = kEpiErrSq;
        pw.epi = kEpiW3Low;
      }
      round2<T>(a, py, pw, red);
      if (!chain_barrier(a.bar)) return false;
      cur ^= 1;
      return true;
    };
    for (int it = 1; it < k1; ++it)
      if (!step(false)) goto timeout;
    // Stage 2 (C = ((I + P)/2)^(1/2)) starts from Y0 = (I + P)/2, W0 = I, E0 = 3I - Y0; its
    // first step is Y1 = Y0 E0 / 2 = (5I + 4P - P^2) / 8 and W1 = E0 / 2 = (5I - P) / 4.
    {
      ChainProd<T> p = prod(a.Y[cur], a.Y[cur], a.Y[cur ^ 1], false, T(-0.125), T(0.625));
      p.Cin = a.Y[cur];
      p.cin_coef = T(0.5);
      p.epi = kEpiStart2 | (k2 == 1 ? kEpiErrSq : 0);
      p.out2 = a.W[cur ^ 1];
      round1<T>(a, p, red);
    }
    if (!chain_barrier(a.bar)) goto timeout;
    cur ^= 1;
    if (k2 == 1) {
      // C^-1 = W[cur] came from the elementwise part of the first step: the W3 halves
      const int tile = 2 * blockIdx.x + (threadIdx.x >> 8);
      const int tm = tile % a.tiles_m, tn = tile / a.tiles_m;
      if (tile < ntile) {
        const int i = tm * 16 + (threadIdx.x & 15), j = tn * 16 + ((threadIdx.x >> 4) & 15);
        if (i < l && j < l) {
          const i64 e = i + (i64)j * l, ll = (i64)l * l;
          const T ci = cld(a.W[cur] + e);
          T *w3 = a.W3 + l + i + (i64)j * 2 * l;
          w3[0] = a.ninv * ci;
          w3[2 * ll] = ci;
          w3[4 * ll] = ci;
        }
      }
    }
    for (int it = 1; it < k2; ++it)
      if (!step(it == k2 - 1)) goto timeout;
    // F := C - D C^-1 (:218) into the F halves of W3
    {
      ChainProd<T> pf = prod(a.D, a.W[cur], nullptr, false, T(-1), T(0));
      pf.Cin = a.Y[cur];
      pf.cin_coef = T(1);
      pf.epi = kEpiW3Up;
      round1<T>(a, pf, red);
    }
    const T err2 = chain_sum<T>(a.part, ntile, sh);
    if (lead) {
      a.scal[0] = sqrt(err2);
      a.scal[3] = __hip_atomic_load(a.bar + kBarAbort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                      ? T(2)
                      : T(0);
    }
    return;
  }
timeout:
  if (threadIdx.x == 0) a.scal[3] = T(2);
}

template <typename T>
__global__ __launch_bounds__(kChainThreads) void kn_chain_kernel(ChainArgs<T> a) {
  __shared__ typename SM<T>::acc red[kChainWaves][64];
  __shared__ T sh[16];
  const int ll = a.l * a.l;
  const int gtid = blockIdx.x * kChainThreads + threadIdx.x, gstride = gridDim.x * kChainThreads;
  // [D | Dg] := sum of the partials / N  (@x@_integrate_kolmogorov_nagumo.c:181-186), fixed order
  {
    const i64 n2 = 2 * (i64)ll;
    for (i64 e = gtid; e < n2; e += gstride) {
      T s = 0;
      for (int q = 0; q < a.nparts; ++q) s += a.parts[q * n2 + e];
      cst(a.D + e, a.ninv * s);  // Dg follows D (ld l)
    }
  }
  if (!chain_barrier(a.bar)) {
    if (threadIdx.x == 0) a.scal[3] = T(2);
    return;
  }
  chain_body<T>(a, red, sh);
}

__global__ void zero_u64_kernel(unsigned long long *p, int n) {
  for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0ull;
}

// Taylor coefficients at 0 of f(x) = (1/2 + (1/4 - x)^(1/2))^(1/2) and g = 1/f (in powers of x),
// from the binomial series of (1 - u)^(1/2) (u = 4x) by power-series square root and reciprocal,
// in long double.
void series_coefficients(long double *cf, long double *cg, int n) {
  std::vector<long double> a(n), w(n), h(n), r(n);
  a[0] = 1.0L;
  for (int k = 1; k < n; ++k) a[k] = a[k - 1] * ((long double)k - 1.5L) / (long double)k;
  w[0] = 2.0L;  // 1 + (1 - u)^(1/2)
  for (int k = 1; k < n; ++k) w[k] = a[k];
  h[0] = std::sqrt(2.0L);  // h = w^(1/2)
  for (int k = 1; k < n; ++k) {
    long double s = w[k];
    for (int j = 1; j < k; ++j) s -= h[j] * h[k - j];
    h[k] = s / (2.0L * h[0]);
  }
  r[0] = 1.0L / h[0];  // r = 1 / h
  for (int k = 1; k < n; ++k) {
    long double s = 0.0L;
    for (int j = 1; j <= k; ++j) s += h[j] * r[k - j];
    r[k] = -s / h[0];
  }
  const long double rt2 = std::sqrt(2.0L);
  long double p4 = 1.0L;  // f = h / sqrt(2), g = sqrt(2) r; u = 4x
  for (int k = 0; k < n; ++k) {
    cf[k] = h[k] / rt2 * p4;
    cg[k] = r[k] * rt2 * p4;
    p4 *= 4.0L;
  }
}

}  // namespace

void zero_u64(unsigned long long *p, int n) {
  hipLaunchKernelGGL(zero_u64_kernel, dim3(1), dim3(64), 0, ctx().stream, p, n);
  LAUNCH_CHECK();
}

template <typename T>
void kn_chain(int l, T ninv, const T *parts, int nparts, T *DDg, const KnChainWork<T> &w,
              T *W3, T *scal, int mode) {
  ISVD_CHECK(l >= 1 && l <= kSmallMaxK, "kn_chain supports 1 <= l <= %d (got %d)", kSmallMaxK,
             l);
  const int tiles_m = (l + 15) / 16;
  ISVD_CHECK(tiles_m * tiles_m <= 64, "kn_chain: too many tiles");
  ISVD_CHECK(nparts >= 1, "kn_chain: no partial sums");
  const i64 ll = (i64)l * l;
  ChainArgs<T> a{};
  a.l = l;
  a.tiles_m = tiles_m;
  a.nparts = nparts;
  a.mode = mode;
  a.ninv = ninv;
  a.parts = parts;
  a.D = DDg;
  a.Dg = DDg + ll;
  a.Z = w.buf;
  a.Z2 = w.buf + ll;
  a.Z3 = w.buf + 2 * ll;
  a.Z4 = w.buf + 3 * ll;
  a.Y[0] = w.buf + 4 * ll;
  a.Y[1] = w.buf + 5 * ll;
  a.W[0] = w.buf + 6 * ll;
  a.W[1] = w.buf + 7 * ll;
  a.E = w.buf + 8 * ll;
  a.part = w.buf + 9 * ll;
  a.coef = w.buf + 9 * ll + kKnChainPart;
  a.W3 = W3;
  a.scal = scal;
  a.bar = w.bar;
  const int nwg = kn_chain_blocks(l);
  hipLaunchKernelGGL(kn_chain_kernel<T>, dim3(nwg), dim3(kChainThreads), 0, ctx().stream, a);
  LAUNCH_CHECK();
}

template <typename T>
void kn_chain_init(const KnChainWork<T> &w, int l) {
  static const std::vector<T> table = [] {
    long double cf[kSeriesTerms], cg[kSeriesTerms];
    series_coefficients(cf, cg, kSeriesTerms);
    std::vector<T> v(2 * kSeriesTerms);
    for (int k = 0; k < kSeriesTerms; ++k) {
      v[k] = (T)cf[k];
      v[kSeriesTerms + k] = (T)cg[k];
    }
    return v;
  }();
  const i64 ll = (i64)l * l;
  HIP_CHECK(hipMemcpyAsync(w.buf + 9 * ll + kKnChainPart, table.data(), sizeof(T) * table.size(),
                           hipMemcpyHostToDevice, ctx().stream));
  zero_u64(w.bar, kKnChainBarWords);
}

template <typename T>
void inv_small(i64 n, T *A, i64 lda, int *info_dev) {
  ISVD_CHECK(n >= 1 && n <= 128, "inv_small supports 1 <= n <= 128 (got %lld)", n);
  const size_t shmem = sizeof(T) * (size_t)n * (size_t)(n + 1);
  auto kern = gj_inverse_lds_kernel<T>;
  HIP_CHECK(hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)shmem));
  hipLaunchKernelGGL(kern, dim3(1), dim3(kGjThreads), shmem, ctx().stream, (int)n, A, lda,
                     info_dev);
  LAUNCH_CHECK();
}

#define INST(T)                                                                               \
  template void small_gemm<T>(char, char, i64, i64, i64, T, const T *, i64, const T *, i64, T, \
                              T *, i64, T);                                                    \
  template void inv_small<T>(i64, T *, i64, int *);                                           \
  template void kn_chain<T>(int, T, const T *, int, T *, const KnChainWork<T> &, T *, T *, int); \
  template void kn_chain_init<T>(const KnChainWork<T> &, int);
INST(double)
INST(float)
#undef INST

}  // namespace isvd
