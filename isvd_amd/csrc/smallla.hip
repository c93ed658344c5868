// l x l (l <= 128) building blocks of the integration loops:
//  - small_gemm: C := alpha op(A) op(B) + beta C + diag I for up to two independent problems
//    per launch, one wave per 16 x 16 output tile (v_mfma_f64_16x16x4f64 / _f32_16x16x4f32),
//    operands read straight from L2 (no LDS, no workspace), so a 128^3 product is one short
//    launch;
//  - kn_matfun: C = (I/2 + (I/4 - Z)^{1/2})^{1/2} and C^-1 by two coupled Newton-Schulz
//    square-root iterations (Higham, Functions of Matrices, 2008, eq. 6.35), valid while
//    ||4 Z||_2 < 1 (the caller checks ||4 Z||_F);
//  - inv_small: Gauss-Jordan inverse with partial pivoting, matrix held in LDS.
#include "la.hpp"

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

// out[0] = ||A - c I||_F (one workgroup, fixed order)
template <typename T>
__global__ __launch_bounds__(1024) void frob_shift_kernel(int n, const T *A, T c, T *out) {
  __shared__ T red[16];
  T part = 0;
  for (int e = threadIdx.x; e < n * n; e += 1024) {
    const T v = A[e] - ((e % n == e / n) ? c : T(0));
    part += v * v;
  }
  for (int off = 32; off > 0; off >>= 1) part += __shfl_xor(part, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = part;
  __syncthreads();
  if (threadIdx.x == 0) {
    T t = 0;
    for (int w = 0; w < 16; ++w) t += red[w];  // fixed order
    out[0] = sqrt(t);
  }
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
  __shared__ int ipiv[128], pos[128];
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
      for (int off = 32; off > 0; off >>= 1) {
        const T ob = __shfl_xor(best, off, 64);
        const int oi = __shfl_xor(bi, off, 64);
        if (ob > best || (ob == best && oi < bi)) { best = ob; bi = oi; }
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
    for (int e = tid; e < n * n; e += kGjThreads) {
      const int r = e / n, c = e % n;
      T v;
      if (r == j)
        v = (c == j) ? ipv : rowj[c] * ipv;
      else
        v = (c == j) ? -colj[r] * ipv : S[r * ld + c] - colj[r] * ipv * rowj[c];
      S[r * ld + c] = v;
    }
    __syncthreads();
  }
  if (bad_sh) {
    if (tid == 0) *info = bad_sh;
    return;
  }
  if (tid == 0) {
    int idx[128];
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

// C, Cinv (n x n, ld n) from Z; work must hold 6 n*n entries.  Returns nothing; the caller
// decides from ||4 Z||_F whether this is valid.
// Start of one Newton-Schulz stage: Y := a * sym(S) + b * I and W := I (one launch)
template <typename T>
__global__ void ns_init_kernel(int n, const T *S, T *Y, T *W, T a, T b) {
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < n * n; e += gridDim.x * blockDim.x) {
    const int i = e % n, j = e / n;
    const T sy = T(0.5) * (S[i + j * n] + S[j + i * n]);
    Y[e] = a * sy + ((i == j) ? b : T(0));
    W[e] = (i == j) ? T(1) : T(0);
  }
}

// End of the matrix functions: C := sym(Cr), Cinv := sym(Cir), and, when err is not null,
// err[0] = ||C - I||_F (one workgroup, fixed reduction order)
template <typename T>
__global__ __launch_bounds__(1024) void ns_finish_kernel(int n, const T *Cr, const T *Cir, T *C,
                                                         T *Cinv, T *err) {
  __shared__ T red[16];
  T part = 0;
  for (int e = threadIdx.x; e < n * n; e += 1024) {
    const int i = e % n, j = e / n;
    const T c = T(0.5) * (Cr[i + j * n] + Cr[j + i * n]);
    C[e] = c;
    Cinv[e] = T(0.5) * (Cir[i + j * n] + Cir[j + i * n]);
    const T d = c - ((i == j) ? T(1) : T(0));
    part += d * d;
  }
  if (!err) return;
  for (int off = 32; off > 0; off >>= 1) part += __shfl_xor(part, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = part;
  __syncthreads();
  if (threadIdx.x == 0) {
    T t = 0;
    for (int w = 0; w < 16; ++w) t += red[w];  // fixed order
    err[0] = sqrt(t);
  }
}

template <typename T>
void kn_matfun(i64 n, const T *Z, T *C, T *Cinv, T *work, int iters1, int iters2, T *err) {
  const int nn = (int)(n * n);
  T *Y = work, *W = work + nn, *E = work + 2 * nn, *Y2 = work + 3 * nn, *W2 = work + 4 * nn;
  // Y_0 and W_0 = I are set by the caller of ns
  auto ns = [&](T *y, T *w, T *y2, T *w2, int iters, T *&yo, T *&wo) {
    for (int it = 0; it < iters; ++it) {
      // E = 3 I - W Y
      small_gemm<T>('N', 'N', n, n, n, T(-1), w, n, y, n, T(0), E, n, T(3));
      // Y' = Y E / 2, W' = E W / 2 (one launch)
      SmallProblem<T> py{y, E, y2, (int)n, (int)n, (int)n, false, false, T(0.5), T(0), T(0)};
      SmallProblem<T> pw{E, w, w2, (int)n, (int)n, (int)n, false, false, T(0.5), T(0), T(0)};
      small_gemm2<T>((int)n, (int)n, (int)n, py, &pw);
      std::swap(y, y2);
      std::swap(w, w2);
    }
    yo = y;
    wo = w;
  };
  // stage 1: P = (I - 4 sym(Z))^{1/2} = 2 (I/4 - Z)^{1/2}
  hipLaunchKernelGGL(ns_init_kernel<T>, dim3(grid_of(nn)), dim3(256), 0, ctx().stream, (int)n,
                     Z, Y, W, T(-4), T(1));
  LAUNCH_CHECK();
  T *P, *Pi;
  ns(Y, W, Y2, W2, iters1, P, Pi);
  // stage 2: C^2 = I/2 + P/2 ; (C, C^-1) = NS square root
  T *Ys = (P == Y) ? Y2 : Y;
  T *Ws = (P == Y) ? W2 : W;
  hipLaunchKernelGGL(ns_init_kernel<T>, dim3(grid_of(nn)), dim3(256), 0, ctx().stream, (int)n,
                     (const T *)P, Ys, Ws, T(0.5), T(0.5));
  LAUNCH_CHECK();
  T *Cr, *Cir;
  T *spare1 = (Ys == Y) ? Y2 : Y;
  T *spare2 = (Ws == W) ? W2 : W;
  ns(Ys, Ws, spare1, spare2, iters2, Cr, Cir);
  // symmetrize the results (and ||C - I||_F)
  hipLaunchKernelGGL(ns_finish_kernel<T>, dim3(1), dim3(1024), 0, ctx().stream, (int)n,
                     (const T *)Cr, (const T *)Cir, C, Cinv, err);
  LAUNCH_CHECK();
}

template <typename T>
void frob_shift(i64 n, const T *A, T c, T *out) {
  hipLaunchKernelGGL(frob_shift_kernel<T>, dim3(1), dim3(1024), 0, ctx().stream, (int)n, A, c,
                     out);
  LAUNCH_CHECK();
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
  template void kn_matfun<T>(i64, const T *, T *, T *, T *, int, int, T *);                    \
  template void frob_shift<T>(i64, const T *, T, T *);                                         \
  template void inv_small<T>(i64, T *, i64, int *);
INST(double)
INST(float)
#undef INST

}  // namespace isvd
