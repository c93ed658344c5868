// Small device kernels: matrix copies / fills and deterministic reductions.
#include "la.hpp"

namespace isvd {

namespace {

template <typename T>
__global__ void omatcopy_kernel(bool trans, i64 m, i64 n, T alpha, const T *A, i64 lda, T *B,
                                i64 ldb, i64 sA, i64 sB) {
  const T *Ab = A + (i64)blockIdx.y * sA;
  T *Bb = B + (i64)blockIdx.y * sB;
  for (i64 e = blockIdx.x * (i64)blockDim.x + threadIdx.x; e < m * n;
       e += (i64)gridDim.x * blockDim.x) {
    const i64 i = e % m, j = e / m;  // result element (i, j)
    Bb[i + j * ldb] = alpha * (trans ? Ab[j + i * lda] : Ab[i + j * lda]);
  }
}

template <typename T>
__global__ void fill_kernel(i64 m, i64 n, T value, T *A, i64 lda, T diag, i64 sA) {
  T *Ab = A + (i64)blockIdx.y * sA;
  for (i64 e = blockIdx.x * (i64)blockDim.x + threadIdx.x; e < m * n;
       e += (i64)gridDim.x * blockDim.x) {
    const i64 i = e % m, j = e / m;
    Ab[i + j * lda] = value + ((i == j) ? diag : T(0));
  }
}

// Two passes with a fixed reduction tree (bitwise reproducible for a given n): workgroup b sums
// the chunk [b chunk, (b + 1) chunk) into part[b]; one workgroup then sums the parts in order.
constexpr int kDotThreads = 256, kDotParts = 512;

template <typename T>
__device__ T block_sum(T s, T *red) {
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  T t = 0;
  for (int w = 0; w < kDotThreads / 64; ++w) t += red[w];
  return t;
}

template <typename T>
__global__ __launch_bounds__(kDotThreads) void dot_parts_kernel(i64 n, i64 chunk, const T *x,
                                                                const T *y, T *part) {
  __shared__ T red[kDotThreads / 64];
  const i64 beg = (i64)blockIdx.x * chunk, end = (beg + chunk < n) ? beg + chunk : n;
  T s = 0;
  for (i64 e = beg + threadIdx.x; e < end; e += kDotThreads) s += x[e] * y[e];
  const T t = block_sum(s, red);
  if (threadIdx.x == 0) part[blockIdx.x] = t;
}

template <typename T>
__global__ __launch_bounds__(kDotThreads) void dot_final_kernel(int nparts, const T *part,
                                                                T *out) {
  __shared__ T red[kDotThreads / 64];
  T s = 0;
  for (int e = threadIdx.x; e < nparts; e += kDotThreads) s += part[e];
  const T t = block_sum(s, red);
  if (threadIdx.x == 0) out[0] = t;
}

}  // namespace

template <typename T>
void omatcopy(char trans, i64 m, i64 n, T alpha, const T *A, i64 lda, T *B, i64 ldb, i64 batch,
              i64 sA, i64 sB) {
  if (m <= 0 || n <= 0 || batch <= 0) return;
  const bool t = (trans == 'T' || trans == 't');
  int blocks = ceil_div(m * n, 256);
  if (blocks > 8192) blocks = 8192;
  hipLaunchKernelGGL(omatcopy_kernel<T>, dim3(blocks, (unsigned)batch), dim3(256), 0,
                     ctx().stream, t, m, n, alpha, A, lda, B, ldb, sA, sB);
  LAUNCH_CHECK();
}

template <typename T>
void fill(i64 m, i64 n, T value, T *A, i64 lda, T diag, i64 batch, i64 sA) {
  if (m <= 0 || n <= 0 || batch <= 0) return;
  int blocks = ceil_div(m * n, 256);
  if (blocks > 8192) blocks = 8192;
  hipLaunchKernelGGL(fill_kernel<T>, dim3(blocks, (unsigned)batch), dim3(256), 0, ctx().stream,
                     m, n, value, A, lda, diag, sA);
  LAUNCH_CHECK();
}

template <typename T>
void dot(i64 n, const T *x, const T *y, T *result) {
  i64 chunk = (n + kDotParts - 1) / kDotParts;
  if (chunk < 4 * kDotThreads) chunk = 4 * kDotThreads;
  const int nparts = (int)((n + chunk - 1) / chunk);
  if (nparts <= 1) {
    hipLaunchKernelGGL(dot_parts_kernel<T>, dim3(1), dim3(kDotThreads), 0, ctx().stream, n,
                       n > 0 ? n : 1, x, y, result);
    LAUNCH_CHECK();
    return;
  }
  DBuf<T> part(nparts);
  hipLaunchKernelGGL(dot_parts_kernel<T>, dim3(nparts), dim3(kDotThreads), 0, ctx().stream, n,
                     chunk, x, y, part.get());
  LAUNCH_CHECK();
  hipLaunchKernelGGL(dot_final_kernel<T>, dim3(1), dim3(kDotThreads), 0, ctx().stream, nparts,
                     (const T *)part.get(), result);
  LAUNCH_CHECK();
}

#define INST(T)                                                                              \
  template void omatcopy<T>(char, i64, i64, T, const T *, i64, T *, i64, i64, i64, i64);    \
  template void fill<T>(i64, i64, T, T *, i64, T, i64, i64);                                \
  template void dot<T>(i64, const T *, const T *, T *);
INST(double)
INST(float)
#undef INST

}  // namespace isvd
