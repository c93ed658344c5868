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

// one workgroup, fixed reduction tree -> bitwise reproducible
template <typename T>
__global__ __launch_bounds__(1024) void dot_kernel(i64 n, const T *x, const T *y, T *out) {
  __shared__ T red[16];
  T s = 0;
  for (i64 e = threadIdx.x; e < n; e += 1024) s += x[e] * y[e];
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    T t = 0;
    for (int w = 0; w < 16; ++w) t += red[w];
    out[0] = t;
  }
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
  hipLaunchKernelGGL(dot_kernel<T>, dim3(1), dim3(1024), 0, ctx().stream, n, x, y, result);
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
