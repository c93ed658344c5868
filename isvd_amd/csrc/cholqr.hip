// CholeskyQR2 for the tall-skinny QR of the pipeline's sketch blocks (and of Z in the TS
// postprocessing): two passes of  W = Y^T Y (MFMA GEMM),  W = R^T R (Cholesky),  Q = Y R^-1
// (MFMA GEMM against the explicit triangular inverse).  One pass leaves an orthogonality error
// of about u kappa(Y)^2; the second pass, on a basis with kappa close to 1, brings it to O(u),
// and the residual ||Y - Q R|| is O(u) ||Y|| as for Householder QR, whenever
// kappa(Y) <~ u^-1/2 (Y. Yamamoto, Y. Nakatsukasa, Y. Yanagisawa, T. Fukaya, "Roundoff error
// analysis of the CholeskyQR2 algorithm", ETNA 44, 2015).  That condition is certified on the
// device: a pivot that is not positive in either factorization, or ||Q1^T Q1 - I||_F > 1/2
// after the first pass, makes the call return false with Y untouched, and the caller runs its
// reference factorization instead.  The l x l steps run in FP64 also for FP32 data.
#include "la.hpp"
#include "mx.hpp"

#include <cstdlib>
#include <type_traits>
#include <vector>

namespace isvd {

namespace {

constexpr int kCqThreads = 1024;
constexpr int kCqGroups = kCqThreads / 128;  // column groups of the trailing update
constexpr int kCqLanes = kCqThreads / 128;   // lanes per row of the triangular inverse

// One workgroup per l x l block: W (full symmetric storage, upper triangle read) -> R (upper
// Cholesky factor, W = R^T R, lower part zero) and Rinv = R^-1 (upper).  With check, first
// ||W - I||_F is formed and more than 1/2 marks the block failed.  status[b] := 1 on failure
// (non-positive or non-finite pivot); the outputs of a failed block are not used.
template <typename T>
__global__ __launch_bounds__(kCqThreads) void chol_inv_kernel(int n, const T *__restrict__ W,
                                                              T *__restrict__ R,
                                                              T *__restrict__ Rinv, int check,
                                                              int *__restrict__ status) {
  extern __shared__ unsigned char smem_raw[];
  T *A = reinterpret_cast<T *>(smem_raw);  // n x n, ld n + 1
  __shared__ T tmp[128];
  __shared__ T red[kCqThreads];
  __shared__ int fail;
  __shared__ T piv;
  const int b = blockIdx.x, t = threadIdx.x, ld = n + 1;
  const T *Wb = W + (i64)b * n * n;
  if (t == 0) fail = 0;
  T dev = 0;
  for (int e = t; e < n * n; e += kCqThreads) {
    const int i = e % n, j = e / n;
    const T v = (i <= j) ? Wb[i + (i64)j * n] : Wb[j + (i64)i * n];
    A[i + j * ld] = v;
    const T d = v - (i == j ? T(1) : T(0));
    dev += d * d;
  }
  red[t] = dev;
  __syncthreads();
  if (check) {
    for (int h = kCqThreads / 2; h > 0; h >>= 1) {
      if (t < h) red[t] += red[t + h];
      __syncthreads();
    }
    if (t == 0 && !(red[0] <= T(0.25))) fail = 1;  // ||W - I||_F > 1/2, or not finite
  }
  // right-looking Cholesky of the upper triangle, in place: row k of R at step k
  for (int k = 0; k < n; ++k) {
    if (t == 0) {
      const T d = A[k + k * ld];
      if (!(d > T(0)) || !(d < T(INFINITY))) fail = 1;
      A[k + k * ld] = (d > T(0)) ? sqrt(d) : T(1);
    }
    __syncthreads();
    const T rkk = A[k + k * ld];
    for (int j = k + 1 + t; j < n; j += kCqThreads) A[k + j * ld] /= rkk;
    __syncthreads();
    // A(i, j) -= R(k, i) R(k, j) for k < i <= j: a fixed row i per thread, columns strided by
    // the number of column groups
    {
      const int i = k + 1 + (t & 127);
      if (i < n) {
        const T rki = A[k + i * ld];
        for (int j = i + ((t >> 7) + kCqGroups - ((i - k - 1) % kCqGroups)) % kCqGroups; j < n;
             j += kCqGroups)
          A[i + j * ld] -= rki * A[k + j * ld];
      }
    }
    __syncthreads();
  }
  T *Rb = R + (i64)b * n * n, *Ib = Rinv + (i64)b * n * n;
  for (int e = t; e < n * n; e += kCqThreads) {
    const int i = e % n, j = e / n;
    Rb[i + (i64)j * n] = (i <= j) ? A[i + j * ld] : T(0);
  }
  __syncthreads();
  // in-place inverse of the upper triangle, column by column (the leading j x j block already
  // holds its inverse T): x = R(0:j, j);  Rinv(0:j, j) = -T x / R(j, j);  Rinv(j, j) = 1 / R(j, j)
  // (the pivot is copied in the first phase: thread 0 replaces A(j, j) in the second, where no
  // other thread reads it)
  // Each row's sum is split over kCqLanes consecutive lanes (k = i + q, i + q + kCqLanes, ...),
  // added in a fixed tree.
  for (int j = 0; j < n; ++j) {
    for (int i = t; i < j; i += kCqThreads) tmp[i] = A[i + j * ld];
    if (t == 0) piv = A[j + j * ld];
    __syncthreads();
    const T rjj = piv;
    {
      const int i = t / kCqLanes, q = t % kCqLanes;
      T s = 0;
      if (i < j)
        for (int k = i + q; k < j; k += kCqLanes) s += A[i + k * ld] * tmp[k];
      for (int off = kCqLanes / 2; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
      if (i < j && q == 0) A[i + j * ld] = -s / rjj;
    }
    if (t == 0) A[j + j * ld] = T(1) / rjj;
    __syncthreads();
  }
  for (int e = t; e < n * n; e += kCqThreads) {
    const int i = e % n, j = e / n;
    Ib[i + (i64)j * n] = (i <= j) ? A[i + j * ld] : T(0);
  }
  if (t == 0) status[b] = fail;
}

template <typename T>
void chol_inv(i64 n, i64 batch, const T *W, T *R, T *Rinv, int check, int *status) {
  const size_t shmem = sizeof(T) * (size_t)n * (size_t)(n + 1);
  HIP_CHECK(hipFuncSetAttribute((const void *)chol_inv_kernel<T>,
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)shmem));
  hipLaunchKernelGGL(chol_inv_kernel<T>, dim3((unsigned)batch), dim3(kCqThreads), shmem,
                     ctx().stream, (int)n, W, R, Rinv, check, status);
  LAUNCH_CHECK();
}

}  // namespace

int cholqr_mode() {
  const char *e = std::getenv("ISVD_CHOLQR");
  if (!e) return 0;
  if (e[0] == '0') return 1;
  if (e[0] == '1') return 2;
  return 0;
}

namespace {

// l x l steps in FP64 for either data type: W (T) -> R, R^-1 (FP64), R^-1 also as T for the
// product with the data
template <typename T>
void chol_step(i64 l, i64 batch, const T *w, double *w64, double *r, double *rinv64, T *rinv,
               int check, int *status) {
  const double *wd = w64;
  if constexpr (std::is_same_v<T, double>) {
    wd = w;
  } else {
    convert<double, float>(l, l * batch, w, l, w64, l);
  }
  chol_inv<double>(l, batch, wd, r, rinv64, check, status);
  if constexpr (!std::is_same_v<T, double>) convert<float, double>(l, l * batch, rinv64, l, rinv, l);
}

}  // namespace

template <typename T>
bool cholqr2_rowmajor(i64 m, i64 l, T *Y, i64 ldy, i64 sY, i64 batch, T *R, i64 ldr, i64 sR) {
  ISVD_CHECK(l >= 1 && l <= 128 && m >= l, "CholeskyQR2: needs 1 <= l <= 128 and m >= l");
  constexpr bool dbl = std::is_same_v<T, double>;
  const i64 ll = l * l, ldz = batch * l;
  DBuf<T> w(batch * ll), z(m * ldz), rinv_t(dbl ? 0 : batch * ll);
  DBuf<double> w64(dbl ? 0 : batch * ll), r1(batch * ll), r2(batch * ll), rinv(batch * ll);
  DBuf<int> st(2 * batch);
  T *rinv_data;
  if constexpr (dbl) rinv_data = rinv.get();
  else rinv_data = rinv_t.get();
  // pass 1: W1 = Y^T Y, W1 = R1^T R1, Z^T = R1^-T Y^T
  gemm<T>('N', 'T', l, l, m, T(1), Y, ldy, Y, ldy, T(0), w.get(), l, batch, sY, sY, ll);
  chol_step<T>(l, batch, w.get(), w64.get(), r1.get(), rinv.get(), rinv_data, 0, st.get());
  gemm<T>('T', 'N', l, m, l, T(1), rinv_data, l, Y, ldy, T(0), z.get(), ldz, batch, ll, sY, l);
  // pass 2 on Z, certified by ||Z^T Z - I||_F <= 1/2
  gemm<T>('N', 'T', l, l, m, T(1), z.get(), ldz, z.get(), ldz, T(0), w.get(), l, batch, l, l, ll);
  chol_step<T>(l, batch, w.get(), w64.get(), r2.get(), rinv.get(), rinv_data, 1,
               st.get() + batch);
  std::vector<int> h(2 * batch);
  read_back(h.data(), st.get(), sizeof(int) * 2 * batch);
  for (int v : h)
    if (v) return false;
  gemm<T>('T', 'N', l, m, l, T(1), rinv_data, l, z.get(), ldz, T(0), Y, ldy, batch, ll, l, sY);
  // R = R2 R1
  if (R) {
    if constexpr (dbl) {
      gemm<double>('N', 'N', l, l, l, 1.0, r2.get(), l, r1.get(), l, 0.0, R, ldr, batch, ll, ll,
                   sR);
    } else {
      gemm<double>('N', 'N', l, l, l, 1.0, r2.get(), l, r1.get(), l, 0.0, w64.get(), l, batch,
                   ll, ll, ll);
      for (i64 b = 0; b < batch; ++b)
        convert<float, double>(l, l, w64.get() + b * ll, l, R + b * sR, ldr);
    }
  }
  return true;
}

template bool cholqr2_rowmajor<double>(i64, i64, double *, i64, i64, i64, double *, i64, i64);
template bool cholqr2_rowmajor<float>(i64, i64, float *, i64, i64, i64, float *, i64, i64);

}  // namespace isvd
