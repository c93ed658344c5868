// Small dense SVD / symmetric eigensolver for the l x l problems of the pipeline (l <= 128):
// the Gramian eigendecompositions of orthogonalization / postprocessing (gesvd of a PSD
// matrix), the SVD of the TSQR R factor, and the KN eigendecomposition of Z.
//
// Method: one-sided (Hestenes) Jacobi with the round-robin parallel ordering.  One workgroup
// owns one matrix, held column-major in LDS (128 x 128 FP64 = 128 KiB of the 160 KiB LDS).
// Each round pairs all columns (n/2 disjoint pairs); a group of lanes per pair computes the
// three dot products with a shuffle reduction, forms the rotation that orthogonalizes the two
// columns, and applies it (and, optionally, to the right factor V kept in global memory).
// A sweep with no rotation above the relative threshold ends the iteration.
#include "la.hpp"

#include "dpp.hpp"

#include <cfloat>

namespace isvd {

namespace {

template <typename T> struct Eps;
template <> struct Eps<double> { static constexpr double v = DBL_EPSILON; };
template <> struct Eps<float> { static constexpr float v = FLT_EPSILON; };

constexpr int kMaxSweeps = 40;
// 16 waves, 16 lanes per column pair at n = 128 (8 rows per lane): a round is a chain of LDS
// reads, three reductions and the scalar rotation formula, and 16 waves hide it best (4 and 8
// waves per matrix were measured slower).
constexpr int kJacobiThreads = 1024;
constexpr int kMaxRpl = 8;  // rows per lane: n <= 128 rows over >= 16 lanes per pair

// round-robin (circle method) pairing of np players (np even), round r in [0, np-1)
__device__ __forceinline__ void rr_pair(int np, int r, int k, int &p, int &q) {
  const int m = np - 1;
  if (k == 0) {
    p = r % m;
    q = m;
  } else {
    p = (r + k) % m;
    q = (r - k + m) % m;
  }
}

// sum over aligned groups of `width` lanes (a power of two <= 64)
template <typename T>
__device__ __forceinline__ T group_sum(T v, int width) {
  return lane_group_sum(v, width);
}

// mode 0: SVD of A (general n x n).  mode 1: symmetric eigendecomposition of the matrix whose
// upper triangle is stored in A, computed as the SVD of sym(A) + c I with c = 2 ||sym(A)||_F.
template <typename T, bool WANT_V>
__global__ __launch_bounds__(kJacobiThreads) void svd_jacobi_kernel(
    int n, int mode, const T *__restrict__ A, i64 lda, i64 sA, T *__restrict__ sig, i64 lds,
    T *__restrict__ U, i64 ldu, i64 sU, T *__restrict__ Vwork, T *__restrict__ V, i64 ldv,
    i64 sV, T *__restrict__ shift_out) {
  extern __shared__ unsigned char smem_raw[];
  // n x n, leading dimension n + 1: the four column pairs a wave works on start in
  // different banks
  T *G = reinterpret_cast<T *>(smem_raw);
  __shared__ T red[kJacobiThreads / 64];
  __shared__ T nrm[128];
  __shared__ int rank_of[128];
  __shared__ int n_rot;
  __shared__ T shift_sh;

  const int b = blockIdx.x;
  const int tid = threadIdx.x;
  const T *Ab = A + (i64)b * sA;
  const int np = (n + 1) & ~1;
  const int npairs = np / 2;
  // lanes per pair: power of two, <= 64
  int tpp = kJacobiThreads / npairs;
  int t2 = 1;
  while (t2 * 2 <= tpp && t2 < 64) t2 *= 2;
  tpp = t2;
  const int grp = tid / tpp, sub = tid % tpp;

  // ---- load ----
  for (int e = tid; e < n * n; e += kJacobiThreads) {
    const int i = e % n, j = e / n;
    T v;
    if (mode == 1) {
      v = (i <= j) ? Ab[i + (i64)j * lda] : Ab[j + (i64)i * lda];  // symmetric from upper
    } else {
      v = Ab[i + (i64)j * lda];
    }
    G[i + j * (n + 1)] = v;
  }
  __syncthreads();
  if (mode == 1) {
    // c = 2 ||A||_F (+1 when A == 0), deterministic block reduction
    T part = 0;
    for (int e = tid; e < n * n; e += kJacobiThreads) {
      const T g = G[e % n + (e / n) * (n + 1)];
      part += g * g;
    }
    part = group_sum(part, 64);
    if ((tid & 63) == 0) red[tid >> 6] = part;
    __syncthreads();
    if (tid == 0) {
      T s = 0;
      for (int w = 0; w < kJacobiThreads / 64; ++w) s += red[w];
      T c = T(2) * sqrt(s);
      if (c == T(0)) c = T(1);
      shift_sh = c;
    }
    __syncthreads();
    for (int i = tid; i < n; i += kJacobiThreads) G[i + i * (n + 1)] += shift_sh;
    __syncthreads();
  }
  T *Vb = WANT_V ? (Vwork + (i64)b * n * n) : nullptr;
  if (WANT_V) {
    for (int e = tid; e < n * n; e += kJacobiThreads) {
      const int i = e % n, j = e / n;
      Vb[e] = (i == j) ? T(1) : T(0);
    }
    __syncthreads();
  }

  const T tol = Eps<T>::v * T(n);
  for (int sweep = 0; sweep < kMaxSweeps; ++sweep) {
    if (tid == 0) n_rot = 0;
    __syncthreads();
    for (int r = 0; r < np - 1; ++r) {
      if (grp < npairs) {
        int p, q;
        rr_pair(np, r, grp, p, q);
        if (p < n && q < n) {
          // the two columns stay in registers between the dot products and the rotation
          T gp[kMaxRpl], gq[kMaxRpl];
          T alpha = 0, beta = 0, gamma = 0;
#pragma unroll
          for (int u = 0; u < kMaxRpl; ++u) {
            const int i = sub + u * tpp;
            gp[u] = (i < n) ? G[i + p * (n + 1)] : T(0);
            gq[u] = (i < n) ? G[i + q * (n + 1)] : T(0);
            alpha += gp[u] * gp[u];
            beta += gq[u] * gq[u];
            gamma += gp[u] * gq[u];
          }
          alpha = group_sum(alpha, tpp);
          beta = group_sum(beta, tpp);
          gamma = group_sum(gamma, tpp);
          // |gamma| > tol sqrt(alpha beta), compared in squares
          if (gamma != T(0) && gamma * gamma > tol * tol * (alpha * beta)) {
            const T zeta = (beta - alpha) / (T(2) * gamma);
            const T t = (zeta >= T(0) ? T(1) : T(-1)) / (fabs(zeta) + sqrt(T(1) + zeta * zeta));
            const T c = T(1) / sqrt(T(1) + t * t);
            const T s = c * t;
#pragma unroll
            for (int u = 0; u < kMaxRpl; ++u) {
              const int i = sub + u * tpp;
              if (i < n) {
                G[i + p * (n + 1)] = c * gp[u] - s * gq[u];
                G[i + q * (n + 1)] = s * gp[u] + c * gq[u];
              }
            }
            if (WANT_V) {
              for (int i = sub; i < n; i += tpp) {
                const T vp = Vb[i + (i64)p * n], vq = Vb[i + (i64)q * n];
                Vb[i + (i64)p * n] = c * vp - s * vq;
                Vb[i + (i64)q * n] = s * vp + c * vq;
              }
            }
            if (sub == 0) atomicAdd(&n_rot, 1);
          }
        }
      }
      __syncthreads();
    }
    if (n_rot == 0) break;
    __syncthreads();
  }

  // ---- singular values (column norms), ranks for descending order ----
  for (int j = tid / 64; j < n; j += kJacobiThreads / 64) {
    T a = 0;
    for (int i = tid % 64; i < n; i += 64) a += G[i + j * (n + 1)] * G[i + j * (n + 1)];
    a = group_sum(a, 64);
    if ((tid & 63) == 0) nrm[j] = sqrt(a);
  }
  __syncthreads();
  for (int j = tid; j < n; j += kJacobiThreads) {
    int rk = 0;
    const T sj = nrm[j];
    for (int i = 0; i < n; ++i) {
      const T si = nrm[i];
      rk += (si > sj) || (si == sj && i < j);
    }
    rank_of[j] = rk;
  }
  __syncthreads();
  for (int j = tid; j < n; j += kJacobiThreads) sig[(i64)b * lds + rank_of[j]] = nrm[j];
  if (mode == 1 && tid == 0 && shift_out) shift_out[b] = shift_sh;
  if (U) {
    T *Ub = U + (i64)b * sU;
    for (int e = tid; e < n * n; e += kJacobiThreads) {
      const int i = e % n, j = e / n;
      const T sj = nrm[j];
      Ub[i + (i64)rank_of[j] * ldu] = (sj > T(0)) ? G[i + j * (n + 1)] / sj : T(0);
    }
  }
  if (WANT_V && V) {
    __syncthreads();
    T *Vo = V + (i64)b * sV;
    for (int e = tid; e < n * n; e += kJacobiThreads) {
      const int i = e % n, j = e / n;
      Vo[i + (i64)rank_of[j] * ldv] = Vb[e];
    }
  }
}

// eigenvalues ascending from the shifted SVD: w[j] = sig[n-1-j] - c, Z[:, j] = U[:, n-1-j]
template <typename T>
__global__ void syev_finish_kernel(int n, const T *sig, const T *shift, const T *Usrt, T *w, T *Z,
                                   i64 ldz) {
  const T c = shift[0];
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < n * n; e += gridDim.x * blockDim.x) {
    const int i = e % n, j = e / n;
    Z[i + (i64)j * ldz] = Usrt[i + (i64)(n - 1 - j) * n];
    if (i == 0) w[j] = sig[n - 1 - j] - c;
  }
}

template <typename T>
void launch_svd(i64 n, int mode, const T *A, i64 lda, i64 sA, T *s, i64 lds, T *U, i64 ldu,
                i64 sU, T *V, i64 ldv, i64 sV, i64 batch, T *shift_out) {
  ISVD_CHECK(n >= 1 && n <= 128, "small SVD supports 1 <= n <= 128 (got %lld)", n);
  const size_t shmem = sizeof(T) * (size_t)n * (size_t)(n + 1);
  DBuf<T> vwork;
  if (V) {
    vwork.alloc(batch * n * n);
    auto kern = svd_jacobi_kernel<T, true>;
    HIP_CHECK(hipFuncSetAttribute((const void *)kern,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)shmem));
    hipLaunchKernelGGL(kern, dim3((unsigned)batch), dim3(kJacobiThreads), shmem, ctx().stream,
                       (int)n, mode, A, lda, sA, s, lds, U, ldu, sU, vwork.get(), V, ldv, sV,
                       shift_out);
  } else {
    auto kern = svd_jacobi_kernel<T, false>;
    HIP_CHECK(hipFuncSetAttribute((const void *)kern,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)shmem));
    hipLaunchKernelGGL(kern, dim3((unsigned)batch), dim3(kJacobiThreads), shmem, ctx().stream,
                       (int)n, mode, A, lda, sA, s, lds, U, ldu, sU, (T *)nullptr, (T *)nullptr,
                       (i64)0, (i64)0, shift_out);
  }
  LAUNCH_CHECK();
}

}  // namespace

template <typename T>
void svd_small(i64 n, T *A, i64 lda, i64 sA, T *s, i64 lds, T *U, i64 ldu, i64 sU, T *V,
               i64 ldv, i64 sV, i64 batch) {
  if (batch <= 0) return;
  launch_svd<T>(n, 0, A, lda, sA, s, lds, U, ldu, sU, V, ldv, sV, batch, nullptr);
}

template <typename T>
void syev_small(i64 n, const T *A, i64 lda, T *w, T *Z, i64 ldz) {
  DBuf<T> sig(n), usrt(n * n), shift(1);
  launch_svd<T>(n, 1, A, lda, 0, sig.get(), n, usrt.get(), n, 0, nullptr, 0, 0, 1, shift.get());
  hipLaunchKernelGGL(syev_finish_kernel<T>, dim3(ceil_div(n * n, 256)), dim3(256), 0,
                     ctx().stream, (int)n, (const T *)sig.get(), (const T *)shift.get(),
                     (const T *)usrt.get(), w, Z, ldz);
  LAUNCH_CHECK();
}

template void svd_small<double>(i64, double *, i64, i64, double *, i64, double *, i64, i64,
                                double *, i64, i64, i64);
template void svd_small<float>(i64, float *, i64, i64, float *, i64, float *, i64, i64, float *,
                               i64, i64, i64);
template void syev_small<double>(i64, const double *, i64, double *, double *, i64);
template void syev_small<float>(i64, const float *, i64, float *, float *, i64);

}  // namespace isvd
