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

#include <type_traits>

#include "dpp.hpp"

#include <cfloat>

namespace isvd {

namespace {

template <typename T> struct Eps;
template <> struct Eps<double> { static constexpr double v = DBL_EPSILON; };
template <> struct Eps<float> { static constexpr float v = FLT_EPSILON; };

constexpr int kMaxSweeps = 40;
// 16 lanes per column pair, each holding every 16th row of the pair's two columns: 4 pairs per
// wave and, at n = 128, 16 waves (4 per SIMD), which hide each other's LDS, DPP, division and
// (with V) global-memory latencies.  A round is bound by instruction issue and latency together:
// 8 lanes per pair (8 waves) and 4 lanes (4 waves) were measured slower at n = 128.
constexpr int kTpp = 16;
constexpr int kMaxThreads = 1024;

// rows per lane for an n x n problem with tpp lanes per pair (a power of two, n <= tpp rows)
inline int jacobi_rpl(i64 n, int tpp) {
  int r = 1;
  while (r * tpp < n) r *= 2;
  return r;
}

// round-robin (circle method) pairing of np players (np even), round r in [0, np-1), table
// k in [0, np/2): player m = np - 1 stays at table 0, the others turn (r, k < m: no modulo)
__device__ __forceinline__ void rr_pair(int np, int r, int k, int &p, int &q) {
  const int m = np - 1;
  if (k == 0) {
    p = r;
    q = m;
  } else {
    p = (r + k >= m) ? r + k - m : r + k;
    q = (r >= k) ? r - k : r - k + m;
  }
}

// mode 0: SVD of A (general n x n).  mode 1: symmetric eigendecomposition of the matrix whose
// upper triangle is stored in A, computed as the SVD of sym(A) + c I with c = 2 ||sym(A)||_F.
// TPP lanes per pair, RPL rows per lane; the columns are padded with zero rows to TPP RPL
// (rotations keep them zero), so the round loop has no row bounds.  Vwork holds V with the same
// padded rows (TPP RPL x n).  blockDim.x = TPP lanes per pair, rounded up to whole waves.
template <typename T, bool WANT_V, int TPP, int RPL>
__global__ __launch_bounds__(kMaxThreads) void svd_jacobi_kernel(
    int n, int mode, const T *__restrict__ A, i64 lda, i64 sA, T *__restrict__ sig, i64 lds,
    T *__restrict__ U, i64 ldu, i64 sU, T *__restrict__ Vwork, T *__restrict__ V, i64 ldv,
    i64 sV, T *__restrict__ shift_out) {
  extern __shared__ unsigned char smem_raw[];
  // n x n, leading dimension n + 1: the columns of the pairs a wave works on start in
  // different banks
  T *G = reinterpret_cast<T *>(smem_raw);
  __shared__ T red[kMaxThreads / 64];
  __shared__ T nrm[128];
  __shared__ int rank_of[128];
  __shared__ int n_rot;
  __shared__ T shift_sh;

  const int b = blockIdx.x;
  const int tid = threadIdx.x;
  const int nt = blockDim.x, waves = nt / 64;
  const T *Ab = A + (i64)b * sA;
  const int np = (n + 1) & ~1;
  const int npairs = np / 2;
  const int grp = tid / TPP, sub = tid % TPP;
  constexpr int R = TPP * RPL;  // padded rows
  constexpr int ld = R + 1;

  // ---- load (rows n .. R-1 zero) ----
  for (int e = tid; e < R * n; e += nt) {
    const int i = e % R, j = e / R;
    T v = T(0);
    if (i < n) {
      if (mode == 1) {
        v = (i <= j) ? Ab[i + (i64)j * lda] : Ab[j + (i64)i * lda];  // symmetric from upper
      } else {
        v = Ab[i + (i64)j * lda];
      }
    }
    G[i + j * ld] = v;
  }
  __syncthreads();
  if (mode == 1) {
    // c = 2 ||A||_F (+1 when A == 0), deterministic block reduction
    T part = 0;
    for (int e = tid; e < n * n; e += nt) {
      const T g = G[e % n + (e / n) * ld];
      part += g * g;
    }
    part = lane_group_sum(part, 64);
    if ((tid & 63) == 0) red[tid >> 6] = part;
    __syncthreads();
    if (tid == 0) {
      T s = 0;
      for (int w = 0; w < waves; ++w) s += red[w];
      T c = T(2) * sqrt(s);
      if (c == T(0)) c = T(1);
      shift_sh = c;
    }
    __syncthreads();
    for (int i = tid; i < n; i += nt) G[i + i * ld] += shift_sh;
    __syncthreads();
  }
  T *Vb = WANT_V ? (Vwork + (i64)b * R * n) : nullptr;
  if (WANT_V) {
    for (int e = tid; e < R * n; e += nt) {
      const int i = e % R, j = e / R;
      Vb[e] = (i == j) ? T(1) : T(0);
    }
    __syncthreads();
  }

  const T tol = Eps<T>::v * T(n);
  for (int sweep = 0; sweep < kMaxSweeps; ++sweep) {
    if (tid == 0) n_rot = 0;
    __syncthreads();
    for (int r = 0; r < np - 1; ++r) {
      int p = n, q = n;
      if (grp < npairs) rr_pair(np, r, grp, p, q);
      if (p < n && q < n) {
        // the V columns are requested first: their latency runs under the dot products
        T vp[RPL], vq[RPL];
        const T *vcp = Vb + (i64)p * R + sub, *vcq = Vb + (i64)q * R + sub;
        if (WANT_V) {
#pragma unroll
          for (int u = 0; u < RPL; ++u) {
            vp[u] = vcp[u * TPP];
            vq[u] = vcq[u * TPP];
          }
        }
        // the two columns stay in registers between the dot products and the rotation
        T *cp = G + p * ld + sub, *cq = G + q * ld + sub;
        T gp[RPL], gq[RPL];
        T alpha = 0, beta = 0, gamma = 0;
#pragma unroll
        for (int u = 0; u < RPL; ++u) {
          gp[u] = cp[u * TPP];
          gq[u] = cq[u * TPP];
          alpha += gp[u] * gp[u];
          beta += gq[u] * gq[u];
          gamma += gp[u] * gq[u];
        }
        alpha = lane_group_sum(alpha, TPP);
        beta = lane_group_sum(beta, TPP);
        gamma = lane_group_sum(gamma, TPP);
        // |gamma| > tol sqrt(alpha beta), compared in squares
        if (gamma != T(0) && gamma * gamma > tol * tol * (alpha * beta)) {
          // t = sign(zeta) / (|zeta| + sqrt(1 + zeta^2)), zeta = (beta - alpha) / (2 gamma); in
          // FP64 multiplied through by |2 gamma| (one division less on the round's dependent
          // chain; the squares stay far from overflow at FP64 range), sign(zeta) = +1 at zeta = 0
          T t;
          if constexpr (std::is_same_v<T, double>) {
            const T d = beta - alpha, g2 = T(2) * gamma;
            const bool pos = d == T(0) || ((d > T(0)) == (g2 > T(0)));
            t = (pos ? fabs(g2) : -fabs(g2)) / (fabs(d) + sqrt(d * d + g2 * g2));
          } else {
            const T zeta = (beta - alpha) / (T(2) * gamma);
            t = (zeta >= T(0) ? T(1) : T(-1)) / (fabs(zeta) + sqrt(T(1) + zeta * zeta));
          }
          const T c = T(1) / sqrt(T(1) + t * t);
          const T s = c * t;
#pragma unroll
          for (int u = 0; u < RPL; ++u) {
            cp[u * TPP] = c * gp[u] - s * gq[u];
            cq[u * TPP] = s * gp[u] + c * gq[u];
          }
          if (WANT_V) {
            T *wp = Vb + (i64)p * R + sub, *wq = Vb + (i64)q * R + sub;
#pragma unroll
            for (int u = 0; u < RPL; ++u) {
              wp[u * TPP] = c * vp[u] - s * vq[u];
              wq[u * TPP] = s * vp[u] + c * vq[u];
            }
          }
          if (sub == 0) atomicAdd(&n_rot, 1);
        }
      }
      __syncthreads();
    }
    if (n_rot == 0) break;
    __syncthreads();
  }

  // ---- singular values (column norms), ranks for descending order ----
  for (int j = tid / 64; j < n; j += waves) {
    T a = 0;
    for (int i = tid % 64; i < n; i += 64) a += G[i + j * ld] * G[i + j * ld];
    a = lane_group_sum(a, 64);
    if ((tid & 63) == 0) nrm[j] = sqrt(a);
  }
  __syncthreads();
  for (int j = tid; j < n; j += nt) {
    int rk = 0;
    const T sj = nrm[j];
    for (int i = 0; i < n; ++i) {
      const T si = nrm[i];
      rk += (si > sj) || (si == sj && i < j);
    }
    rank_of[j] = rk;
  }
  __syncthreads();
  for (int j = tid; j < n; j += nt) sig[(i64)b * lds + rank_of[j]] = nrm[j];
  if (mode == 1 && tid == 0 && shift_out) shift_out[b] = shift_sh;
  if (U) {
    T *Ub = U + (i64)b * sU;
    for (int e = tid; e < n * n; e += nt) {
      const int i = e % n, j = e / n;
      const T sj = nrm[j];
      Ub[i + (i64)rank_of[j] * ldu] = (sj > T(0)) ? G[i + j * ld] / sj : T(0);
    }
  }
  if (WANT_V && V) {
    __syncthreads();
    T *Vo = V + (i64)b * sV;
    for (int e = tid; e < n * n; e += nt) {
      const int i = e % n, j = e / n;
      Vo[i + (i64)rank_of[j] * ldv] = Vb[i + (i64)j * R];
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

template <typename T, bool WANT_V, int TPP, int RPL>
void launch_rpl(int nt, i64 n, int mode, const T *A, i64 lda, i64 sA, T *s, i64 lds, T *U,
                i64 ldu, i64 sU, T *vwork, T *V, i64 ldv, i64 sV, i64 batch, T *shift_out) {
  const size_t shmem = sizeof(T) * (size_t)(TPP * RPL + 1) * (size_t)n;
  auto kern = svd_jacobi_kernel<T, WANT_V, TPP, RPL>;
  HIP_CHECK(hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)shmem));
  hipLaunchKernelGGL(kern, dim3((unsigned)batch), dim3(nt), shmem, ctx().stream, (int)n, mode, A,
                     lda, sA, s, lds, U, ldu, sU, vwork, V, ldv, sV, shift_out);
  LAUNCH_CHECK();
}

template <typename T, bool WANT_V>
void launch_v(i64 n, int mode, const T *A, i64 lda, i64 sA, T *s, i64 lds, T *U, i64 ldu, i64 sU,
              T *vwork, T *V, i64 ldv, i64 sV, i64 batch, T *shift_out) {
  constexpr int TPP = kTpp;
  const i64 npairs = (n + 1) / 2;
  const int nt = (int)((npairs * TPP + 63) / 64 * 64);
#define ISVD_JAC(R) \
  launch_rpl<T, WANT_V, TPP, R>(nt, n, mode, A, lda, sA, s, lds, U, ldu, sU, vwork, V, ldv, sV,   \
                                batch, shift_out)
  switch (jacobi_rpl(n, TPP)) {
    case 1: ISVD_JAC(1); break;
    case 2: ISVD_JAC(2); break;
    case 4: ISVD_JAC(4); break;
    case 8: ISVD_JAC(8); break;
    default: ISVD_JAC(128 / TPP); break;
  }
#undef ISVD_JAC
}

template <typename T>
void launch_svd(i64 n, int mode, const T *A, i64 lda, i64 sA, T *s, i64 lds, T *U, i64 ldu,
                i64 sU, T *V, i64 ldv, i64 sV, i64 batch, T *shift_out) {
  ISVD_CHECK(n >= 1 && n <= 128, "small SVD supports 1 <= n <= 128 (got %lld)", n);
  if (V) {
    // V with its rows padded like the columns of the matrix (at most 128 rows)
    DBuf<T> vwork(batch * (i64)kTpp * jacobi_rpl(n, kTpp) * n);
    launch_v<T, true>(n, mode, A, lda, sA, s, lds, U, ldu, sU, vwork.get(), V, ldv, sV, batch,
                      shift_out);
  } else {
    launch_v<T, false>(n, mode, A, lda, sA, s, lds, U, ldu, sU, (T *)nullptr, (T *)nullptr,
                       (i64)0, (i64)0, batch, shift_out);
  }
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
