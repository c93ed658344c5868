// The four iSVD stages on one GPU.  Each function cites the reference code it follows.
#include "stages.hpp"

#include <cmath>
#include <cstdlib>
#include <cstring>
#include <type_traits>

#include "la.hpp"
#include "mx.hpp"

namespace isvd {

namespace {

// ---------------------------------------------------------------------------------------------
// Philox-4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11) counter-based generator + Box-Muller.
// Element e of Omega^T (column-major Nl x n) uses counter (e, 0, 0, 0) and key (seed_lo,
// seed_hi); the first two 32-bit words make u1, the last two u2 (53-bit uniforms in (0, 1]).
// ---------------------------------------------------------------------------------------------
__device__ __forceinline__ void philox_round(unsigned (&c)[4], unsigned (&k)[2]) {
  const unsigned long long p0 = 0xD2511F53ull * c[0];
  const unsigned long long p1 = 0xCD9E8D57ull * c[2];
  const unsigned hi0 = (unsigned)(p0 >> 32), lo0 = (unsigned)p0;
  const unsigned hi1 = (unsigned)(p1 >> 32), lo1 = (unsigned)p1;
  const unsigned n0 = hi1 ^ c[1] ^ k[0];
  const unsigned n2 = hi0 ^ c[3] ^ k[1];
  c[0] = n0; c[1] = lo1; c[2] = n2; c[3] = lo0;
  k[0] += 0x9E3779B9u;
  k[1] += 0xBB67AE85u;
}

template <typename T>
__global__ void gaussian_kernel(i64 rows, i64 cols, T *out, i64 ld, unsigned long long seed) {
  const i64 total = rows * cols;
  for (i64 e = blockIdx.x * (i64)blockDim.x + threadIdx.x; e < total;
       e += (i64)gridDim.x * blockDim.x) {
    unsigned c[4] = {(unsigned)e, (unsigned)(e >> 32), 0u, 0u};
    unsigned k[2] = {(unsigned)seed, (unsigned)(seed >> 32)};
#pragma unroll
    for (int r = 0; r < 10; ++r) philox_round(c, k);
    const unsigned long long a = ((unsigned long long)c[0] << 32) | c[1];
    const unsigned long long b = ((unsigned long long)c[2] << 32) | c[3];
    const double u1 = ((a >> 11) + 1) * (1.0 / 9007199254740992.0);
    const double u2 = ((b >> 11) + 1) * (1.0 / 9007199254740992.0);
    const double z = sqrt(-2.0 * log(u1)) * cos(6.283185307179586476925 * u2);
    const i64 i = e % rows, j = e / rows;
    out[i + j * ld] = (T)z;
  }
}

// W[:, j] *= 1 / sqrt(s[j])  for j < ncols (batched; vSqrt + Dism of the Gramian stage)
template <typename T>
__global__ void sqrt_and_colscale_kernel(int l, int ncols, const T *s, i64 lds, T *W, i64 ldw,
                                         i64 sW) {
  const int b = blockIdx.y;
  const T *sb = s + (i64)b * lds;
  T *Wb = W + (i64)b * sW;
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < l * ncols;
       e += gridDim.x * blockDim.x) {
    const int i = e % l, j = e / l;
    Wb[i + (i64)j * ldw] *= T(1) / sqrt(sb[j]);
  }
}

template <typename T>
__global__ void sqrt_inplace_kernel(int n, T *s) {
  const int b = blockIdx.y;
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < n; e += gridDim.x * blockDim.x)
    s[(i64)b * n + e] = sqrt(s[(i64)b * n + e]);
}

// KN (@x@_integrate_kolmogorov_nagumo.c:196-212, :230-236): from eig(Z) = (w, V):
//   s = sqrt(1/2 + sqrt(1/4 - w)); ss = sqrt(s); ZS = V diag(ss); ZI = V diag(1/ss);
//   err = ||s - 1||_2 (one workgroup, fixed reduction order)
template <typename T>
__global__ __launch_bounds__(256) void kn_scalars_kernel(int l, const T *w, const T *V, T *ZS,
                                                         T *ZI, T *err) {
  __shared__ T ss[128];
  __shared__ T red[4];
  T part = 0;
  for (int j = threadIdx.x; j < l; j += 256) {
    const T s = sqrt(T(0.5) + sqrt(T(0.25) - w[j]));
    ss[j] = sqrt(s);
    part += (s - T(1)) * (s - T(1));
  }
  for (int off = 32; off > 0; off >>= 1) part += __shfl_xor(part, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = part;
  __syncthreads();
  if (threadIdx.x == 0) err[0] = sqrt(red[0] + red[1] + red[2] + red[3]);
  for (int e = threadIdx.x; e < l * l; e += 256) {
    const int j = e / l;
    ZS[e] = V[e] * ss[j];
    ZI[e] = V[e] / ss[j];
  }
}

// A := upper(A) mirrored to the lower triangle (syrk 'U' result used through symm 'U')
template <typename T>
__global__ void sym_from_upper_kernel(int n, T *A, i64 lda) {
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < n * n; e += gridDim.x * blockDim.x) {
    const int i = e % n, j = e / n;
    if (i > j) A[i + (i64)j * lda] = A[j + (i64)i * lda];
  }
}

// WY step operators in closed form.  With C = [Dc/2 - I/tau, I/2; -Dgc/2, -Dc/2 - I/tau]
// (@x@_integrate_wen_yin.c:249-259), eliminating with the C_12 = I/2 block gives
//   Fgc = [inv(C)_12 Dc - inv(C)_11] = -(2/tau) S^-1,
//   Fc  = [inv(C)_22 Dc - inv(C)_21] + I = (2/tau) (Dc - (2/tau) I) S^-1 - I,
// with S = -(Dgc - Dc^2)/2 - (2/tau^2) I, which is singular exactly when C is.  S^-1 is an
// l x l Gauss-Jordan inverse (the reference inverts the 2l x 2l C with getrf + getri).
// out := a * sym(A) + b * I
template <typename T>
__global__ void sym_scale_kernel(int n, const T *A, T *out, T a, T b) {
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < n * n; e += gridDim.x * blockDim.x) {
    const int i = e % n, j = e / n;
    out[e] = a * T(0.5) * (A[i + j * n] + A[j + i * n]) + ((i == j) ? b : T(0));
  }
}

// F := F + a * S + b * I   (l x l)
template <typename T>
__global__ void axpy_diag_kernel(int l, T *F, const T *S, T a, T b) {
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < l * l; e += gridDim.x * blockDim.x) {
    const int i = e % l, j = e / l;
    F[e] += a * S[e] + ((i == j) ? b : T(0));
  }
}

// sum |diag(A)| - ||B||_F^2  (mu of @x@_integrate_wen_yin.c:232, :307); one workgroup
template <typename T>
__global__ __launch_bounds__(256) void wy_mu_kernel(int l, const T *dgc, const T *dc, T *mu) {
  __shared__ T red[4];
  T part = 0;
  for (int e = threadIdx.x; e < l * l; e += 256) {
    const int i = e % l, j = e / l;
    if (i == j) part += fabs(dgc[e]);
    part -= dc[e] * dc[e];
  }
  for (int off = 32; off > 0; off >>= 1) part += __shfl_xor(part, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = part;
  __syncthreads();
  if (threadIdx.x == 0) mu[0] = red[0] + red[1] + red[2] + red[3];
}

// c = a - b (elementwise, n entries)
template <typename T>
__global__ void sub_kernel(i64 n, const T *a, const T *b, T *c) {
  for (i64 e = blockIdx.x * (i64)blockDim.x + threadIdx.x; e < n;
       e += (i64)gridDim.x * blockDim.x)
    c[e] = a[e] - b[e];
}

template <typename T>
__global__ void diag_add_kernel(int n, T *A, i64 lda, T v) {
  for (int i = threadIdx.x; i < n; i += blockDim.x) A[i + (i64)i * lda] += v;
}

// Dism side 'R' (src/lib/libisvd/la/blas/dism.c:26-31): W[:, j] *= 1 / s[j]
template <typename T>
__global__ void colscale_inv_kernel(int rows, int ncols, const T *s, T *W, i64 ldw) {
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < rows * ncols;
       e += gridDim.x * blockDim.x) {
    const int i = e % rows, j = e / rows;
    W[i + (i64)j * ldw] *= T(1) / s[j];
  }
}

template <typename T>
__global__ void transpose_small_kernel(int n, const T *A, i64 lda, T *B, i64 ldb) {
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < n * n; e += gridDim.x * blockDim.x) {
    const int i = e % n, j = e / n;
    B[j + (i64)i * ldb] = A[i + (i64)j * lda];
  }
}

template <typename T>
T read_scalar(const T *dptr) {
  T h;
  read_back(&h, dptr, sizeof(T));
  return h;
}

// ISVD_TRACE=1 prints one line per integration iteration (and line-search trip) to stderr.
bool trace_on() {
  static const bool on = [] {
    const char *e = std::getenv("ISVD_TRACE");
    return e && e[0] == '1';
  }();
  return on;
}

int grid_for(i64 n) {
  i64 b = (n + 255) / 256;
  if (b > 8192) b = 8192;
  if (b < 1) b = 1;
  return (int)b;
}

}  // namespace

// ---------------------------------------------------------------------------------------------
template <typename T>
void gaussian_omega(i64 rows, i64 cols, T *omegat, i64 ld, unsigned long long seed) {
  hipLaunchKernelGGL(gaussian_kernel<T>, dim3(grid_for(rows * cols)), dim3(256), 0,
                     ctx().stream, rows, cols, omegat, ld, seed);
  LAUNCH_CHECK();
}

// Sketch: Y^T := Omega^T * op(A)^T   (sketchBlockRow, src/lib/libisvd/core/stage/@x@_sketch.h:
// 136-139; sketchBlockCol :74-82 is the same product at P = 1).
template <typename T>
void stage_sketch(const Dims &d, char ordera, const T *a, i64 lda, const T *omegat, i64 ldo,
                  T *yst, i64 ldyst) {
  const char transa = (ordera == 'C') ? 'T' : 'N';
  gemm<T>('N', transa, d.Nl, d.m, d.n, T(1), omegat, ldo, a, lda, T(0), yst, ldyst);
}

// A host-resident A staged by column blocks of `cols` columns (the reference's GPU stages,
// src/lib/libisvd/gpu/stage/@x@_sketch_gpu.h:104-125, @x@_postprocess_gpu.h:121-140).  Block
// c0 .. c0 + nc of A is an m x nc column-major block for ordera 'C', or the nc x m block of the
// row-major array (rows c0 .. c0 + nc of A^T) for 'R'; it lands in `buf` with leading dimension
// m ('C') or nc ('R').
template <typename T>
static i64 stage_block(char ordera, i64 m, const T *a_host, i64 lda, i64 c0, i64 nc, T *buf) {
  if (ordera == 'C') {
    HIP_CHECK(hipMemcpy2DAsync(buf, sizeof(T) * m, a_host + c0 * lda, sizeof(T) * lda,
                               sizeof(T) * m, nc, hipMemcpyHostToDevice, ctx().stream));
    return m;
  }
  HIP_CHECK(hipMemcpy2DAsync(buf, sizeof(T) * nc, a_host + c0, sizeof(T) * lda, sizeof(T) * nc,
                             m, hipMemcpyHostToDevice, ctx().stream));
  return nc;
}

// Y^T := Omega^T op(A)^T accumulated over column blocks of a host-resident A
// (@x@_sketch_gpu.h:116-125: Y^T += Omega_b^T A_b^T)
template <typename T>
void stage_sketch_host_blocks(const Dims &d, char ordera, const T *a_host, i64 lda,
                              const T *omegat, i64 ldo, T *yst, i64 ldyst, i64 cols) {
  DBuf<T> buf(d.m * cols);
  for (i64 c0 = 0; c0 < d.n; c0 += cols) {
    const i64 nc = (d.n - c0) < cols ? (d.n - c0) : cols;
    const i64 ldb = stage_block<T>(ordera, d.m, a_host, lda, c0, nc, buf.get());
    gemm<T>('N', ordera == 'C' ? 'T' : 'N', d.Nl, d.m, nc, T(1), omegat + c0 * ldo, ldo,
            buf.get(), ldb, c0 ? T(1) : T(0), yst, ldyst);
  }
}

// Z^T := Qbar^T op(A) by column blocks of a host-resident A (@x@_postprocess_gpu.h:121-140)
template <typename T>
void project_host_blocks(const Dims &d, char ordera, const T *a_host, i64 lda, const T *qt,
                         i64 ldqt, T *zt, i64 cols) {
  DBuf<T> buf(d.m * cols);
  for (i64 c0 = 0; c0 < d.n; c0 += cols) {
    const i64 nc = (d.n - c0) < cols ? (d.n - c0) : cols;
    const i64 ldb = stage_block<T>(ordera, d.m, a_host, lda, c0, nc, buf.get());
    gemm<T>('N', ordera == 'C' ? 'N' : 'T', d.l, nc, d.m, T(1), qt, ldqt, buf.get(), ldb, T(0),
            zt + c0 * d.l, d.l);
  }
}

// Gramian orthogonalization (src/lib/libisvd/core/stage/@x@_orthogonalize_gramian.c:81-98):
// Q_i = Y_i U_i S_i^-1 from W_i = Y_i^T Y_i = U_i S_i^2 U_i^T.  With cholqr (KN follows, see
// stage_orth_tsqr), the Cholesky form of the same Gramian orthogonalization, done twice
// (CholeskyQR2: Q_i = Y_i R_i^-1, the same span, another orthonormal basis of it), where the
// device certifies it.
template <typename T>
void stage_orth_gramian(const Dims &d, T *yst, i64 ldyst, bool cholqr) {
  const int mode = cholqr_mode();
  if (((cholqr && mode == 0) || mode == 2) &&
      cholqr2_rowmajor<T>(d.m, d.l, yst, ldyst, d.l, d.N, nullptr, 0, 0))
    return;
  const i64 l = d.l, N = d.N, m = d.m;
  DBuf<T> w(l * d.Nl), u(l * d.Nl), s(l * N), ycopy(ldyst * m);
  // W_i := Y_i^T Y_i  (:81-84)
  gemm<T>('N', 'T', l, l, m, T(1), yst, ldyst, yst, ldyst, T(0), w.get(), l, N, l, l, l * l);
  // gesvd('O','N') of every W_i  (:87-90)
  svd_small<T>(l, w.get(), l, l * l, s.get(), l, u.get(), l, l * l, nullptr, 0, 0, N);
  // s := sqrt(s); U_i := U_i / s_i  (:91, :94)
  hipLaunchKernelGGL(sqrt_and_colscale_kernel<T>, dim3(grid_for(l * l), (unsigned)N), dim3(256),
                     0, ctx().stream, (int)l, (int)l, (const T *)s.get(), l, u.get(), l, l * l);
  LAUNCH_CHECK();
  // Q_i^T := U_i^T Y_i^T  (:95-98)
  HIP_CHECK(hipMemcpyAsync(ycopy.get(), yst, sizeof(T) * ldyst * m, hipMemcpyDeviceToDevice,
                           ctx().stream));
  gemm<T>('T', 'N', l, m, l, T(1), u.get(), l, ycopy.get(), ldyst, T(0), yst, ldyst, N, l * l, l,
          l);
}

// Tall-skinny QR orthogonalization (@x@_orthogonalize_tall_skinny_qr.c:127-200): every sketch
// Y_i (row-major m x l inside yst) is replaced by the orthonormal factor of its QR.  With
// cholqr, by CholeskyQR2 on the matrix cores when the device certifies it; else (and for
// Wen-Yin, whose stopping point within its tolerance moves with rounding-level changes of the
// bases) by the Householder tree.
template <typename T>
void stage_orth_tsqr(const Dims &d, T *yst, i64 ldyst, bool cholqr) {
  const int mode = cholqr_mode();
  if (((cholqr && mode == 0) || mode == 2) &&
      cholqr2_rowmajor<T>(d.m, d.l, yst, ldyst, d.l, d.N, nullptr, 0, 0))
    return;
  tsqr_rowmajor<T>(d.m, d.l, yst, ldyst, d.l, d.N, nullptr, 0, 0);
}

// Kolmogorov-Nagumo integration (@x@_integrate_kolmogorov_nagumo.c:155-276), iterated in the
// coordinates of the sketch basis: every iterate satisfies Q_c = Qs X_c, and
//   Bg = Qs^T (Qs B / N) = M B / N with M = Qs^T Qs (Nl x Nl, formed once),
//   X_+ = X_c F + (B_c / N) C^-1   (the image of Q_+ = Q_c F + G_c C^-1),
// so each iteration touches only Nl x l and l x l matrices; Qbar = Qs X is formed at the end.
// W3 := [F; C^-1 / N], [F; C^-1], [F; C^-1]  (three 2l x l blocks, ld 2l)
template <typename T>
__global__ void kn_assemble_kernel(int l, const T *F, const T *Ci, T ninv, T *W) {
  const i64 blk = (i64)2 * l * l;
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < l * l; e += gridDim.x * blockDim.x) {
    const int i = e % l, j = e / l;
    const T f = F[e], ci = Ci[e];
    T *w = W + (i64)j * 2 * l;
    w[i] = f;
    w[l + i] = ninv * ci;
    w[blk + i] = f;
    w[blk + l + i] = ci;
    w[2 * blk + i] = f;
    w[2 * blk + l + i] = ci;
  }
}

// Runs the enclosed launches on another stream instead of ctx().stream.
struct OnStream {
  hipStream_t saved;
  explicit OnStream(hipStream_t s) : saved(ctx().stream) { ctx().stream = s; }
  ~OnStream() { ctx().stream = saved; }
};

// At scope exit (also when an error unwinds it) the first stream waits for everything queued on
// aux, so that buffers freed in stream order on the first are no longer in use.
struct JoinAux {
  ~JoinAux() {
    if (hipEventRecord(ctx().join, ctx().aux) == hipSuccess)
      (void)hipStreamWaitEvent(ctx().stream, ctx().join, 0);
  }
};

// The way kn_chain takes C = f(Z): ISVD_KN_EIG=1 the eigendecomposition of Z in every KN
// iteration, ISVD_KN_NS=1 Newton-Schulz whenever it converges (the paths otherwise taken only
// when rho is too large for the series, or for Newton-Schulz); default automatic.
int kn_mode() {
  const char *e = std::getenv("ISVD_KN_EIG");
  if (e && e[0] == '1') return 1;
  const char *n = std::getenv("ISVD_KN_NS");
  if (n && n[0] == '1') return 2;
  return 0;
}

// The KN iteration in sketch coordinates.  A state [X | B | Bg | G] (Nl x 4l) holds X_c,
// B_c = Qs^T Q_c, Bg_c = M B_c / N and G_c = M Bg_c / N; [D | Dg] = B_c^T [B_c | Bg_c] / N.
// Iteration it, from state s to state s+1 (three states in a ring):
//   second stream: G_c := M Bg_c / N (waits for Bg_c only);
//   kn_chain: [D | Dg] from the Gram partials, Z = Dg - D^T D, rho, C and C^-1, F = C - D C^-1,
//     W3 = [F; C^-1/N], [F; C^-1], [F; C^-1];
//   [X_+ | B_+] := [X_c B_c | B_c Bg_c] W3 (beside G_c), then, once G_c is there,
//   Bg_+ := [Bg_c G_c] W3 (B_+ and Bg_+ are M X_+ and M B_+ / N because M X_c = B_c and
//   M Bg_c / N = G_c), and the Gram partials of the new state;
//   err = ||C - I||_F copied to a pinned slot.
// The host queues iteration it + 1 before it waits for err of iteration it, so the GPU does not
// idle between iterations; the work queued past the last iteration is discarded.  When rho is
// out of reach of the chain (status 1), iteration it is done again from its state (intact in the
// ring) with C and C^-1 from the eigendecomposition of Z.
template <typename T>
void integrate_kn_coords(const Dims &d, const KnArgs &args, const T *M, T *x_out, i64 *iter_out,
                         double *err_out, const SplitSym *ms) {
  const i64 l = d.l, Nl = d.Nl, ll = l * l;
  ISVD_CHECK(!ms || (std::is_same_v<T, float> && ms->n == Nl),
             "KN: the BF16-split M is for the FP32 loop with n = Nl");
  // out := alpha M in (Nl x l, ld Nl)
  auto m_apply = [&](const T *in, T alpha, T *out) {
    if constexpr (std::is_same_v<T, float>) {
      if (ms) {
        split_sym_apply(*ms, l, in, Nl, alpha, out, Nl);
        return;
      }
    }
    // (in^T M)^T: M is exactly symmetric
    gemm_ct<T>('T', 'N', l, Nl, Nl, alpha, in, Nl, M, Nl, T(0), out, Nl);
  };
  const T ninv = T(1) / T(d.N);
  const int mode = kn_mode();
  // [D | Dg] as K-chunk partial sums over about 256 workgroups of 64 x 64 tiles
  const i64 dtiles = ((l + 63) / 64) * ((2 * l + 63) / 64);
  int want = (int)(256 / dtiles);
  if (want < 1) want = 1;
  if (want > Nl / 64) want = (int)(Nl / 64 > 0 ? Nl / 64 : 1);
  const i64 sz = Nl * 4 * l;
  DBuf<T> sbuf(3 * sz), ddg(2 * ll), w3(6 * ll), scal(4), parts((i64)(want + 1) * 2 * ll);
  DBuf<T> cwork(kn_chain_work_size(l));
  DBuf<unsigned long long> bar(kKnChainBarWords);
  KnChainWork<T> chain;
  chain.buf = cwork.get();
  chain.bar = bar.get();
  kn_chain_init<T>(chain, (int)l);
  // eigendecomposition path
  DBuf<T> w(l), zv(ll), zs(ll), zi(ll), c(ll), cinv(ll);
  const JoinAux join_aux;  // destroyed before the buffers above
  auto X = [&](i64 q) { return sbuf.get() + (q % 3) * sz; };
  auto B = [&](i64 q) { return X(q) + Nl * l; };
  auto Bg = [&](i64 q) { return X(q) + 2 * Nl * l; };
  auto G = [&](i64 q) { return X(q) + 3 * Nl * l; };
  // D := B^T B / N (:181), Dg := B^T Bg / N (:184-186): the K-chunk partials, summed by kn_chain
  int nparts = 0;
  auto gram = [&](i64 q) {
    nparts = gemm_partials<T>('T', 'N', l, 2 * l, Nl, B(q), Nl, B(q), Nl, parts.get(), want);
  };
  // G_q := M Bg_q / N on the second stream, after the fork event (recorded once Bg_q is written).
  // Its workgroups fill a CU's registers each: it leaves the chain's CUs free, or the chain,
  // launched beside it, would wait for G's workgroups to finish before its own could start.
  auto g_product = [&](i64 q) {
    HIP_CHECK(hipStreamWaitEvent(ctx().aux, pinned_slot_event((int)(q & 1)), 0));
    {
      OnStream on(ctx().aux);
      GridCap cap(ctx().cus - kn_chain_blocks(l));
      m_apply(Bg(q), ninv, G(q));
    }
    HIP_CHECK(hipEventRecord(ctx().join, ctx().aux));
  };
  // after the join, one batched launch: X_+ := X_c F + (B_c / N) C^-1 (image of :221-222),
  // B_+ := B_c F + Bg_c C^-1 (:225-226), Bg_+ := Bg_c F + G_c C^-1 ([X | B], [B | Bg] and
  // [Bg | G] are consecutive column blocks of the state); then the event of pinned slot
  // (q + 1) % 2, which is both the fork for the next G (state q + 1 is written) and the host's
  // signal that the chain of iteration q + 1 has stored its scalars in that slot (one marker on
  // the library stream per iteration); the next Gram.  (Updating X and B before the join would
  // run them beside G on the few CUs G leaves to the chain.)
  auto updates = [&](i64 q) {
    HIP_CHECK(hipStreamWaitEvent(ctx().stream, ctx().join, 0));
    gemm_shortk<T>('N', 'N', Nl, l, 2 * l, T(1), X(q), Nl, w3.get(), 2 * l, T(0), X(q + 1), Nl,
                   3, Nl * l, 2 * ll, Nl * l);
    pinned_slot_record((int)((q + 1) & 1));
    gram(q + 1);
  };
  // iteration it from state it - 1 (queued; the chain stores err, rho and its status straight
  // into pinned slot it % 2, whose event the updates behind it record)
  T *slot[2] = {static_cast<T *>(pinned_slot(0)), static_cast<T *>(pinned_slot(1))};
  auto enqueue = [&](i64 it, int m) {
    const i64 q = it - 1;
    g_product(q);
    // [D | Dg]; C and C^-1 (:193-212): C = (I/2 + (I/4 - Z)^{1/2})^{1/2}; F (:218); W3
    kn_chain<T>((int)l, ninv, parts.get(), nparts, ddg.get(), chain, w3.get(), slot[it & 1], m);
    updates(q);
  };
  // X_c := [I; 0] (Qc := Q_0, :156);  B_c := Qs^T Q_0 = M[:, 0:l]  (:159-160);
  // Bg_c := M B_c / N (:174-178)
  fill<T>(Nl, l, T(0), X(0), Nl, T(1));
  if constexpr (std::is_same_v<T, float>) {
    if (ms) split_sym_columns(*ms, l, B(0), Nl);
    else omatcopy<T>('N', Nl, l, T(1), M, Nl, B(0), Nl);
  } else {
    omatcopy<T>('N', Nl, l, T(1), M, Nl, B(0), Nl);
  }
  m_apply(B(0), ninv, Bg(0));
  pinned_slot_record(0);  // the fork for G_0
  gram(0);
  enqueue(1, mode);
  i64 it;
  double error = -INFINITY;
  T h[4];
  for (it = 1;; ++it) {
    const bool more = it < args.maxit;
    if (more) enqueue(it + 1, mode);
    read_back_wait((int)(it & 1), h, 4 * sizeof(T));
    ISVD_CHECK(h[3] != T(2), "KN: a grid barrier of the l x l chain timed out");
    const bool use_eig = h[3] == T(1);
    if (use_eig) {
      // iteration it again from its state, with C = V diag(s) V^T,
      // s = sqrt(1/2 + sqrt(1/4 - eig(Z))) and err = ||s - 1||_2 (G_c of the state is valid);
      // everything queued so far is finished first
      HIP_CHECK(hipStreamSynchronize(ctx().aux));
      const i64 q = it - 1;
      gram(q);
      kn_chain<T>((int)l, ninv, parts.get(), nparts, ddg.get(), chain, w3.get(), scal.get(), 1);
      syev_small<T>(l, chain.Z(), l, w.get(), zv.get(), l);
      hipLaunchKernelGGL(kn_scalars_kernel<T>, dim3(1), dim3(256), 0, ctx().stream, (int)l,
                         (const T *)w.get(), (const T *)zv.get(), zs.get(), zi.get(), scal.get());
      LAUNCH_CHECK();
      small_gemm<T>('N', 'T', l, l, l, T(1), zs.get(), l, zs.get(), l, T(0), c.get(), l);
      small_gemm<T>('N', 'T', l, l, l, T(1), zi.get(), l, zi.get(), l, T(0), cinv.get(), l);
      hipLaunchKernelGGL(sym_from_upper_kernel<T>, dim3(grid_for(ll)), dim3(256), 0,
                         ctx().stream, (int)l, cinv.get(), l);
      LAUNCH_CHECK();
      // F [in C] := C - D inv(C)  (:218)
      small_gemm<T>('N', 'N', l, l, l, T(-1), ddg.get(), l, cinv.get(), l, T(1), c.get(), l);
      hipLaunchKernelGGL(kn_assemble_kernel<T>, dim3(grid_for(ll)), dim3(256), 0, ctx().stream,
                         (int)l, (const T *)c.get(), (const T *)cinv.get(), ninv, w3.get());
      LAUNCH_CHECK();
      HIP_CHECK(hipEventRecord(ctx().join, ctx().stream));  // G_c is already in the state
      updates(q);
      read_back(h, scal.get(), sizeof(T));
      if (more) enqueue(it + 1, mode);  // the queued one ran from a discarded state
    }
    error = (double)h[0];
    if (trace_on())
      std::fprintf(stderr, "isvd KN iter %lld err %.17g rho %.3g (%s)\n", it, error,
                   (double)h[2], use_eig ? "eig" : (h[1] == T(0) ? "series" : "newton-schulz"));
    // convergence (:230-241)
    if (error <= args.tol) break;
    if (it == args.maxit) {
      ++it;
      break;
    }
  }
  // the result is the state after the last iteration; the queued iteration past it writes the
  // next state of the ring only
  const i64 last = (error <= args.tol) ? it : it - 1;
  omatcopy<T>('N', Nl, l, T(1), X(last), Nl, x_out, Nl);
  *iter_out = it;
  *err_out = error;
}

template <typename T>
void stage_integrate_kn(const Dims &d, const KnArgs &args, const T *yst, i64 ldyst, T *qt,
                        i64 ldqt, i64 *iter_out, double *err_out) {
  const i64 l = d.l, Nl = d.Nl, m = d.m;
  DBuf<T> M(Nl * Nl), x(Nl * l);
  // M := Qs^T Qs
  syrk_full<T>('N', Nl, m, T(1), yst, ldyst, M.get(), Nl);
  integrate_kn_coords<T>(d, args, M.get(), x.get(), iter_out, err_out);
  // Qbar^T := X_+^T Qs^T  (:250-253)
  gemm<T>('T', 'N', l, m, Nl, T(1), x.get(), Nl, yst, ldyst, T(0), qt, ldqt);
}

// Wen-Yin integration (@x@_integrate_wen_yin.c:201-384), same coordinates as KN:
// Q_c = Qs xq, G_c = Qs gq (gq = B_c / N), X_c = Qs xx, <Qs a, Qs b> = <a, M b>; the
// 2l x 2l inverse of the line search is replaced by its closed form (sym_scale_kernel).
template <typename T>
void integrate_wy_coords(const Dims &d, const WyArgs &args, const T *M, T *x_out, i64 *iter_out,
                         double *mu_out) {
  const i64 l = d.l, Nl = d.Nl;
  const T ninv = T(1) / T(d.N);
  const T half_ninv = T(0.5) / T(d.N);
  DBuf<T> xq(Nl * l), xqp(Nl * l), xx(Nl * l), xxp(Nl * l), bc(Nl * l), bp(Nl * l),
      bgc(Nl * l), gq(Nl * l), d1(Nl * l), d2(Nl * l), md1(Nl * l);
  // scal: [0] mu, [1] phi, [2] phit, [3] the inverse's pivot status (an int in the slot),
  // [4..5] t; the values the host needs at one point are read back together
  DBuf<T> dc(l * l), dgc(l * l), zm(l * l), si(l * l), fcb(l * l), fgcb(l * l), scal(6);
  int *info = reinterpret_cast<int *>(scal.get() + 3);
  // init (:201-235)
  fill<T>(Nl, l, T(0), xq.get(), Nl, T(1));
  omatcopy<T>('N', Nl, l, T(1), M, Nl, bc.get(), Nl);
  gemm<T>('T', 'N', l, l, Nl, ninv, bc.get(), Nl, bc.get(), Nl, T(0), dc.get(), l);
  omatcopy<T>('N', Nl, l, ninv, bc.get(), Nl, gq.get(), Nl);
  gemm_ct<T>('T', 'N', l, Nl, Nl, T(1), gq.get(), Nl, M, Nl, T(0), bgc.get(), Nl);
  gemm<T>('T', 'N', l, l, Nl, ninv, bc.get(), Nl, bgc.get(), Nl, T(0), dgc.get(), l);
  omatcopy<T>('N', Nl, l, T(1), gq.get(), Nl, xx.get(), Nl);
  gemm<T>('N', 'N', Nl, l, l, T(-1), xq.get(), Nl, dc.get(), l, T(1), xx.get(), Nl);
  hipLaunchKernelGGL(wy_mu_kernel<T>, dim3(1), dim3(256), 0, ctx().stream, (int)l,
                     (const T *)dgc.get(), (const T *)dc.get(), scal.get());
  LAUNCH_CHECK();
  double mu = (double)read_scalar(scal.get());
  double taug = args.tau0, zeta = 1.0;
  dot<T>(Nl * l, bc.get(), bc.get(), scal.get() + 1);
  double phi = (double)half_ninv * (double)read_scalar(scal.get() + 1);
  T *fc = fcb.get();
  T *fgc = fgcb.get();
  i64 it;
  for (it = 1;; ++it) {
    // step size (:243-282); Z = Dgc - Dc^2 once per iteration, then per trip the closed form
    small_gemm<T>('T', 'N', l, l, l, T(-1), dc.get(), l, dc.get(), l, T(0), zm.get(), l);
    hipLaunchKernelGGL(axpy_diag_kernel<T>, dim3(grid_for(l * l)), dim3(256), 0, ctx().stream,
                       (int)l, zm.get(), (const T *)dgc.get(), T(1), T(0));
    LAUNCH_CHECK();
    double tau = taug, phit = phi;
    for (i64 tauiter = 1; tauiter <= args.taumaxit; ++tauiter, tau *= args.beta) {
      const T it_ = (T)(1.0 / tau);
      // S = -Z/2 - (2/tau^2) I ; S^-1 ; Fgc = -(2/tau) S^-1 ; Fc = (2/tau) Dc S^-1 - (4/tau^2) S^-1 - I
      hipLaunchKernelGGL(sym_scale_kernel<T>, dim3(grid_for(l * l)), dim3(256), 0, ctx().stream,
                         (int)l, (const T *)zm.get(), si.get(), T(-0.5), T(-2) * it_ * it_);
      LAUNCH_CHECK();
      inv_small<T>(l, si.get(), l, info);
      hipLaunchKernelGGL(sym_scale_kernel<T>, dim3(grid_for(l * l)), dim3(256), 0, ctx().stream,
                         (int)l, (const T *)si.get(), fgc, T(-2) * it_, T(0));
      LAUNCH_CHECK();
      small_gemm<T>('N', 'N', l, l, l, T(2) * it_, dc.get(), l, si.get(), l, T(0), fc, l, T(-1));
      hipLaunchKernelGGL(axpy_diag_kernel<T>, dim3(grid_for(l * l)), dim3(256), 0, ctx().stream,
                         (int)l, fc, (const T *)si.get(), T(-4) * it_ * it_, T(0));
      LAUNCH_CHECK();
      // B+ := Bc Fc + Bgc Fgc  (:270-271)
      gemm<T>('N', 'N', Nl, l, l, T(1), bc.get(), Nl, fc, l, T(0), bp.get(), Nl);
      gemm<T>('N', 'N', Nl, l, l, T(1), bgc.get(), Nl, fgc, l, T(1), bp.get(), Nl);
      dot<T>(Nl * l, bp.get(), bp.get(), scal.get() + 2);
      T hv[2];
      read_back(hv, scal.get() + 2, sizeof(hv));
      int h_info = 0;
      std::memcpy(&h_info, &hv[1], sizeof(int));
      phit = (double)half_ninv * (double)hv[0];
      ISVD_CHECK(h_info == 0, "Wen-Yin step matrix is singular (pivot %d)", h_info);
      if (trace_on())
        std::fprintf(stderr, "isvd WY iter %lld trip %lld tau %.17g phit %.17g phi %.17g\n", it,
                     tauiter, tau, phit, phi);
      if (phit >= phi + tau * args.sigma * mu) break;
    }
    // (:287-288)
    phi = (args.eta * zeta * phi + phit) / (args.eta * zeta + 1);
    zeta = args.eta * zeta + 1;
    // Q+ := Qc Fc + Gc Fgc  (:291-292)
    gemm<T>('N', 'N', Nl, l, l, T(1), xq.get(), Nl, fc, l, T(0), xqp.get(), Nl);
    gemm<T>('N', 'N', Nl, l, l, T(1), gq.get(), Nl, fgc, l, T(1), xqp.get(), Nl);
    // D+ := B+^T B+ / N (:295); G+ := Qs B+ / N (:298); Bg+ := Qs^T G+ (:301-303)
    gemm<T>('T', 'N', l, l, Nl, ninv, bp.get(), Nl, bp.get(), Nl, T(0), dc.get(), l);
    omatcopy<T>('N', Nl, l, ninv, bp.get(), Nl, gq.get(), Nl);
    gemm_ct<T>('T', 'N', l, Nl, Nl, T(1), gq.get(), Nl, M, Nl, T(0), bgc.get(), Nl);
    // Dg+ := B+^T Bg+ / N (:306); mu (:307)
    gemm<T>('T', 'N', l, l, Nl, ninv, bp.get(), Nl, bgc.get(), Nl, T(0), dgc.get(), l);
    hipLaunchKernelGGL(wy_mu_kernel<T>, dim3(1), dim3(256), 0, ctx().stream, (int)l,
                       (const T *)dgc.get(), (const T *)dc.get(), scal.get());
    LAUNCH_CHECK();
    // X+ := G+ - Q+ D+ (:324-325); Delta1, Delta2 (:328-329), queued before mu is read back so
    // that mu and t come back together (after the last iteration they go unused)
    omatcopy<T>('N', Nl, l, T(1), gq.get(), Nl, xxp.get(), Nl);
    gemm<T>('N', 'N', Nl, l, l, T(-1), xqp.get(), Nl, dc.get(), l, T(1), xxp.get(), Nl);
    hipLaunchKernelGGL(sub_kernel<T>, dim3(grid_for(Nl * l)), dim3(256), 0, ctx().stream,
                       Nl * l, (const T *)xq.get(), (const T *)xqp.get(), d1.get());
    hipLaunchKernelGGL(sub_kernel<T>, dim3(grid_for(Nl * l)), dim3(256), 0, ctx().stream,
                       Nl * l, (const T *)xx.get(), (const T *)xxp.get(), d2.get());
    LAUNCH_CHECK();
    // t (:334-342) in sketch coordinates: <D_a, D_b> = <d_a, M d_b>, M symmetric.  Odd
    // iterations take <D1, D1>, <D1, D2>, even ones <D2, D1>, <D2, D2>: both are
    // [<d1, v>, <d2, v>] with v = M d1 (odd) or M d2 (even), one product with M
    gemm_ct<T>('T', 'N', l, Nl, Nl, T(1), (it % 2) ? d1.get() : d2.get(), Nl, M, Nl, T(0),
               md1.get(), Nl);
    dot<T>(Nl * l, d1.get(), md1.get(), scal.get() + 4);
    dot<T>(Nl * l, d2.get(), md1.get(), scal.get() + 5);
    T hs[6];
    read_back(hs, scal.get(), sizeof(hs));
    mu = (double)hs[0];
    if (trace_on()) std::fprintf(stderr, "isvd WY iter %lld mu %.17g\n", it, mu);
    // convergence (:312-319)
    if (mu < args.tol * args.tol) break;
    if (it == args.maxit) {
      ++it;
      break;
    }
    const T t[2] = {hs[4], hs[5]};
    taug = std::fabs((double)t[0] / (double)t[1]);
    if (taug < args.taumin) taug = args.taumin;
    if (taug > args.taumax) taug = args.taumax;
    std::swap(xq.p, xqp.p);
    std::swap(xx.p, xxp.p);
    std::swap(bc.p, bp.p);
  }
  omatcopy<T>('N', Nl, l, T(1), xqp.get(), Nl, x_out, Nl);
  *iter_out = it;
  *mu_out = mu;
}

template <typename T>
void stage_integrate_wy(const Dims &d, const WyArgs &args, const T *yst, i64 ldyst, T *qt,
                        i64 ldqt, i64 *iter_out, double *mu_out) {
  const i64 l = d.l, Nl = d.Nl, m = d.m;
  DBuf<T> M(Nl * Nl), x(Nl * l);
  syrk_full<T>('N', Nl, m, T(1), yst, ldyst, M.get(), Nl);
  integrate_wy_coords<T>(d, args, M.get(), x.get(), iter_out, mu_out);
  // Qbar^T := xq+^T Qs^T  (:359-362)
  gemm<T>('T', 'N', l, m, Nl, T(1), x.get(), Nl, yst, ldyst, T(0), qt, ldqt);
}

// Q_i^T rows r := tmp rows r / sqrt(2 (1 + s_i[r])) for the h merged blocks
template <typename T>
__global__ void hr_scale_kernel(i64 hl, i64 m, const T *tmp, const T *s, T *yst, i64 ldyst) {
  for (i64 e = blockIdx.x * (i64)blockDim.x + threadIdx.x; e < hl * m;
       e += (i64)gridDim.x * blockDim.x) {
    const i64 r = e % hl, c = e / hl;
    yst[r + c * ldyst] = tmp[r + c * hl] / sqrt(T(2) * (T(1) + s[r]));
  }
}

// Hierarchical-reduction integration (@x@_integrate_hierarchical_reduction.c:87-135): pairwise
// merges of the sketch bases, h = floor(h2 / 2) pairs per level, all pairs of a level in one
// batched product, one batched l x l SVD and one batched update.  yst is overwritten.
template <typename T>
void stage_integrate_hr(const Dims &d, T *yst, i64 ldyst, T *qt, i64 ldqt) {
  const i64 l = d.l, N = d.N, m = d.m;
  const i64 hmax = N / 2 > 0 ? N / 2 : 1;
  DBuf<T> bs(l * l * hmax), wb(l * l * hmax), tb(l * l * hmax), s(l * hmax), tmp(hmax * l * m);
  for (i64 h2 = N; h2 > 1; h2 = (h2 + 1) / 2) {
    const i64 h = h2 / 2;
    // B_i := Q_i^T Q_{i+h}  (:93-95)
    gemm<T>('N', 'T', l, l, m, T(1), yst, ldyst, yst + h * l, ldyst, T(0), bs.get(), l, h, l, l,
            l * l);
    // svd(B_i) = W_i S_i T_i^T  (:112)
    svd_small<T>(l, bs.get(), l, l * l, s.get(), l, wb.get(), l, l * l, tb.get(), l, l * l, h);
    // Q_i^T := W_i^T Q_i^T + T_i^T Q_{i+h}^T  (:115-117), then / sqrt(2 (I + S_i))  (:120-122)
    gemm<T>('T', 'N', l, m, l, T(1), wb.get(), l, yst, ldyst, T(0), tmp.get(), h * l, h, l * l,
            l, l);
    gemm<T>('T', 'N', l, m, l, T(1), tb.get(), l, yst + h * l, ldyst, T(1), tmp.get(), h * l, h,
            l * l, l, l);
    hipLaunchKernelGGL(hr_scale_kernel<T>, dim3(grid_for(h * l * m)), dim3(256), 0, ctx().stream,
                       h * l, m, (const T *)tmp.get(), (const T *)s.get(), yst, ldyst);
    LAUNCH_CHECK();
  }
  // Qbar := Q_0  (:127)
  omatcopy<T>('N', l, m, T(1), yst, ldyst, qt, ldqt);
}

// Z^T := Qbar^T op(A)  (projectBlockRow, src/lib/libisvd/core/stage/@x@_postprocess.h:154-157)
template <typename T>
static void project(const Dims &d, char ordera, const T *a, i64 lda, const T *qt, i64 ldqt,
                    T *zt) {
  const char transa = (ordera == 'C') ? 'N' : 'T';
  gemm<T>('N', transa, d.l, d.n, d.m, T(1), qt, ldqt, a, lda, T(0), zt, d.l);
}

// Gramian postprocessing (src/lib/libisvd/core/stage/@x@_postprocess_gramian.c:135-185).
template <typename T>
void post_gramian_z(const Dims &d, const T *qt, i64 ldqt, T *zt, T *s, T *ut, i64 ldut, T *vt,
                    i64 ldvt, bool want_u, bool want_v) {
  const i64 l = d.l, k = d.k;
  DBuf<T> w(l * l), uw(l * l);
  // W := Z^T Z (:145); gesvd('O','N') + vSqrt (:149-151)
  gemm<T>('N', 'T', l, l, d.n, T(1), zt, l, zt, l, T(0), w.get(), l);
  svd_small<T>(l, w.get(), l, 0, s, l, uw.get(), l, 0, nullptr, 0, 0, 1);
  hipLaunchKernelGGL(sqrt_inplace_kernel<T>, dim3(1, 1), dim3(256), 0, ctx().stream, (int)l, s);
  LAUNCH_CHECK();
  // U^T := W[:, :k]^T Qbar^T (:157-158)
  if (want_u) gemm<T>('T', 'N', k, d.m, l, T(1), uw.get(), l, qt, ldqt, T(0), ut, ldut);
  // V^T := (W[:, :k] / S)^T Z^T (:172-174)
  if (want_v) {
    hipLaunchKernelGGL(colscale_inv_kernel<T>, dim3(grid_for(l * k)), dim3(256), 0, ctx().stream,
                       (int)l, (int)k, (const T *)s, uw.get(), l);
    LAUNCH_CHECK();
    gemm<T>('T', 'N', k, d.n, l, T(1), uw.get(), l, zt, l, T(0), vt, ldvt);
  }
}

template <typename T>
void stage_post_gramian(const Dims &d, char ordera, const T *a, i64 lda, const T *qt, i64 ldqt,
                        T *s, T *ut, i64 ldut, T *vt, i64 ldvt, bool want_u, bool want_v) {
  DBuf<T> zt(d.l * d.n);
  project<T>(d, ordera, a, lda, qt, ldqt, zt.get());
  post_gramian_z<T>(d, qt, ldqt, zt.get(), s, ut, ldut, vt, ldvt, want_u, want_v);
}

// Tall-skinny-QR postprocessing (src/lib/libisvd/core/stage/@x@_postprocess_tall_skinny_qr.c:
// 184-296) at P = 1: LQ of Z^T (as the QR of Z by TSQR), SVD of L = R^T, then
//   U = Qbar U_L[:, :k],  V = Q_lq^T V_L[:, :k].
// The reference forms U from V_L and V from the rows of U_L (:263-285); that agrees with the
// above only in projector form and only when k == l, and for k < l its V even depends on the
// sign convention of gesvd, so it is not reproduced.
template <typename T>
void post_tsqr_z(const Dims &d, const T *qt, i64 ldqt, T *zt, T *s, T *ut, i64 ldut, T *vt,
                 i64 ldvt, bool want_u, bool want_v) {
  const i64 l = d.l, k = d.k;
  ISVD_CHECK(d.n >= l, "tall-skinny QR postprocessing needs n >= l (n = %lld, l = %lld)", d.n, l);
  DBuf<T> r(l * l), lmat(l * l), ul(l * l), vl(l * l);
  // Z (row-major n x l in zt) = Q_z R  <=>  Z^T = R^T Q_z^T = L Q_lq   (:199-208)
  if (cholqr_mode() == 1 || !cholqr2_rowmajor<T>(d.n, l, zt, l, 0, 1, r.get(), l, 0))
    tsqr_rowmajor<T>(d.n, l, zt, l, 0, 1, r.get(), l, 0);
  hipLaunchKernelGGL(transpose_small_kernel<T>, dim3(grid_for(l * l)), dim3(256), 0,
                     ctx().stream, (int)l, (const T *)r.get(), l, lmat.get(), l);
  LAUNCH_CHECK();
  // gesvd(L): L = U_L S V_L^T  (:252-258)
  svd_small<T>(l, lmat.get(), l, 0, s, l, ul.get(), l, 0, vl.get(), l, 0, 1);
  // U^T := U_L[:, :k]^T Qbar^T  (:263-265)
  if (want_u) gemm<T>('T', 'N', k, d.m, l, T(1), ul.get(), l, qt, ldqt, T(0), ut, ldut);
  // V^T := V_L[:, :k]^T Q_lq    (:278-284; Q_lq = Q_z^T is zt as an l x n col-major matrix)
  if (want_v) gemm<T>('T', 'N', k, d.n, l, T(1), vl.get(), l, zt, l, T(0), vt, ldvt);
}

template <typename T>
void stage_post_tsqr(const Dims &d, char ordera, const T *a, i64 lda, const T *qt, i64 ldqt,
                     T *s, T *ut, i64 ldut, T *vt, i64 ldvt, bool want_u, bool want_v) {
  DBuf<T> zt(d.l * d.n);
  project<T>(d, ordera, a, lda, qt, ldqt, zt.get());
  post_tsqr_z<T>(d, qt, ldqt, zt.get(), s, ut, ldut, vt, ldvt, want_u, want_v);
}

// Symmetric postprocessing (src/lib/libisvd/core/stage/@x@_postprocess_symmetric.c:150-184):
// W = Z^T Qbar, syev(W) (eigenvalues ascending), U = Qbar W[:, :k].
template <typename T>
void stage_post_symmetric(const Dims &d, char ordera, const T *a, i64 lda, const T *qt, i64 ldqt,
                          T *s, T *ut, i64 ldut, bool want_u) {
  const i64 l = d.l, k = d.k;
  ISVD_CHECK(d.m == d.n, "symmetric postprocessing needs m == n (m = %lld, n = %lld)", d.m, d.n);
  DBuf<T> zt(l * d.n), w(l * l), wv(l * l);
  project<T>(d, ordera, a, lda, qt, ldqt, zt.get());
  // W := Z^T Q  (:154)
  gemm<T>('N', 'T', l, l, d.n, T(1), zt.get(), l, qt, ldqt, T(0), w.get(), l);
  // eig(W) = W S W^T  (:159)
  syev_small<T>(l, w.get(), l, s, wv.get(), l);
  // U^T := W[:, :k]^T Q^T  (:167)
  if (want_u) gemm<T>('T', 'N', k, d.m, l, T(1), wv.get(), l, qt, ldqt, T(0), ut, ldut);
}

#define INST(T)                                                                                \
  template void stage_integrate_hr<T>(const Dims &, T *, i64, T *, i64);                       \
  template void stage_post_symmetric<T>(const Dims &, char, const T *, i64, const T *, i64,    \
                                        T *, T *, i64, bool);                                  \
  template void gaussian_omega<T>(i64, i64, T *, i64, unsigned long long);                    \
  template void stage_sketch<T>(const Dims &, char, const T *, i64, const T *, i64, T *, i64); \
  template void stage_sketch_host_blocks<T>(const Dims &, char, const T *, i64, const T *, i64,  \
                                            T *, i64, i64);                                    \
  template void project_host_blocks<T>(const Dims &, char, const T *, i64, const T *, i64, T *,  \
                                       i64);                                                   \
  template void stage_orth_gramian<T>(const Dims &, T *, i64, bool);                          \
  template void stage_orth_tsqr<T>(const Dims &, T *, i64, bool);                             \
  template void stage_integrate_kn<T>(const Dims &, const KnArgs &, const T *, i64, T *, i64,  \
                                      i64 *, double *);                                        \
  template void stage_integrate_wy<T>(const Dims &, const WyArgs &, const T *, i64, T *, i64,  \
                                      i64 *, double *);                                        \
  template void stage_post_gramian<T>(const Dims &, char, const T *, i64, const T *, i64, T *, \
                                      T *, i64, T *, i64, bool, bool);                         \
  template void stage_post_tsqr<T>(const Dims &, char, const T *, i64, const T *, i64, T *,    \
                                   T *, i64, T *, i64, bool, bool);                            \
  template void integrate_kn_coords<T>(const Dims &, const KnArgs &, const T *, T *, i64 *,    \
                                       double *, const SplitSym *);                            \
  template void integrate_wy_coords<T>(const Dims &, const WyArgs &, const T *, T *, i64 *,    \
                                       double *);                                              \
  template void post_gramian_z<T>(const Dims &, const T *, i64, T *, T *, T *, i64, T *, i64,  \
                                  bool, bool);                                                 \
  template void post_tsqr_z<T>(const Dims &, const T *, i64, T *, T *, T *, i64, T *, i64,     \
                               bool, bool);
INST(double)
INST(float)
#undef INST

}  // namespace isvd
