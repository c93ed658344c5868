// BF16 conversions for the mixed-precision pipeline: round-to-nearest-even BF16 of FP32 data,
// the materialized BF16 splits x = h0 + h1 (+ h2) that turn the compensated products into
// plain BF16 GEMMs (bf16_gemm.hip), and the small kernels that combine their results.
#include "mx.hpp"

namespace isvd {

namespace {

__device__ __forceinline__ uint32_t bf16_rne(float x) {
  const uint32_t u = __float_as_uint(x);
  return (u + 0x7FFFu + ((u >> 16) & 1u)) >> 16;  // finite inputs
}
__device__ __forceinline__ float bf16_val(uint32_t h) { return __uint_as_float(h << 16); }

template <typename Tout, typename Tin>
__global__ void convert_kernel(i64 rows, i64 cols, const Tin *src, i64 lds, Tout *dst, i64 ldd) {
  for (i64 e = blockIdx.x * (i64)blockDim.x + threadIdx.x; e < rows * cols;
       e += (i64)gridDim.x * blockDim.x) {
    const i64 i = e % rows, j = e / rows;
    dst[i + j * ldd] = (Tout)src[i + j * lds];
  }
}

// dst rows p*rows + i of column j := term p of the BF16 split of src(i, j)
__global__ void split_kernel(i64 rows, i64 cols, int parts, const float *src, i64 lds,
                             uint16_t *dst, i64 ldd) {
  for (i64 e = blockIdx.x * (i64)blockDim.x + threadIdx.x; e < rows * cols;
       e += (i64)gridDim.x * blockDim.x) {
    const i64 i = e % rows, j = e / rows;
    float r = src[i + j * lds];
    for (int q = 0; q < parts; ++q) {
      const uint32_t h = bf16_rne(r);
      dst[q * rows + i + j * ldd] = (uint16_t)h;
      r -= bf16_val(h);
    }
  }
}

// dst(i, j) := sum over p of src(p*rows + i, j), in the order p = 0, 1, ...
__global__ void sum_blocks_kernel(i64 rows, i64 cols, int parts, const float *src, i64 lds,
                                  float *dst, i64 ldd) {
  for (i64 e = blockIdx.x * (i64)blockDim.x + threadIdx.x; e < rows * cols;
       e += (i64)gridDim.x * blockDim.x) {
    const i64 i = e % rows, j = e / rows;
    float v = src[i + j * lds];
    for (int q = 1; q < parts; ++q) v += src[q * rows + i + j * lds];
    dst[i + j * ldd] = v;
  }
}

// M(i, j) := HH(min(i, j), max(i, j)) + (LH(i, j) + LH(j, i)) with HH = rows 0..n-1 (its
// upper triangle) and LH = rows n..2n-1 of C (2n x n): hi*hi + hi*lo + lo*hi, exactly symmetric
__global__ void gram_combine_kernel(i64 n, const float *C, i64 ldc, float *M, i64 ldm) {
  for (i64 e = blockIdx.x * (i64)blockDim.x + threadIdx.x; e < n * n;
       e += (i64)gridDim.x * blockDim.x) {
    const i64 i = e % n, j = e / n;
    const float hh = (i <= j) ? C[i + j * ldc] : C[j + i * ldc];
    const float lh = C[n + i + j * ldc] + C[n + j + i * ldc];
    M[i + j * ldm] = hh + lh;
  }
}

__global__ void f32_to_bf16_kernel(i64 count, const float *src, uint16_t *dst) {
  for (i64 e = blockIdx.x * (i64)blockDim.x + threadIdx.x; e < count;
       e += (i64)gridDim.x * blockDim.x)
    dst[e] = (uint16_t)bf16_rne(src[e]);
}

// bop (2n x 3l, ld 2n) from B (n x l): B = b0 + b1 + b2, its BF16 terms (exact for FP32), laid
// out as the column blocks [b0; b0], [b1; b1], [b2; 0], so that
// [Mh | Ml] bop = [Mh b0 + Ml b0 | Mh b1 + Ml b1 | Mh b2]
__global__ void split_sym_rhs_kernel(i64 n, i64 l, const float *B, i64 ldb, uint16_t *bop) {
  const i64 ld = 2 * n;
  for (i64 e = blockIdx.x * (i64)blockDim.x + threadIdx.x; e < n * l;
       e += (i64)gridDim.x * blockDim.x) {
    const i64 i = e % n, j = e / n;
    float r = B[i + j * ldb];
    const uint32_t h0 = bf16_rne(r);
    r -= bf16_val(h0);
    const uint32_t h1 = bf16_rne(r);
    r -= bf16_val(h1);
    const uint32_t h2 = bf16_rne(r);
    uint16_t *c0 = bop + j * ld, *c1 = bop + (l + j) * ld, *c2 = bop + (2 * l + j) * ld;
    c0[i] = (uint16_t)h0;
    c0[n + i] = (uint16_t)h0;
    c1[i] = (uint16_t)h1;
    c1[n + i] = (uint16_t)h1;
    c2[i] = (uint16_t)h2;
    c2[n + i] = 0;
  }
}

// G(i, j) := alpha * the sum over chunks s = 0, 1, ... and, within a chunk, over the column
// blocks c = 0, 1, 2 of W[s](i, c l + j) (W[s] n x 3l dense), in that fixed order
__global__ void split_sym_fold_kernel(i64 n, i64 l, int parts, const float *W, float alpha,
                                      float *G, i64 ldg) {
  const i64 blk = n * l;
  for (i64 e = blockIdx.x * (i64)blockDim.x + threadIdx.x; e < blk;
       e += (i64)gridDim.x * blockDim.x) {
    const i64 i = e % n, j = e / n;
    const float *w = W + i + j * n;
    float v = 0.f;
    for (int q = 0; q < parts; ++q, w += 3 * blk) {
      v += w[0];
      v += w[blk];
      v += w[2 * blk];
    }
    G[i + j * ldg] = alpha * v;
  }
}

// B(i, j) := Mh(i, j) + Ml(i, j) (exact in FP32) for S = [Mh; Ml] (2n x n, ld 2n), j < l
__global__ void split_sym_columns_kernel(i64 n, i64 l, const uint16_t *S, float *B, i64 ldb) {
  for (i64 e = blockIdx.x * (i64)blockDim.x + threadIdx.x; e < n * l;
       e += (i64)gridDim.x * blockDim.x) {
    const i64 i = e % n, j = e / n;
    B[i + j * ldb] = bf16_val(S[i + j * 2 * n]) + bf16_val(S[n + i + j * 2 * n]);
  }
}

int grid_of(i64 n) {
  i64 b = (n + 255) / 256;
  return (int)(b < 1 ? 1 : (b > 8192 ? 8192 : b));
}

}  // namespace

template <typename Tout, typename Tin>
void convert(i64 rows, i64 cols, const Tin *src, i64 lds, Tout *dst, i64 ldd) {
  if (rows <= 0 || cols <= 0) return;
  hipLaunchKernelGGL((convert_kernel<Tout, Tin>), dim3(grid_of(rows * cols)), dim3(256), 0,
                     ctx().stream, rows, cols, src, lds, dst, ldd);
  LAUNCH_CHECK();
}

void f32_to_bf16(i64 count, const float *src, uint16_t *dst) {
  if (count <= 0) return;
  hipLaunchKernelGGL(f32_to_bf16_kernel, dim3(grid_of(count)), dim3(256), 0, ctx().stream, count,
                     src, dst);
  LAUNCH_CHECK();
}

template void convert<double, float>(i64, i64, const float *, i64, double *, i64);
template void convert<float, double>(i64, i64, const double *, i64, float *, i64);


void split_bf16(i64 rows, i64 cols, int parts, const float *src, i64 lds, uint16_t *dst,
                i64 ldd) {
  if (rows <= 0 || cols <= 0) return;
  ISVD_CHECK(ldd >= parts * rows, "split_bf16: ldd too small");
  hipLaunchKernelGGL(split_kernel, dim3(grid_of(rows * cols)), dim3(256), 0, ctx().stream, rows,
                     cols, parts, src, lds, dst, ldd);
  LAUNCH_CHECK();
}

void sum_blocks(i64 rows, i64 cols, int parts, const float *src, i64 lds, float *dst, i64 ldd) {
  if (rows <= 0 || cols <= 0) return;
  hipLaunchKernelGGL(sum_blocks_kernel, dim3(grid_of(rows * cols)), dim3(256), 0, ctx().stream,
                     rows, cols, parts, src, lds, dst, ldd);
  LAUNCH_CHECK();
}

void gram_combine(i64 n, const float *C, i64 ldc, float *M, i64 ldm) {
  if (n <= 0) return;
  hipLaunchKernelGGL(gram_combine_kernel, dim3(grid_of(n * n)), dim3(256), 0, ctx().stream, n,
                     C, ldc, M, ldm);
  LAUNCH_CHECK();
}

i64 split_sym_gw_size(i64 n, i64 l, int max_split) { return (i64)max_split * n * 3 * l; }

void split_sym_apply(const SplitSym &m, i64 l, const float *B, i64 ldb, float alpha, float *G,
                     i64 ldg) {
  const i64 n = m.n;
  if (n <= 0 || l <= 0) return;
  hipLaunchKernelGGL(split_sym_rhs_kernel, dim3(grid_of(n * l)), dim3(256), 0, ctx().stream, n,
                     l, B, ldb, m.bop);
  LAUNCH_CHECK();
  // [Mh | Ml] = S^T (Mh, Ml symmetric), K = 2n
  const int parts =
      gemm_bf16_partials('T', 'N', n, 3 * l, 2 * n, m.S, 2 * n, m.bop, 2 * n, m.gw, m.max_split);
  hipLaunchKernelGGL(split_sym_fold_kernel, dim3(grid_of(n * l)), dim3(256), 0, ctx().stream, n,
                     l, parts, (const float *)m.gw, alpha, G, ldg);
  LAUNCH_CHECK();
}

void split_sym_columns(const SplitSym &m, i64 l, float *B, i64 ldb) {
  if (m.n <= 0 || l <= 0) return;
  hipLaunchKernelGGL(split_sym_columns_kernel, dim3(grid_of(m.n * l)), dim3(256), 0,
                     ctx().stream, m.n, l, m.S, B, ldb);
  LAUNCH_CHECK();
}

}  // namespace isvd
