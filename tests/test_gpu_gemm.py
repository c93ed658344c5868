"""Sketch products Y = A Omega at shapes that take every path of the GEMM behind the sketch
stage, against numpy's FP64 product of the same (FP64 or FP32) values, through the C ABI.

Paths: direct-to-LDS tiles with op(B) K-contiguous (row-major A) and MN-contiguous
(column-major A); edge tiles in both dimensions; a partial last K step; split-K and a single
K chunk; odd sizes, which take the register-staged kernel.  The bound is elementwise
K 2^-52 (|A| |Omega|), which every summation order meets; a wrong layout, chunk swizzle, K
pairing or edge clamp gives errors of the order of the entries themselves.  The FP32 cases take
the register-staged FP32 kernel, whose results leave through the 16-byte vector stores of its
epilogue when the output's leading dimension allows it (edge tiles, split-K partials) and
through scalar stores otherwise (odd Nl); their bound is K 2^-23 (|A| |Omega|)."""
import numpy as np
import pytest

from isvd_amd import capi, pipeline

pytestmark = pytest.mark.gpu

SHAPES = [
    (1000, 520, 25, 12),   # Nl = 300: edge tiles, K = 32 steps + 8, split-K
    (2048, 4096, 64, 4),   # Nl = 256: whole tiles, split-K
    (3000, 120, 50, 6),    # K = 120 < 128: one K chunk, 7 steps + 8
    (777, 301, 21, 5),     # odd Nl and m: register-staged kernel
]


@pytest.mark.parametrize("ordera", ["R", "C"])
@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_sketch_product(torch_cuda, ordera, shape, dtype):
    torch = torch_cuda
    m, n, l, N = shape
    rng = np.random.default_rng(m + n + l)
    a = (rng.standard_normal((m, n)) * np.exp2(rng.integers(-4, 5, size=(m, 1)))).astype(dtype)
    om = rng.standard_normal((n, l * N)).astype(dtype)
    p = capi.create_param(m, n, l, 0, N)
    a_arg = torch.from_numpy(np.ascontiguousarray(a if ordera == "R" else a.T)).cuda()
    y = pipeline.sketch(p, a_arg, omega=torch.from_numpy(om).cuda(), ordera=ordera)
    y = y.cpu().numpy().astype(np.float64)
    a64, om64 = a.astype(np.float64), om.astype(np.float64)
    ref = a64 @ om64
    eps = 2.0 ** -52 if dtype == np.float64 else 2.0 ** -23
    bound = n * eps * (np.abs(a64) @ np.abs(om64))
    err = np.abs(y - ref)
    assert (err <= bound).all(), float((err / bound).max())
