"""FP64 sketch products Y = A Omega at shapes that take every path of the GEMM behind the
sketch stage, against numpy's FP64 product, through the C ABI.

Paths: direct-to-LDS tiles with op(B) K-contiguous (row-major A) and MN-contiguous
(column-major A); edge tiles in both dimensions; a partial last K step; split-K and a single
K chunk; odd sizes, which take the register-staged kernel.  The bound is elementwise
K 2^-52 (|A| |Omega|), which every summation order meets; a wrong layout, chunk swizzle, K
pairing or edge clamp gives errors of the order of the entries themselves."""
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
def test_sketch_product_fp64(torch_cuda, ordera, shape):
    torch = torch_cuda
    m, n, l, N = shape
    rng = np.random.default_rng(m + n + l)
    a = rng.standard_normal((m, n)) * np.exp2(rng.integers(-4, 5, size=(m, 1)))
    om = rng.standard_normal((n, l * N))
    p = capi.create_param(m, n, l, 0, N)
    a_arg = torch.from_numpy(np.ascontiguousarray(a if ordera == "R" else a.T)).cuda()
    y = pipeline.sketch(p, a_arg, omega=torch.from_numpy(om).cuda(), ordera=ordera)
    y = y.cpu().numpy()
    ref = a @ om
    bound = n * 2.0 ** -52 * (np.abs(a) @ np.abs(om))
    err = np.abs(y - ref)
    assert (err <= bound).all(), float((err / bound).max())
