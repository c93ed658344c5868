"""Tall-skinny QR orthogonalization (TSQR tree) at shapes that reach every path of its kernels,
against numpy's Householder QR of the same blocks, through the C ABI.

Shapes: 5 leaves (tree levels 5 -> 3 -> 2 -> 1, a lone last child at two levels), 12 leaves of
250 rows (12 -> 6 -> 3 -> 2 -> 1), l = 37 (column groups of 4 that are partial, row accesses
element by element), and one leaf taller than l.  Each block's Q must be orthonormal and span
the block (equal projectors); the projector is fixed by the block alone, so a wrong reflector,
a lost lone child or a misplaced column gives O(1) errors."""
import numpy as np
import pytest

from isvd_amd import capi, pipeline

pytestmark = pytest.mark.gpu

TOL = {np.float64: 1e-11, np.float32: 2e-4}


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("shape", [(1280, 128, 2), (3000, 100, 3), (777, 37, 3), (130, 128, 1)])
def test_tsqr_orthogonalization_matches_numpy_qr(torch_cuda, dtype, shape):
    torch = torch_cuda
    m, l, N = shape
    rng = np.random.default_rng(m + l)
    # columns scaled over two decades: a moderately conditioned sketch
    y = rng.standard_normal((m, l * N)) * np.exp2(rng.uniform(-3, 3, size=(1, l * N)))
    p = capi.create_param(m, m, l, 0, N)
    yd = torch.from_numpy(np.ascontiguousarray(y, dtype=dtype)).cuda()
    pipeline.orthogonalize(p, yd, "TS")
    q = yd.double().cpu().numpy()
    y32 = y.astype(dtype).astype(np.float64)  # the values the library saw
    for i in range(N):
        qi = q[:, i * l:(i + 1) * l]
        qn, _ = np.linalg.qr(y32[:, i * l:(i + 1) * l])
        assert np.abs(qi.T @ qi - np.eye(l)).max() < TOL[dtype]
        assert np.abs(qi @ qi.T - qn @ qn.T).max() < TOL[dtype] * 10
