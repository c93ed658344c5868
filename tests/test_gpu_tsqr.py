"""Tall-skinny QR orthogonalization at shapes that reach every path of its kernels, against
numpy's Householder QR of the same blocks, through the C ABI.  Both factorizations run: the
Householder TSQR tree (what the stage entry point takes by default) and CholeskyQR2 (what the
driver takes before KN when the device certifies it; ISVD_CHOLQR=1 makes the stage entry points
take it), including a block too ill-conditioned for CholeskyQR2, which must fall back, and the
Cholesky form of the Gramian orthogonalization (GR) in the FP64, FP32 and mixed families.

Shapes: 5 leaves (tree levels 5 -> 3 -> 2 -> 1, a lone last child at two levels), 12 leaves of
250 rows (12 -> 6 -> 3 -> 2 -> 1), l = 37 (column groups of 4 that are partial, row accesses
element by element), one leaf taller than l, and 1000 rows at l = 128 (an edge tile of the
FP32 direct-to-LDS product Y R^-1 of CholeskyQR2).  Each block's Q must be orthonormal and span
the block (equal projectors); the projector is fixed by the block alone, so a wrong reflector,
a lost lone child, a misplaced column or a wrong triangular factor gives O(1) errors."""
import numpy as np
import pytest

from isvd_amd import capi, pipeline

pytestmark = pytest.mark.gpu

TOL = {np.float64: 1e-11, np.float32: 2e-4}


@pytest.mark.parametrize("path", ["householder", "cholqr"])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("shape", [(1280, 128, 2), (3000, 100, 3), (777, 37, 3), (130, 128, 1),
                                   (1000, 128, 3)])
def test_tsqr_orthogonalization_matches_numpy_qr(torch_cuda, monkeypatch, path, dtype, shape):
    torch = torch_cuda
    monkeypatch.setenv("ISVD_CHOLQR", "0" if path == "householder" else "1")
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


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_ill_conditioned_block_falls_back_to_householder(torch_cuda, monkeypatch, dtype):
    """A block with kappa = 1e12 (FP64) / 1e5 (FP32), beyond u^-1/2, where CholeskyQR2's first
    Gram loses the small directions: the certification must refuse it, and the Householder tree
    must still return an orthonormal basis of the block's span (numpy's QR of the same values)."""
    torch = torch_cuda
    monkeypatch.setenv("ISVD_CHOLQR", "1")
    m, l, N = 2048, 64, 2
    kappa = 1e12 if dtype == np.float64 else 1e5
    rng = np.random.default_rng(7)
    blocks = []
    for _ in range(N):
        u, _ = np.linalg.qr(rng.standard_normal((m, l)))
        v, _ = np.linalg.qr(rng.standard_normal((l, l)))
        blocks.append((u * np.logspace(0, -np.log10(kappa), l)) @ v.T)
    y = np.concatenate(blocks, axis=1)
    p = capi.create_param(m, m, l, 0, N)
    yd = torch.from_numpy(np.ascontiguousarray(y, dtype=dtype)).cuda()
    pipeline.orthogonalize(p, yd, "TS")
    q = yd.double().cpu().numpy()
    y32 = y.astype(dtype).astype(np.float64)
    for i in range(N):
        qi = q[:, i * l:(i + 1) * l]
        qn, _ = np.linalg.qr(y32[:, i * l:(i + 1) * l])
        assert np.abs(qi.T @ qi - np.eye(l)).max() < TOL[dtype]
        # the span of the trailing directions is fixed only to about u kappa by the data
        assert np.abs(qi @ qi.T - qn @ qn.T).max() < {np.float64: 1e-3, np.float32: 5e-2}[dtype]


@pytest.mark.parametrize("family", ["d", "s", "b"])
@pytest.mark.parametrize("shape", [(1280, 128, 2), (777, 37, 3), (1000, 128, 3)])
def test_cholqr2_gramian_orthogonalization(torch_cuda, monkeypatch, family, shape):
    """GR in its Cholesky form (what the driver takes before KN): an orthonormal basis of each
    block's span, whatever basis of it (equal projectors with numpy's QR)."""
    torch = torch_cuda
    monkeypatch.setenv("ISVD_CHOLQR", "1")
    m, l, N = shape
    dtype = np.float64 if family == "d" else np.float32
    rng = np.random.default_rng(m + 3 * l)
    y = rng.standard_normal((m, l * N)) * np.exp2(rng.uniform(-3, 3, size=(1, l * N)))
    p = capi.create_param(m, m, l, 0, N)
    yd = torch.from_numpy(np.ascontiguousarray(y, dtype=dtype)).cuda()
    pipeline.orthogonalize(p, yd, "GR", mixed=(family == "b"))
    q = yd.double().cpu().numpy()
    y32 = y.astype(dtype).astype(np.float64)
    for i in range(N):
        qi = q[:, i * l:(i + 1) * l]
        qn, _ = np.linalg.qr(y32[:, i * l:(i + 1) * l])
        assert np.abs(qi.T @ qi - np.eye(l)).max() < TOL[dtype]
        assert np.abs(qi @ qi.T - qn @ qn.T).max() < TOL[dtype] * 10
