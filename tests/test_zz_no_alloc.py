"""The no-allocation Gramian postprocessing (isvd_dPostprocessGramianNoAlloc) through the C ABI:
the same results as isvd_dPostprocessGramian on two calls in a row (the second smaller, so
stale scratch contents from the first would show), and clear errors for what it cannot take."""
import numpy as np
import pytest

from conftest import lowrank_matrix
from isvd_amd import capi

pytestmark = pytest.mark.gpu


def _post(torch, lib, fn, prm, a, qbar, ordera):
    l, m, n = prm.dim_sketch, prm.nrow, prm.ncol
    s = torch.full((l,), np.nan, dtype=torch.float64, device="cuda")
    u = torch.full((m, l), np.nan, dtype=torch.float64, device="cuda")
    v = torch.full((n, l), np.nan, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    getattr(lib, fn)(prm, None, 0, None, 0, b"R", ordera.encode(), a.data_ptr(), a.shape[1],
                     qbar.data_ptr(), l, s.data_ptr(), u.data_ptr(), l, v.data_ptr(), l, 0, 0)
    err = capi.last_error()
    assert not err, err
    k = prm.rank
    return s.cpu().numpy(), u[:, :k].cpu().numpy(), v[:, :k].cpu().numpy()


@pytest.mark.parametrize("ordera", ["C", "R"])
def test_no_alloc_matches_allocating_entry_point(torch_cuda, ordera):
    torch = torch_cuda
    lib = capi.load()
    cases = [(3000, 1500, 20, 12, 0.93, 5), (1000, 700, 10, 6, 0.8, 6)]  # second call smaller
    for m, n, k, p, decay, seed in cases:
        l = k + p
        print(f"no-alloc postprocess m={m} n={n} k={k} l={l} ordera={ordera}", flush=True)
        a = lowrank_matrix(m, n, decay ** np.arange(min(m, n)), seed)
        rng = np.random.default_rng(seed + 100)
        qbar, _ = np.linalg.qr(a @ rng.standard_normal((n, l)))
        a_dev = torch.from_numpy(np.ascontiguousarray(a.T if ordera == "C" else a)).cuda()
        q_dev = torch.from_numpy(qbar).cuda()
        prm = capi.create_param(m, n, k, p, 1)
        s0, u0, v0 = _post(torch, lib, "isvd_dPostprocessGramian", prm, a_dev, q_dev, ordera)
        s1, u1, v1 = _post(torch, lib, "isvd_dPostprocessGramianNoAlloc", prm, a_dev, q_dev,
                           ordera)
        # same kernels, same split-K choices at these sizes: identical bits
        assert np.array_equal(s0, s1) and np.array_equal(u0, u1) and np.array_equal(v0, v1)
        # and the singular values of Qbar^T A (the textbook SVD), FP64 bar
        ref = np.linalg.svd(qbar.T @ a, compute_uv=False)
        assert np.max(np.abs(s1[:k] - ref[:k]) / ref[0]) < 1e-10


def test_no_alloc_rejects_what_it_cannot_take(torch_cuda):
    torch = torch_cuda
    lib = capi.load()
    l = 128
    small = torch.zeros(1024, dtype=torch.float64, device="cuda")
    out = torch.zeros(200 * l, dtype=torch.float64, device="cuda")
    # Z^T alone (8 l n bytes) beyond the 64 MiB region: rejected before any device work
    prm = capi.create_param(200, 70000, 100, 28, 1)
    lib.isvd_dPostprocessGramianNoAlloc(prm, None, 0, None, 0, b"R", b"R", small.data_ptr(),
                                        70000, small.data_ptr(), l, out.data_ptr(),
                                        out.data_ptr(), l, out.data_ptr(), l, 0, 0)
    assert "scratch region" in capi.last_error()
    # host arrays would need staging buffers
    prm = capi.create_param(300, 200, 10, 6, 1)
    a = np.zeros((300, 200))
    q = np.zeros((300, 16))
    s, u, v = np.zeros(16), np.zeros((300, 16)), np.zeros((200, 16))
    lib.isvd_dPostprocessGramianNoAlloc(prm, None, 0, None, 0, b"R", b"R", a.ctypes.data, 200,
                                        q.ctypes.data, 16, s.ctypes.data, u.ctypes.data, 16,
                                        v.ctypes.data, 16, 0, 0)
    assert "device memory" in capi.last_error()
