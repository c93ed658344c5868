"""The reference derivative used by the adjoint's GPU tests, pinned on the CPU: the closed form
over all n singular triplets against FP64 central differences of the dense-SVD loss (numpy
LAPACK), at the differentiation workload's spectrum s_i = 2 * 0.9^i + 0.01 (512 x 256, k = 8)
and at a rank-deficient A (left and right null directions)."""
import numpy as np
import pytest

from adjoint_ref import dense_loss, dense_svd_grad, full_rank_problem

torch = pytest.importorskip("torch")


@pytest.mark.parametrize("rank", [256, 40])
def test_dense_svd_grad_matches_central_differences(rank):
    m, n, k = 512, 256, 8
    a, u, s, v = full_rank_problem(m, n)
    if rank < n:  # zero the tail: repeated zero singular values in the null space
        s = s.copy()
        s[rank:] = 0.0
        a = (u * s) @ v.T
    rng = np.random.default_rng(5)
    w, GU, GV = rng.standard_normal(k), rng.standard_normal((m, k)), rng.standard_normal((n, k))
    t = torch.from_numpy
    ref = dense_svd_grad(torch, t(u), t(s), t(v), t(w), t(GU), t(GV), k).numpy()
    h = 1e-5
    for _ in range(3):
        e = rng.standard_normal((m, n))
        e /= np.linalg.norm(e)
        fd = (dense_loss(a + h * e, k, w, GU, GV, u[:, :k])
              - dense_loss(a - h * e, k, w, GU, GV, u[:, :k])) / (2 * h)
        assert abs(fd - np.sum(ref * e)) <= 1e-6 * max(1.0, abs(fd)), (fd, np.sum(ref * e))


def test_dense_svd_grad_drops_a_plausible_mistake():
    """Dropping the coupling to the discarded pairs (the rank-l truncation) is 4% off at this
    spectrum: central differences tell the two apart."""
    m, n, k, l = 512, 256, 8, 16
    a, u, s, v = full_rank_problem(m, n)
    rng = np.random.default_rng(5)
    w, GU, GV = rng.standard_normal(k), rng.standard_normal((m, k)), rng.standard_normal((n, k))
    t = torch.from_numpy
    full = dense_svd_grad(torch, t(u), t(s), t(v), t(w), t(GU), t(GV), k).numpy()
    s_l = s.copy()
    s_l[l:] = 0.0
    trunc = dense_svd_grad(torch, t(u), t(s_l), t(v), t(w), t(GU), t(GV), k).numpy()
    assert np.linalg.norm(trunc - full) / np.linalg.norm(full) > 1e-2
