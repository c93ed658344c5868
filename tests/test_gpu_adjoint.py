"""Adjoint of the rank-k result (isvd_?IsvdBackward) and the torch.autograd.Function over it.

What pins it:
  - FP64 central finite differences of L(svd_k(A)) with LAPACK's dense SVD (numpy) on a
    512 x 256 matrix of rank l = 16, k = 8: there the l computed triplets are exact and the
    complement of span(U) carries no singular value, so the adjoint is the exact derivative;
  - torch.autograd.gradcheck of the C-ABI forward (the four stages) against the C-ABI backward
    on a matrix with n = l: every sketch then spans range(A), so the forward is the exact SVD
    under any perturbation and its finite differences are the true Jacobian;
  - at a full-rank spectrum (s_i = 2 * 0.9^i + 0.01): the adjoint formula evaluated in numpy
    FP64 from the GPU's own l triplets (the kernels' arithmetic), FP64 and FP32.
Loss throughout: L = sum_i w_i s_i + <G_U, U> + <G_V, V> with seeded N(0, 1) w, G_U, G_V."""
import numpy as np
import pytest

from conftest import lowrank_matrix
from isvd_amd import capi, pipeline

pytestmark = pytest.mark.gpu


def dev(torch, x, dtype=np.float64):
    return torch.from_numpy(np.ascontiguousarray(x, dtype=dtype)).cuda()


def forward(torch, a, k, p, N, seed, dtype=np.float64, alg_p="TS"):
    m, n = a.shape
    rng = np.random.default_rng(seed)
    om = dev(torch, rng.standard_normal((n, (k + p) * N)), dtype)
    ad = dev(torch, a, dtype)
    prm = capi.create_param(m, n, k, p, N)
    y = pipeline.sketch(prm, ad, omega=om)
    pipeline.orthogonalize(prm, y, "TS")
    qbar, _, _ = pipeline.integrate(prm, y, "KN")
    s, u, v = pipeline.postprocess(prm, ad, qbar, alg_p)
    return prm, ad, qbar, s, u, v


def adjoint_formula(u, s, v, gs, gu, gv, k):
    """dA from all l triplets (numpy FP64), gradients padded with zeros from k to l."""
    l = s.shape[0]
    gu_l = np.zeros_like(u); gu_l[:, :k] = gu
    gv_l = np.zeros_like(v); gv_l[:, :k] = gv
    gs_l = np.zeros(l); gs_l[:k] = gs
    J, K = u.T @ gu_l, v.T @ gv_l
    E = s[None, :] ** 2 - s[:, None] ** 2
    F = np.where(np.eye(l, dtype=bool), 0.0, 1.0 / np.where(np.eye(l, dtype=bool), 1.0, E))
    mid = (F * (J - J.T)) * s[None, :] + s[:, None] * (F * (K - K.T)) + np.diag(gs_l)
    return (u @ mid @ v.T + (gu_l - u @ J) / s[None, :] @ v.T
            + u @ ((gv_l - v @ K) / s[None, :]).T)


@pytest.mark.parametrize("alg_p", ["TS", "GR"])
def test_adjoint_matches_central_differences_of_dense_svd(torch_cuda, alg_p):
    m, n, k, p, N = 512, 256, 8, 8, 4
    l = k + p
    a = lowrank_matrix(m, n, 2 * 0.9 ** np.arange(l) + 0.01, seed=21)  # rank l
    rng = np.random.default_rng(5)
    w, GU, GV = rng.standard_normal(k), rng.standard_normal((m, k)), rng.standard_normal((n, k))
    prm, ad, qbar, s, u, v = forward(torch_cuda, a, k, p, N, seed=6, alg_p=alg_p)
    uk, vk = u[:, :k].cpu().numpy(), v[:, :k].cpu().numpy()
    da = pipeline.isvd_backward(prm, ad, qbar, dev(torch_cuda, w), dev(torch_cuda, GU),
                                dev(torch_cuda, GV), alg_p).cpu().numpy()

    def loss(x):
        U, S, Vt = np.linalg.svd(x, full_matrices=False)
        U, V = U[:, :k], Vt[:k].T
        sg = np.sign(np.sum(U * uk, axis=0))  # the forward's sign of each singular pair
        return w @ S[:k] + np.sum(GU * (U * sg)) + np.sum(GV * (V * sg))

    h = 1e-5
    for t in range(4):
        e = rng.standard_normal((m, n))
        e /= np.linalg.norm(e)
        fd = (loss(a + h * e) - loss(a - h * e)) / (2 * h)
        an = float(np.sum(da * e))
        assert abs(fd - an) <= 1e-6 * max(1.0, abs(fd)), (t, fd, an)


def test_autograd_gradcheck_through_the_c_abi(torch_cuda):
    torch = torch_cuda
    from isvd_amd.autograd import ISVDFunction
    m, n, k, p, N = 30, 10, 3, 7, 2  # l = n
    a = lowrank_matrix(m, n, np.linspace(3.0, 1.0, n), seed=8)
    rng = np.random.default_rng(9)
    om = dev(torch, rng.standard_normal((n, (k + p) * N)))
    ad = dev(torch, a).requires_grad_(True)

    def fn(x):
        return ISVDFunction.apply(x, om, k, p, N, ("GP", "TS", "KN", "TS"))

    assert torch.autograd.gradcheck(fn, (ad,), eps=1e-6, atol=1e-6, rtol=1e-4)


@pytest.mark.parametrize("dtype,tol", [(np.float64, 1e-10), (np.float32, 2e-3)])
def test_adjoint_kernels_match_formula_full_rank(torch_cuda, dtype, tol):
    m, n, k, p, N = 2048, 512, 16, 8, 8
    l = k + p
    a = lowrank_matrix(m, n, 2 * 0.9 ** np.arange(n) + 0.01, seed=31)
    rng = np.random.default_rng(32)
    w, GU, GV = rng.standard_normal(k), rng.standard_normal((m, k)), rng.standard_normal((n, k))
    prm, ad, qbar, _, _, _ = forward(torch_cuda, a, k, p, N, seed=33, dtype=dtype)
    da = pipeline.isvd_backward(prm, ad, qbar, dev(torch_cuda, w, dtype),
                                dev(torch_cuda, GU, dtype), dev(torch_cuda, GV, dtype), "TS")
    # all l triplets, exactly as the backward recomputes them
    prm_l = capi.create_param(m, n, l, 0, N)
    s, u, v = pipeline.postprocess(prm_l, ad, qbar, "TS")
    ref = adjoint_formula(u.double().cpu().numpy(), s.double().cpu().numpy(),
                          v.double().cpu().numpy(), w, GU, GV, k)
    err = np.linalg.norm(da.double().cpu().numpy() - ref) / np.linalg.norm(ref)
    assert err < tol, err


def test_autograd_loss_backward_runs_fp32(torch_cuda):
    """The differentiation workload's loss through the autograd wrapper in FP32 (smaller
    shape): the gradient is finite and has A's shape."""
    torch = torch_cuda
    from isvd_amd.autograd import isvd
    m, n, k, p, N = 4096, 1024, 32, 16, 8
    a = lowrank_matrix(m, n, 2 * 0.9 ** np.arange(n) + 0.01, seed=41, dtype=np.float32)
    rng = np.random.default_rng(42)
    om = dev(torch, rng.standard_normal((n, (k + p) * N)), np.float32)
    ad = dev(torch, a, np.float32).requires_grad_(True)
    w = dev(torch, rng.standard_normal(k), np.float32)
    gU = dev(torch, rng.standard_normal((m, k)), np.float32)
    gV = dev(torch, rng.standard_normal((n, k)), np.float32)
    s, u, v = isvd(ad, om, k, p, N)
    loss = (w * s).sum() + (gU * u).sum() + (gV * v).sum()
    loss.backward()
    g = ad.grad
    assert g.shape == (m, n) and bool(torch.isfinite(g).all())
