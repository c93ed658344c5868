"""Adjoint of the rank-k result (isvd_?IsvdBackward) and the torch.autograd.Function over it.

The reference gradient is the derivative of L(U_k, S_k, V_k) for the top-k triplets of the
dense SVD of A, L = sum_i w_i s_i + <G_U, U> + <G_V, V> (seeded N(0, 1) w, G_U, G_V):
adjoint_ref.dense_svd_grad() evaluates the textbook closed form over ALL n singular triplets
of A, itself pinned by FP64 central differences in test_adjoint_reference.py.  What pins each
part of the GPU adjoint:
  - the adjoint at a full-rank spectrum (s_i = 2 * 0.9^i + 0.01, the differentiation workload's
    spectrum, 512 x 256, k = 8, l = 16) with Qbar = the exact top-l left singular vectors:
    the backward must then be the dense-SVD derivative itself (its complement series carries
    the coupling to the 240 discarded singular pairs) -- FP64 to 1e-10, FP32 to its measured
    precision, and FP64 central differences;
  - with Qbar from the iSVD forward: the same derivative evaluated at the computed subspace,
    within a bound set by the forward's own subspace error;
  - rank(A) = l: central differences (the complement series stops at its first term);
  - torch.autograd.gradcheck of the C-ABI forward (the four stages) against the C-ABI backward
    with n = l (the forward is then the exact SVD under any perturbation);
  - the differentiation workload at full size (65536 x 8192 FP32, k = 64, l = 96): A is built
    from known singular factors, so the dense-SVD gradient is computed exactly in FP64 from
    them; the FP32 backward with the exact top-l subspace must match it."""
import numpy as np
import pytest

from adjoint_ref import dense_loss, dense_svd_grad, full_rank_problem
from conftest import lowrank_matrix, principal_angles
from isvd_amd import capi, pipeline

pytestmark = pytest.mark.gpu


def dev(torch, x, dtype=np.float64):
    return torch.from_numpy(np.ascontiguousarray(x, dtype=dtype)).cuda()


def align(u, v, u_fwd, k):
    """Exact factors with the signs of the first k singular pairs set to the forward's (the
    gradients refer to the U and V the forward returned)."""
    sg = np.sign(np.sum(u_fwd[:, :k] * u[:, :k], axis=0))
    u, v = u.copy(), v.copy()
    u[:, :k] *= sg
    v[:, :k] *= sg
    return u, v


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


@pytest.mark.parametrize("alg_p", ["TS", "GR"])
@pytest.mark.parametrize("dtype,tol", [(np.float64, 1e-10), (np.float32, 4e-6)])
def test_adjoint_full_rank_exact_subspace_is_the_dense_svd_derivative(torch_cuda, alg_p, dtype,
                                                                      tol):
    """Full-rank A, Qbar = exact top-l: the backward is the dense-SVD derivative (measured FP64
    1.7e-15, FP32 4e-7 relative on gfx950; the FP32 bound allows 10x)."""
    torch = torch_cuda
    m, n, k, p = 512, 256, 8, 8
    l = k + p
    a, u, s, v = full_rank_problem(m, n)
    rng = np.random.default_rng(5)
    w, GU, GV = rng.standard_normal(k), rng.standard_normal((m, k)), rng.standard_normal((n, k))
    prm = capi.create_param(m, n, k, p, 1)
    ad = dev(torch, a, dtype)
    qbar = dev(torch, u[:, :l], dtype)
    da = pipeline.isvd_backward(prm, ad, qbar, dev(torch, w, dtype), dev(torch, GU, dtype),
                                dev(torch, GV, dtype), alg_p).double().cpu().numpy()
    _, u_fwd, _ = pipeline.postprocess(prm, ad, qbar, alg_p)  # the forward's U for this Qbar
    u, v = align(u, v, u_fwd.double().cpu().numpy(), k)
    t = lambda x: torch.from_numpy(x)
    ref = dense_svd_grad(torch, t(u), t(s), t(v), t(w), t(GU), t(GV), k).numpy()
    err = np.linalg.norm(da - ref) / np.linalg.norm(ref)
    print(f"exact subspace {alg_p} {np.dtype(dtype).name}: rel err vs dense-SVD gradient {err:.2e}")
    assert err < tol, err
    if dtype == np.float64:
        h = 1e-5
        for _ in range(2):
            e = rng.standard_normal((m, n))
            e /= np.linalg.norm(e)
            fd = (dense_loss(a + h * e, k, w, GU, GV, u[:, :k])
                  - dense_loss(a - h * e, k, w, GU, GV, u[:, :k])) / (2 * h)
            assert abs(fd - float(np.sum(da * e))) <= 1e-6 * max(1.0, abs(fd))


@pytest.mark.parametrize("p,N", [(8, 4), (56, 8)])
def test_adjoint_full_rank_isvd_subspace(torch_cuda, p, N):
    """Full-rank A, Qbar from the iSVD forward: the derivative at the computed subspace.  Its
    distance to the dense-SVD derivative follows the forward's own error: the relative
    difference stays below 1.5 times the largest principal angle between the computed and the
    exact top-k singular subspaces (measured ratio 0.64 at l = 16 and at l = 64)."""
    torch = torch_cuda
    m, n, k = 512, 256, 8
    a, u, s, v = full_rank_problem(m, n)
    rng = np.random.default_rng(5)
    w, GU, GV = rng.standard_normal(k), rng.standard_normal((m, k)), rng.standard_normal((n, k))
    prm, ad, qbar, s_g, u_g, v_g = forward(torch, a, k, p, N, seed=6)
    da = pipeline.isvd_backward(prm, ad, qbar, dev(torch, w), dev(torch, GU), dev(torch, GV),
                                "TS").cpu().numpy()
    uu, vv = align(u, v, u_g.cpu().numpy(), k)
    t = lambda x: torch.from_numpy(x)
    ref = dense_svd_grad(torch, t(uu), t(s), t(vv), t(w), t(GU), t(GV), k).numpy()
    err = np.linalg.norm(da - ref) / np.linalg.norm(ref)
    ang = max(principal_angles(u_g[:, :k].cpu().numpy(), u[:, :k]).max(),
              principal_angles(v_g[:, :k].cpu().numpy(), v[:, :k]).max())
    print(f"iSVD subspace l={k + p} N={N}: rel err {err:.2e}, forward angle {ang:.2e}")
    assert err < 1.5 * ang + 1e-10, (err, ang)


@pytest.mark.parametrize("alg_p", ["TS", "GR"])
def test_adjoint_rank_l_matches_central_differences(torch_cuda, alg_p):
    """rank(A) = l: the complement carries no singular value."""
    m, n, k, p, N = 512, 256, 8, 8, 4
    l = k + p
    a = lowrank_matrix(m, n, 2 * 0.9 ** np.arange(l) + 0.01, seed=21)
    rng = np.random.default_rng(5)
    w, GU, GV = rng.standard_normal(k), rng.standard_normal((m, k)), rng.standard_normal((n, k))
    prm, ad, qbar, s, u, v = forward(torch_cuda, a, k, p, N, seed=6, alg_p=alg_p)
    uk = u[:, :k].cpu().numpy()
    da = pipeline.isvd_backward(prm, ad, qbar, dev(torch_cuda, w), dev(torch_cuda, GU),
                                dev(torch_cuda, GV), alg_p).cpu().numpy()
    h = 1e-5
    for t in range(4):
        e = rng.standard_normal((m, n))
        e /= np.linalg.norm(e)
        fd = (dense_loss(a + h * e, k, w, GU, GV, uk) - dense_loss(a - h * e, k, w, GU, GV, uk)) / (2 * h)
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


def exact_factors(torch, m, n, seed):
    """Seeded orthonormal U (m x n), V (n x n) in FP64 on the GPU (CholeskyQR2)."""
    g = torch.Generator(device="cuda").manual_seed(seed)

    def orth(rows, cols):
        x = torch.randn(rows, cols, generator=g, device="cuda", dtype=torch.float64)
        eye = torch.eye(cols, device="cuda", dtype=torch.float64)
        for _ in range(2):  # x := x R^-1 with R^T R = x^T x (R = c^T)
            c = torch.linalg.cholesky(x.T @ x)
            x = x @ torch.linalg.solve_triangular(c.T, eye, upper=True)
        return x

    return orth(m, n), orth(n, n)


def test_differentiation_workload_full_size(torch_cuda):
    """65536 x 8192 FP32, s_i = 2 * 0.9^i + 0.01, k = 64, l = 96: the FP32 backward with the
    exact top-l subspace against the FP64 dense-SVD gradient computed from the known factors;
    and the autograd path (iSVD forward, N = 16) runs and gives a finite gradient."""
    torch = torch_cuda
    m, n, k, p, N = 65536, 8192, 64, 32, 16
    l = k + p
    u, v = exact_factors(torch, m, n, 2025)
    s = 2 * 0.9 ** torch.arange(n, device="cuda", dtype=torch.float64) + 0.01
    a64 = (u * s[None, :]) @ v.T
    a32 = a64.float()
    del a64
    g = torch.Generator(device="cuda").manual_seed(7)
    w = torch.randn(k, generator=g, device="cuda", dtype=torch.float64)
    gU = torch.randn(m, k, generator=g, device="cuda", dtype=torch.float64)
    gV = torch.randn(n, k, generator=g, device="cuda", dtype=torch.float64)
    prm = capi.create_param(m, n, k, p, N)
    print(f"differentiation workload {m}x{n} k={k} l={l}: backward with exact Qbar", flush=True)
    qbar = u[:, :l].float().contiguous()
    da = pipeline.isvd_backward(prm, a32, qbar, w.float(), gU.float(), gV.float(), "TS")
    # signs of the forward's U for this Qbar, applied to the exact factors
    _, u_fwd, _ = pipeline.postprocess(prm, a32, qbar, "TS")
    sg = torch.sign((u_fwd[:, :k].double() * u[:, :k]).sum(0))
    u[:, :k] *= sg
    v[:, :k] *= sg
    ref = dense_svd_grad(torch, u, s, v, w, gU, gV, k)
    err = float(torch.linalg.norm(da.double() - ref) / torch.linalg.norm(ref))
    print(f"  FP32 backward vs FP64 dense-SVD gradient: rel err {err:.2e}")
    del ref, da
    assert err < 2e-3, err
    from isvd_amd.autograd import isvd
    om = torch.randn(n, l * N, generator=g, device="cuda", dtype=torch.float32)
    ad = a32.clone().requires_grad_(True)
    s_, u_, v_ = isvd(ad, om, k, p, N)
    loss = (w.float() * s_).sum() + (gU.float() * u_).sum() + (gV.float() * v_).sum()
    loss.backward()
    assert ad.grad.shape == (m, n) and bool(torch.isfinite(ad.grad).all())
    u_, v_ = u_.detach(), v_.detach()
    assert float((u_.T @ u_ - torch.eye(k, device="cuda")).abs().max()) < 1e-4
    assert float((v_.T @ v_ - torch.eye(k, device="cuda")).abs().max()) < 1e-4
