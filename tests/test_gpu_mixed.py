"""Mixed-precision ("b" family) stages and driver on the GPU, through the C ABI.

A and Omega are BF16; the sketches, bases and outputs FP32.  What pins each part:
  - BF16 MFMA products (sketch): the FP64 product of the same BF16 values, elementwise, within
    the worst-case FP32 accumulation bound (K/16 + 32) u sum|a||b| (u = 2^-24; products of
    BF16 values are exact in FP32; each 32 x 32 x 16 instruction adds 16 products to the FP32
    accumulator, and the split-K partial sums add at most 16 more roundings);
  - Gram of the stacked bases (split BF16 x3: hi*hi + hi*lo + lo*hi of two-term splits) and the
    projection Qbar^T A (three-term split of Qbar against the exact BF16 A): through the stages
    that use them, against the FP64 stages on the same values;
  - the whole pipeline: against the FP64 driver (isvd_dIsvdOmega) on the identical BF16 values
    of A and Omega widened to FP64 -- sigma relative error <= 1e-5 and principal angles of U, V
    <= 1e-4 (the accuracy target of the mixed path), for a fast (0.95^i) and a slow (i^-0.5)
    singular-value decay; and against the reference's golden vectors where they apply."""
import numpy as np
import pytest

from conftest import lowrank_matrix, principal_angles
from isvd_amd import capi, pipeline

pytestmark = pytest.mark.gpu

U32 = 2.0 ** -24


def bf16_dev(torch, x):
    """numpy FP64 -> torch BF16 on the GPU (round to nearest even)."""
    return torch.from_numpy(np.ascontiguousarray(x, dtype=np.float32)).cuda().to(torch.bfloat16)


def values(t):
    return t.double().cpu().numpy()


def dev(torch, x, dtype):
    return torch.from_numpy(np.ascontiguousarray(x, dtype=dtype)).cuda()


@pytest.mark.parametrize("ordera", ["C", "R"])
@pytest.mark.parametrize("shape", [(1024, 512, 16, 8), (1001, 997, 11, 12), (4096, 2048, 128, 2),
                                   (130, 4100, 32, 3)])
def test_sketch_bf16_within_fp32_accumulation_bound(torch_cuda, ordera, shape):
    m, n, l, N = shape
    rng = np.random.default_rng(m + n)
    a = rng.standard_normal((m, n))
    om = rng.standard_normal((n, N * l))
    a_b = bf16_dev(torch_cuda, a if ordera == "R" else a.T)
    om_b = bf16_dev(torch_cuda, om)
    p = capi.create_param(m, n, l, 0, N)
    y = pipeline.sketch(p, a_b, omega=om_b, ordera=ordera)
    assert str(y.dtype) == "torch.float32"
    a64 = values(a_b) if ordera == "R" else values(a_b).T
    om64 = values(om_b)
    ref = a64 @ om64
    bound = (n / 16 + 32) * U32 * (np.abs(a64) @ np.abs(om64))
    err = np.abs(y.double().cpu().numpy() - ref)
    assert (err <= bound).all(), float((err / bound).max())


def test_sketch_golden_bf16(golden, torch_cuda):
    """The reference's sketch fixture (a . o_gp), with A and Omega rounded to BF16."""
    a, om = golden["a"], golden["o_gp"]
    m, n = a.shape
    p = capi.create_param(m, n, 11, 0, 12)
    a_b, om_b = bf16_dev(torch_cuda, a), bf16_dev(torch_cuda, om)
    y = pipeline.sketch(p, a_b, omega=om_b).double().cpu().numpy()
    # rounding A and Omega to BF16 perturbs each product by <= 2 * 2^-9 relative
    bound = 2.0 ** -7 * (np.abs(a) @ np.abs(om))
    assert (np.abs(y - golden["ys_gp"]) <= bound).all()
    # and against the exact product of the rounded values, to FP32 accumulation accuracy
    ref = values(a_b) @ values(om_b)
    assert (np.abs(y - ref) <= (n / 16 + 32) * U32 * (np.abs(values(a_b)) @ np.abs(values(om_b)))).all()


@pytest.mark.parametrize("alg", ["GR", "TS"])
def test_orthogonalize_mixed_golden(golden, torch_cuda, alg):
    """Gramian twice in FP32 (FP64 eigensolves) / FP32 TSQR: same projectors as the
    reference's fixture, orthonormal to FP32 accuracy."""
    p = capi.create_param(101, 101, 11, 0, 12)
    y = dev(torch_cuda, golden["ys"], np.float32)
    pipeline.orthogonalize(p, y, alg, mixed=True)
    q = y.double().cpu().numpy()
    for i in range(12):
        qi, gi = q[:, 11 * i:11 * (i + 1)], golden["qs"][:, 11 * i:11 * (i + 1)]
        assert np.abs(qi @ qi.T - gi @ gi.T).max() < 2e-5
        assert np.abs(qi.T @ qi - np.eye(11)).max() < 2e-6


def test_orthogonalize_gramian_twice_repairs_ill_conditioning(torch_cuda):
    """cond(Y_i) = 1e3: one FP32 Gramian pass loses ~u cond^2 = 6e-2 of orthogonality, the
    second pass restores it."""
    m, l, N = 4096, 32, 4
    rng = np.random.default_rng(7)
    blocks = []
    for i in range(N):
        u, _ = np.linalg.qr(rng.standard_normal((m, l)))
        v, _ = np.linalg.qr(rng.standard_normal((l, l)))
        blocks.append((u * np.logspace(0, -3, l)[None, :]) @ v.T)
    ys = np.concatenate(blocks, axis=1)
    p = capi.create_param(m, m, l, 0, N)
    y = dev(torch_cuda, ys, np.float32)
    pipeline.orthogonalize(p, y, "GR", mixed=True)
    q = y.double().cpu().numpy()
    for i in range(N):
        qi = q[:, l * i:l * (i + 1)]
        assert np.abs(qi.T @ qi - np.eye(l)).max() < 1e-5
        assert principal_angles(qi, ys[:, l * i:l * (i + 1)]).max() < 1e-4


@pytest.mark.parametrize("alg,it_ref", [("KN", 81), ("WY", 21)])
def test_integrate_mixed_golden(golden, torch_cuda, alg, it_ref):
    """Split-BF16 Gram + FP64 iterations on the reference's fixture bases."""
    p = capi.create_param(101, 101, 11, 0, 12)
    qs = dev(torch_cuda, golden["qs"], np.float32)
    argv = [256, 1e-4] if alg == "KN" else [256, 1e-3]
    q, it, _ = pipeline.integrate(p, qs, alg, argv, mixed=True)
    assert abs(it - it_ref) <= 1
    ref = golden["q_kn" if alg == "KN" else "q_wy"]
    qh = q.double().cpu().numpy()
    assert np.abs(qh @ qh.T - ref @ ref.T).max() < 1e-4


def test_integrate_mixed_kn_at_the_workload_sketch_count(torch_cuda):
    """The FP32 KN loop at N = 32, l = 128 (Nl = 4096, the 131072 x 16384 workload's sketch
    count and size): M held as its two BF16 terms, every product with M on the BF16 matrix
    cores (five terms against the exact three-term split of the right-hand side, K-chunk
    partials folded in a fixed order).  Against the FP64 stage on the same FP64-orthonormal
    bases: the same iteration count (+-1) and Qbar's subspace within the mixed path's 1e-4
    target."""
    torch = torch_cuda
    m, n, l, N = 8192, 4096, 128, 32
    rng = np.random.default_rng(21)
    a = lowrank_matrix(m, n, 1.0 / np.arange(1, 2049), seed=12, noise=1e-6)
    y = a @ rng.standard_normal((n, N * l))
    qs = np.concatenate([np.linalg.qr(y[:, i * l:(i + 1) * l])[0] for i in range(N)], axis=1)
    p = capi.create_param(m, n, 100, l - 100, N)
    q64, it64, _ = pipeline.integrate(p, dev(torch, qs, np.float64), "KN", [256, 1e-4])
    q32, it32, _ = pipeline.integrate(p, dev(torch, qs, np.float32), "KN", [256, 1e-4],
                                      mixed=True)
    ang = principal_angles(q32.double().cpu().numpy(), q64.cpu().numpy()).max()
    print(f"mixed KN N=32: iter {it32} (FP64 {it64}), angle {ang:.2e}")
    assert it64 < 256 and abs(it32 - it64) <= 1
    assert ang < 1e-4


@pytest.mark.parametrize("alg", ["GR", "TS"])
def test_postprocess_mixed_golden(golden, torch_cuda, alg):
    """sigma of Qbar^T A with A rounded to BF16 against the FP64 postprocess of the same
    rounded A (the reference fixture s.mtx only pins the unrounded A)."""
    a, q = golden["a"], golden["q"]
    p = capi.create_param(101, 1001, 11, 0, 1)
    a_b = bf16_dev(torch_cuda, a)
    s, u, v = pipeline.postprocess(p, a_b, dev(torch_cuda, q, np.float32), alg)
    s64, u64, v64 = pipeline.postprocess(p, dev(torch_cuda, values(a_b), np.float64),
                                         dev(torch_cuda, q.astype(np.float32), np.float64), alg)
    sh, sr = s.double().cpu().numpy(), s64.cpu().numpy()
    assert np.abs(sh - sr).max() / sr.max() < 1e-5
    uh, ur = u.double().cpu().numpy(), u64.cpu().numpy()
    vh, vr = v.double().cpu().numpy(), v64.cpu().numpy()
    assert np.abs(uh @ uh.T - ur @ ur.T).max() < 1e-5
    assert np.abs(vh @ vh.T - vr @ vr.T).max() < 1e-5
    # and the reference's fixture, up to the BF16 rounding of A
    s_ref = golden["s"].ravel()
    assert np.abs(sh - s_ref).max() / s_ref.max() < 2.0 ** -7


def _pair(torch, m, n, k, p, N, spectrum, seed, algs, noise=1e-6):
    a = lowrank_matrix(m, n, spectrum, seed, noise=noise)
    rng = np.random.default_rng(seed + 1)
    om = rng.standard_normal((n, (k + p) * N))
    a_b, om_b = bf16_dev(torch, a), bf16_dev(torch, om)
    prm = capi.create_param(m, n, k, p, N)
    rb = pipeline.isvd(prm, a_b, omega=om_b, algs=algs)
    rd = pipeline.isvd(prm, dev(torch, values(a_b), np.float64),
                       omega=dev(torch, values(om_b), np.float64), algs=algs)
    return rb, rd, values(a_b)


@pytest.mark.parametrize("decay", ["fast", "slow"])
@pytest.mark.parametrize("algs", [("GP", "GR", "KN", "GR"), ("GP", "TS", "KN", "TS"),
                                  ("GP", "GR", "WY", "TS")])
def test_mixed_driver_matches_fp64_driver(torch_cuda, decay, algs):
    m, n, k, p, N = 8192, 2048, 20, 12, 8
    r = 512
    spectrum = 0.95 ** np.arange(r) if decay == "fast" else (np.arange(r) + 1.0) ** -0.5
    rb, rd, _ = _pair(torch_cuda, m, n, k, p, N, spectrum, 11, algs)
    assert str(rb["s"].dtype) == "torch.float32"
    # KN (tol 1e-4 on ||C - I||) converges to its fixed point, so the two drivers differ by
    # precision only; WY stops at ||grad|| < 1e-3, a tolerance-sized distance from the fixed
    # point that two trajectories reach at different places
    s_tol, ang_tol = (1e-5, 1e-4) if algs[2] == "KN" else (1e-3, 1e-2)
    sb = rb["s"][:k].double().cpu().numpy()
    sd = rd["s"][:k].cpu().numpy()
    assert np.abs(sb - sd).max() / sd.min() < s_tol
    ub, ud = rb["u"][:, :k].double().cpu().numpy(), rd["u"][:, :k].cpu().numpy()
    vb, vd = rb["v"][:, :k].double().cpu().numpy(), rd["v"][:, :k].cpu().numpy()
    assert principal_angles(ub, ud).max() < ang_tol
    assert principal_angles(vb, vd).max() < ang_tol
    assert np.abs(ub.T @ ub - np.eye(k)).max() < 1e-5
    if algs[2] == "KN":  # WY's nonmonotone line search branches on rounding-level differences
        assert abs(rb["iter"] - rd["iter"]) <= 2


def test_mixed_driver_seeded_is_deterministic(torch_cuda):
    m, n, k, p, N = 2000, 1000, 20, 12, 16
    a = lowrank_matrix(m, n, 0.9 ** np.arange(1000), 3)
    a_b = bf16_dev(torch_cuda, a)
    prm = capi.create_param(m, n, k, p, N)
    r1 = pipeline.isvd(prm, a_b, seed=5, algs=("GP", "GR", "KN", "GR"))
    r2 = pipeline.isvd(prm, a_b, seed=5, algs=("GP", "GR", "KN", "GR"))
    assert torch_cuda.equal(r1["s"], r2["s"])
    assert torch_cuda.equal(r1["u"][:, :k], r2["u"][:, :k])  # the first k columns are output
    # a randomized estimate of the leading singular values (sanity, not parity)
    s_true = np.linalg.svd(values(a_b), compute_uv=False)[:k]
    assert np.abs(r1["s"][:k].double().cpu().numpy() - s_true).max() / s_true[0] < 5e-2
