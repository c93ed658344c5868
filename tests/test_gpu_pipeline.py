"""End-to-end GPU parity of the driver (isvd_?IsvdOmega / isvd_?Isvd) against the CPU oracle
on the same A and host-generated Omega.

Bar (FP64): singular values to 1e-10 relative, U / V subspaces to a principal angle of 1e-8.
Shapes: the 2000 x 1000 CPU-runnable workload (k=20, p=12, N=16, GP/GR/KN/GR) in full, and
reduced instances of the 131072 x 16384 workload (same l = 128, same selectors) sized so the
oracle finishes in seconds."""
import numpy as np
import pytest

import oracle as O
from conftest import lowrank_matrix, principal_angles
from isvd_amd import capi, pipeline

pytestmark = pytest.mark.gpu


def host(t):
    return t.double().cpu().numpy()


def run_both(torch, a, om, k, p, N, algs, dtype=np.float64, ordera="R", argv_i=None):
    m, n = a.shape
    prm = capi.create_param(m, n, k, p, N)
    tdt = torch.float64 if dtype == np.float64 else torch.float32
    a_arg = torch.from_numpy(np.ascontiguousarray(a if ordera == "R" else a.T)).to("cuda", tdt)
    om_d = torch.from_numpy(np.ascontiguousarray(om)).to("cuda", tdt)
    g = pipeline.isvd(prm, a_arg, omega=om_d, algs=algs, ordera=ordera, argv_i=argv_i)
    kw = {}
    if argv_i:
        kw = dict(maxit=argv_i[0], tol=argv_i[1]) if len(argv_i) >= 2 else dict(maxit=argv_i[0])
    r = O.isvd(a, om, k, N, algs[1], algs[2], algs[3], argv_i=kw)
    return g, r


def check(g, r, k, s_tol=1e-10, ang_tol=1e-8, iter_slack=0):
    s = host(g["s"])
    assert abs(g["iter"] - r["iter"]) <= iter_slack, (g["iter"], r["iter"])
    assert np.abs(s[:k] - r["s"][:k]).max() <= s_tol * r["s"][0], \
        np.abs(s[:k] - r["s"][:k]).max() / r["s"][0]
    assert principal_angles(host(g["u"])[:, :k], r["u"]).max() < ang_tol
    assert principal_angles(host(g["v"])[:, :k], r["v"]).max() < ang_tol


def test_cpu_runnable_workload(torch_cuda):
    # 2000 x 1000, s_i = 0.9^i, k=20, p=12, N=16; GP / GR / KN / GR (mirrors demo/cdemo.c:77-105)
    a = lowrank_matrix(2000, 1000, 0.9 ** np.arange(1000), seed=0)
    om = np.random.default_rng(1).standard_normal((1000, 16 * 32))
    for ordera in ("R", "C"):
        g, r = run_both(torch_cuda, a, om, 20, 12, 16, ("GP", "GR", "KN", "GR"), ordera=ordera)
        check(g, r, 20)
        assert all(t > 0 for t in g["time"])


@pytest.mark.parametrize("algs", [("GP", "TS", "KN", "TS"), ("GP", "TS", "WY", "TS"),
                                  ("GP", "GR", "WY", "GR")])
def test_reduced_gpu_scale_workload(torch_cuda, algs):
    # same l = 128 as the 131072 x 16384 workload; s_i = 1/i (i < 2048) + 1e-6 noise
    m, n, k, p, N = 8192, 2048, 100, 28, 4
    a = lowrank_matrix(m, n, 1.0 / np.arange(1, 2049)[:n], seed=2, noise=1e-6)
    om = np.random.default_rng(3).standard_normal((n, N * (k + p)))
    g, r = run_both(torch_cuda, a, om, k, p, N, algs)
    if algs[2] == "KN":
        check(g, r, k)
    else:
        # Wen-Yin's BB step / nonmonotone line search amplify rounding: the oracle itself moves
        # sigma by ~3e-11 under a 1e-15 perturbation of Qs at this shape, and the Gramian
        # orthogonalization perturbs Qs by ~u cond(Y_i)^2; both runs stop at mu < tol^2
        check(g, r, k, s_tol=1e-8, ang_tol=1e-6, iter_slack=2)


def test_single_precision_driver(torch_cuda):
    a = lowrank_matrix(2000, 1000, 0.9 ** np.arange(1000), seed=4)
    om = np.random.default_rng(5).standard_normal((1000, 16 * 32))
    g, r = run_both(torch_cuda, a, om, 20, 12, 16, ("GP", "GR", "KN", "GR"), dtype=np.float32)
    s = host(g["s"])
    assert np.abs(s[:20] - r["s"][:20]).max() <= 1e-3 * r["s"][0]
    assert principal_angles(host(g["u"])[:, :10], r["u"][:, :10]).max() < 1e-2


def test_seeded_driver_is_deterministic_and_accurate(torch_cuda):
    # Omega drawn on the GPU from the seed: same seed -> bitwise identical result
    spec = 0.8 ** np.arange(300)
    a = lowrank_matrix(3000, 600, spec, seed=6)
    prm = capi.create_param(3000, 600, 10, 6, 8)
    a_d = torch_cuda.from_numpy(a).cuda()
    g1 = pipeline.isvd(prm, a_d, seed=42, algs=("GP", "TS", "KN", "TS"))
    g2 = pipeline.isvd(prm, a_d, seed=42, algs=("GP", "TS", "KN", "TS"))
    g3 = pipeline.isvd(prm, a_d, seed=43, algs=("GP", "TS", "KN", "TS"))
    assert torch_cuda.equal(g1["s"], g2["s"])
    assert torch_cuda.equal(g1["u"][:, :10], g2["u"][:, :10])  # columns >= k are not written
    assert not torch_cuda.equal(g1["s"], g3["s"])
    # rank-10 approximation of a slowly decaying spectrum with l = 16: sigma within 1e-3
    assert np.abs(host(g1["s"])[:10] - spec[:10]).max() < 1e-3


def test_seeded_omega_is_standard_gaussian(torch_cuda):
    # the sketch with A = I returns Omega^T itself: check its moments
    n, Nl = 512, 4 * 64
    prm = capi.create_param(n, n, 60, 4, 4)
    eye = torch_cuda.eye(n, dtype=torch_cuda.float64, device="cuda")
    y = host(pipeline.sketch(prm, eye, seed=7))  # Y = I * Omega = Omega (n x Nl)
    assert abs(y.mean()) < 0.01 and abs(y.std() - 1) < 0.01
    assert abs(np.mean(y ** 4) - 3) < 0.1


# ---- the 131072 x 16384 workload's sketch count and size (N = 32, l = 128, Nl = 4096) on a
# reduced A, so that the oracle (full-space KN / WY iterations over m x Nl bases) finishes in
# about 10-20 s per integration on the CPU
N32 = dict(m=8192, n=4096, k=100, p=28, N=32)


def n32_problem():
    m, n, k, p, N = (N32[x] for x in ("m", "n", "k", "p", "N"))
    a = lowrank_matrix(m, n, 1.0 / np.arange(1, 2049), seed=12, noise=1e-6)
    om = np.random.default_rng(13).standard_normal((n, N * (k + p)))
    return a, om


@pytest.mark.parametrize("ordera", ["C", "R"])
def test_n32_kolmogorov_nagumo_matches_oracle(torch_cuda, ordera):
    """GP / TS / KN / TS at N = 32, l = 128: the FP64 bar (sigma 1e-10, angles 1e-8, same
    iteration count).  Exercises the Nl x Nl = 4096 x 4096 Gram (upper tiles, split-K), the
    transposed-store M B product and the 32-sketch TSQR tree."""
    a, om = n32_problem()
    g, r = run_both(torch_cuda, a, om, N32["k"], N32["p"], N32["N"], ("GP", "TS", "KN", "TS"),
                    ordera=ordera)
    print(f"N=32 KN ordera={ordera}: iter {g['iter']} (oracle {r['iter']})")
    check(g, r, N32["k"])


def test_n32_wen_yin_within_the_oracles_own_rounding_sensitivity(torch_cuda):
    """GP / TS / WY / TS at N = 32, l = 128.  Wen-Yin's Barzilai-Borwein steps and nonmonotone
    line search amplify rounding, so the tolerance is measured here: the oracle is run twice,
    on Qs and on Qs perturbed by one unit roundoff (2^-52 relative, elementwise), and the GPU
    result (whose products round differently everywhere) must lie within 10x that spread of
    the oracle (never looser than sigma 1e-8 / angle 1e-6, never tighter than the FP64 bar),
    with the iteration count within the spread plus one."""
    k, N = N32["k"], N32["N"]
    a, om = n32_problem()
    g, r = run_both(torch_cuda, a, om, k, N32["p"], N, ("GP", "TS", "WY", "TS"))
    rng = np.random.default_rng(14)
    qs2 = r["qs"] * (1.0 + 2.0 ** -52 * rng.standard_normal(r["qs"].shape))
    q2, it2, _ = O.integrate_wen_yin(qs2, N)
    s2, u2, v2 = O.postprocess_tall_skinny_qr(a, q2, k)
    ds = np.abs(s2[:k] - r["s"][:k]).max() / r["s"][0]
    da = max(principal_angles(u2, r["u"]).max(), principal_angles(v2, r["v"]).max())
    s_tol = min(max(10 * ds, 1e-10), 1e-8)
    a_tol = min(max(10 * da, 1e-8), 1e-6)
    slack = abs(it2 - r["iter"]) + 1
    s = host(g["s"])
    gs = np.abs(s[:k] - r["s"][:k]).max() / r["s"][0]
    ga = max(principal_angles(host(g["u"])[:, :k], r["u"]).max(),
             principal_angles(host(g["v"])[:, :k], r["v"]).max())
    print(f"N=32 WY: oracle self-spread sigma {ds:.2e} angle {da:.2e} iter {it2} vs "
          f"{r['iter']}; GPU sigma {gs:.2e} angle {ga:.2e} iter {g['iter']}")
    assert abs(g["iter"] - r["iter"]) <= slack
    assert gs <= s_tol and ga <= a_tol, (gs, s_tol, ga, a_tol)


def test_n32_mixed_precision_matches_fp64_oracle(torch_cuda):
    """The b family (BF16 A and Omega; GR twice in FP32, split-BF16 Gram, FP64 l x l solves,
    three-term BF16 projection) at N = 32, l = 128 against the FP64 oracle on the identical
    BF16 values: sigma within 1e-5 relative and principal angles within 1e-4 (the mixed path's
    accuracy target)."""
    torch = torch_cuda
    k, p, N = N32["k"], N32["p"], N32["N"]
    a, om = n32_problem()
    a_b = torch.from_numpy(a.astype(np.float32)).cuda().to(torch.bfloat16)
    om_b = torch.from_numpy(om.astype(np.float32)).cuda().to(torch.bfloat16)
    a64, om64 = host(a_b), host(om_b)
    prm = capi.create_param(a.shape[0], a.shape[1], k, p, N)
    algs = ("GP", "GR", "KN", "TS")
    g = pipeline.isvd(prm, a_b, omega=om_b, algs=algs)
    r = O.isvd(a64, om64, k, N, "GR", "KN", "TS")
    s = host(g["s"])
    gs = np.abs(s[:k] - r["s"][:k]).max() / r["s"][0]
    ga = max(principal_angles(host(g["u"])[:, :k], r["u"]).max(),
             principal_angles(host(g["v"])[:, :k], r["v"]).max())
    print(f"N=32 mixed vs FP64 oracle: sigma {gs:.2e} angle {ga:.2e} iter {g['iter']} "
          f"(oracle {r['iter']})")
    assert gs <= 1e-5 and ga <= 1e-4, (gs, ga)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_seeded_omega_is_the_philox_stream(torch_cuda, dtype):
    """Omega drawn from the seed (read back as Y = I Omega, exact) against the CPU restatement
    (oracle/philox.py, pinned to Random123's and rocRAND's Philox outputs).  The Philox words
    are integer arithmetic; the Box-Muller transform uses the device's log / cos, so FP64 is
    held to 4 ulp of each element and FP32 (the FP64 value rounded once) to identity except
    where that rounding falls within those 4 ulp of a tie."""
    from oracle import philox as P
    torch = torch_cuda
    n, Nl, seed = 1024, 8 * 64, -12345
    prm = capi.create_param(n, n, 60, 4, 8)
    tdt = torch.float64 if dtype == np.float64 else torch.float32
    eye = torch.eye(n, dtype=tdt, device="cuda")
    y = pipeline.sketch(prm, eye, seed=seed).cpu().numpy()
    ref = P.omega(seed, n, Nl)
    if dtype == np.float64:
        ulp = np.spacing(np.abs(ref))
        assert (np.abs(y - ref) <= 4 * ulp).all(), float((np.abs(y - ref) / ulp).max())
    else:
        r32 = ref.astype(np.float32)
        differ = y != r32
        assert differ.sum() <= 2, int(differ.sum())
        assert (np.abs(y[differ] - r32[differ]) <= np.spacing(np.abs(r32[differ]))).all()
