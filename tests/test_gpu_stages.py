"""GPU parity of every stage against the reference's golden vectors (data/*.mtx, l = 11,
N = 12, m = 101, n = 1001) and the CPU oracle, called through the C ABI.

Tolerances: the reference tests' derr = 1e-8 (double) / serr = 1e-3 (single)
(check/check/libisvd/core/stage/@x@_*.cxx), entrywise where the reference compares entrywise
(KN/WY Qbar, sigma) and on projectors where it compares projectors (orthogonalization, U/V)."""
import numpy as np
import pytest

import oracle as O
from isvd_amd import capi, pipeline

pytestmark = pytest.mark.gpu

TOL = {np.float64: 1e-8, np.float32: 1e-3}


def proj(q):
    return q @ q.T


def dev(torch, x, dtype):
    return torch.from_numpy(np.ascontiguousarray(x, dtype=dtype)).cuda()


def host(t):
    return t.double().cpu().numpy() if hasattr(t, "cpu") else np.asarray(t, dtype=np.float64)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("ordera", ["C", "R"])
@pytest.mark.parametrize("dista", ["C", "R"])
def test_sketch_golden(golden, torch_cuda, dtype, ordera, dista):
    a, om = golden["a"], golden["o_gp"]
    m, n = a.shape
    p = capi.create_param(m, n, 11, 0, 12)
    a_arg = dev(torch_cuda, a if ordera == "R" else a.T, dtype)
    y = pipeline.sketch(p, a_arg, omega=dev(torch_cuda, om, dtype), ordera=ordera, dista=dista)
    tol = TOL[dtype] * (1 if dtype == np.float64 else np.abs(golden["ys_gp"]).max())
    assert np.abs(host(y) - golden["ys_gp"]).max() < tol


def test_sketch_host_pointers_match_device(golden, torch_cuda):
    a, om = golden["a"], golden["o_gp"]
    p = capi.create_param(101, 1001, 11, 0, 12)
    y_host = pipeline.sketch(p, np.ascontiguousarray(a), omega=np.ascontiguousarray(om))
    y_dev = pipeline.sketch(p, dev(torch_cuda, a, np.float64), omega=dev(torch_cuda, om, np.float64))
    assert isinstance(y_host, np.ndarray)
    assert np.array_equal(y_host, host(y_dev))


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("alg", ["GR", "TS"])
def test_orthogonalize_golden(golden, torch_cuda, dtype, alg):
    ys = golden["ys"]
    p = capi.create_param(101, 101, 11, 0, 12)
    y = dev(torch_cuda, ys, dtype)
    pipeline.orthogonalize(p, y, alg)
    q = host(y)
    assert np.abs(proj(q) - proj(golden["qs"])).max() < TOL[dtype]
    # the Gramian method squares cond(Y_i): its orthogonality is bounded by the oracle's own
    # loss of orthogonality on the same input (TS is orthogonal to rounding)
    q_ref = (O.orthogonalize_gramian if alg == "GR" else O.orthogonalize_tall_skinny_qr)(
        ys.astype(dtype), 12).astype(np.float64)
    for i in range(12):
        qi, ri = q[:, 11 * i:11 * (i + 1)], q_ref[:, 11 * i:11 * (i + 1)]
        ref_err = np.abs(ri.T @ ri - np.eye(11)).max()
        bound = 10 * ref_err + (1e-13 if dtype == np.float64 else 1e-5)
        assert np.abs(qi.T @ qi - np.eye(11)).max() < bound


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_kn_golden(golden, torch_cuda, dtype):
    p = capi.create_param(101, 101, 11, 0, 12)
    q, it, err = pipeline.integrate(p, dev(torch_cuda, golden["qs"], dtype), "KN", [256, 1e-4])
    if dtype == np.float64:
        assert it == 81
        assert err <= 1e-4
    assert np.abs(host(q) - golden["q_kn"]).max() < TOL[dtype]


# C = (I/2 + (I/4 - Z)^(1/2))^(1/2) and C^-1 are taken from a Taylor series of Z when rho is
# small, by Newton-Schulz otherwise, and from the eigendecomposition of Z when rho is out of
# reach of both.  ISVD_KN_NS=1 keeps the series out, ISVD_KN_EIG=1 takes the eigendecomposition
# in every iteration: each path must reproduce the reference's vectors and iteration counts.
KN_PATHS = ["ISVD_KN_NS", "ISVD_KN_EIG"]


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("path", KN_PATHS)
def test_kn_golden_other_paths(golden, torch_cuda, dtype, path, monkeypatch):
    monkeypatch.setenv(path, "1")
    p = capi.create_param(101, 101, 11, 0, 12)
    q, it, err = pipeline.integrate(p, dev(torch_cuda, golden["qs"], dtype), "KN", [256, 1e-4])
    if dtype == np.float64:
        assert it == 81
        assert err <= 1e-4
    assert np.abs(host(q) - golden["q_kn"]).max() < TOL[dtype]


@pytest.mark.parametrize("path", KN_PATHS)
def test_kn_l128_other_paths_match_oracle(torch_cuda, path, monkeypatch):
    from conftest import lowrank_matrix
    monkeypatch.setenv(path, "1")
    m, n, k, p, N = 4096, 1024, 100, 28, 4
    a = lowrank_matrix(m, n, 1.0 / np.arange(1, n + 1), seed=5, noise=1e-6)
    qs = O.orthogonalize_tall_skinny_qr(a @ np.random.default_rng(6).standard_normal((n, 512)), N)
    q_ref, it_ref, _ = O.integrate_kolmogorov_nagumo(qs, N)
    q, it, _ = pipeline.integrate(capi.create_param(m, n, k, p, N), dev(torch_cuda, qs, np.float64),
                                  "KN")
    assert it == it_ref
    assert np.abs(host(q) - q_ref).max() < 1e-10


def test_wy_golden(golden, torch_cuda):
    # single-precision WY is excluded by the reference too (check/CMakeLists.txt:86-88)
    p = capi.create_param(101, 101, 11, 0, 12)
    q, it, mu = pipeline.integrate(p, dev(torch_cuda, golden["qs"], np.float64), "WY",
                                   [256, 1e-3])
    assert it == 21
    assert np.abs(host(q) - golden["q_wy"]).max() < 1e-8


@pytest.mark.parametrize("alg", ["KN", "WY"])
def test_integrate_l128_same_input_as_oracle(torch_cuda, alg):
    # the integration stage at l = 128 on exactly the oracle's Qs: same iteration count and
    # Qbar entrywise
    from conftest import lowrank_matrix
    m, n, k, p, N = 8192, 2048, 100, 28, 4
    a = lowrank_matrix(m, n, 1.0 / np.arange(1, n + 1), seed=2, noise=1e-6)
    qs = O.orthogonalize_tall_skinny_qr(a @ np.random.default_rng(3).standard_normal((n, 512)), N)
    fn = O.integrate_kolmogorov_nagumo if alg == "KN" else O.integrate_wen_yin
    q_ref, it_ref, _ = fn(qs, N)
    prm = capi.create_param(m, n, k, p, N)
    q, it, _ = pipeline.integrate(prm, dev(torch_cuda, qs, np.float64), alg)
    assert it == it_ref
    # KN is a contraction: rounding-level agreement.  WY's BB steps amplify rounding: on this
    # instance the oracle's own Qbar moves by ~1e-8 (entrywise) under a 1e-14 perturbation of Qs
    assert np.abs(host(q) - q_ref).max() < (1e-10 if alg == "KN" else 1e-7)


def test_integrate_matches_oracle_trajectory(torch_cuda):
    # same Qs, same iteration count and Qbar as the oracle at a random instance
    rng = np.random.default_rng(0)
    a = rng.standard_normal((600, 200)) * (0.9 ** np.arange(200))[None, :]
    qs = O.orthogonalize_gramian(a @ rng.standard_normal((200, 8 * 12)), 8)
    p = capi.create_param(600, 200, 8, 4, 8)
    for alg, fn in (("KN", O.integrate_kolmogorov_nagumo), ("WY", O.integrate_wen_yin)):
        q_ref, it_ref, _ = fn(qs, 8)
        q, it, _ = pipeline.integrate(p, dev(torch_cuda, qs, np.float64), alg)
        assert it == it_ref
        assert np.abs(host(q) - q_ref).max() < 1e-8


@pytest.mark.parametrize("alg", ["GR", "TS"])
@pytest.mark.parametrize("ordera", ["C", "R"])
@pytest.mark.parametrize("dista", ["C", "R"])
@pytest.mark.parametrize("job", ["gather", "block", "none"])
def test_postprocess_golden(golden, torch_cuda, alg, ordera, dista, job):
    # check/.../@x@_postprocess_gramian.cxx:4-51: every dista x ordera x job combination
    a, qbar = golden["a"], golden["q"]
    m, n = a.shape
    p = capi.create_param(m, n, 11, 0, 1)
    lib = capi.load()
    torch = torch_cuda
    a_arg = dev(torch, a if ordera == "R" else a.T, np.float64)
    q_arg = dev(torch, qbar, np.float64)
    s = torch.zeros(11, dtype=torch.float64, device="cuda")
    ld = 11 if job == "gather" else 13  # block jobs accept ld >= l
    u = torch.zeros((m, ld), dtype=torch.float64, device="cuda")
    v = torch.zeros((n, ld), dtype=torch.float64, device="cuda")
    root = {"gather": 0, "block": -1, "none": -2}[job]
    fn = getattr(lib, "isvd_dPostprocessGramian" if alg == "GR" else "isvd_dPostprocessTallSkinnyQr")
    fn(p, None, 0, None, 0, dista.encode(), ordera.encode(), a_arg.data_ptr(), a_arg.shape[1],
       q_arg.data_ptr(), 11, s.data_ptr(), u.data_ptr(), ld, v.data_ptr(), ld, root, root)
    assert capi.last_error() == ""
    assert np.abs(host(s) - golden["s"][:, 0]).max() < 1e-8
    if job == "none":
        assert float(u.abs().max()) == 0.0 and float(v.abs().max()) == 0.0
    else:
        uu, vv = host(u)[:, :11], host(v)[:, :11]
        assert np.abs(proj(uu) - proj(golden["u"])).max() < 1e-8
        assert np.abs(proj(vv) - proj(golden["v"])).max() < 1e-8
        if job == "block":
            assert float(u[:, 11:].abs().max()) == 0.0  # ld padding untouched


@pytest.mark.parametrize("alg", ["GR", "TS"])
def test_postprocess_k_less_than_l_matches_textbook_svd(torch_cuda, alg):
    # k < l: U, V must be the top-k singular vectors of Qbar Qbar^T A (the reference's TS
    # path swaps factors here; see test_oracle.test_post_tsqr_reference_swap_defect)
    from conftest import lowrank_matrix, principal_angles
    a = lowrank_matrix(300, 180, 0.8 ** np.arange(60), seed=3, noise=1e-4)
    qbar, _ = np.linalg.qr(a @ np.random.default_rng(4).standard_normal((180, 16)))
    p = capi.create_param(300, 180, 6, 10, 1)
    s, u, v = pipeline.postprocess(p, dev(torch_cuda, a, np.float64),
                                   dev(torch_cuda, qbar, np.float64), alg)
    s_ref, u_ref, v_ref = O.postprocess_tall_skinny_qr(a, qbar, 6)
    assert np.abs(host(s) - s_ref).max() < 1e-10 * s_ref[0]
    assert principal_angles(host(u)[:, :6], u_ref).max() < 1e-8
    assert principal_angles(host(v)[:, :6], v_ref).max() < 1e-8


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_hr_golden(golden, torch_cuda, dtype):
    # check/.../@x@_integrate_hierarchical_reduction.cxx: projector of Qbar vs q_hr.mtx
    p = capi.create_param(101, 101, 11, 0, 12)
    q, _, _ = pipeline.integrate(p, dev(torch_cuda, golden["qs"], dtype), "HR")
    assert np.abs(proj(host(q)) - proj(golden["q_hr"])).max() < TOL[dtype]


def test_hr_matches_oracle_l128(torch_cuda):
    # odd level sizes (N = 7: 7 -> 4 -> 2 -> 1) follow the reference's loop
    from conftest import lowrank_matrix
    m, l, N = 3000, 128, 7
    a = lowrank_matrix(m, 900, 0.97 ** np.arange(900), seed=12)
    y = a @ np.random.default_rng(13).standard_normal((900, N * l))
    qs = O.orthogonalize_tall_skinny_qr(y, N)
    p = capi.create_param(m, m, 100, 28, N)
    q, _, _ = pipeline.integrate(p, dev(torch_cuda, qs, np.float64), "HR")
    q_ref = O.integrate_hierarchical_reduction(qs, N)
    assert np.abs(proj(host(q)) - proj(q_ref)).max() < 1e-10


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("ordera", ["C", "R"])
@pytest.mark.parametrize("job", ["gather", "u_only", "v_only"])
def test_postprocess_symmetric_golden(golden, torch_cuda, dtype, ordera, job):
    # check/.../@x@_postprocess_symmetric.hxx: s vs s_sym.mtx (ascending), U projector
    a, qbar = golden["a_sym"], golden["q"]
    m = a.shape[0]
    p = capi.create_param(m, m, 11, 0, 1)
    lib = capi.load()
    torch = torch_cuda
    tdt = torch.float64 if dtype == np.float64 else torch.float32
    a_arg = dev(torch, a if ordera == "R" else a.T, dtype)
    q_arg = dev(torch, qbar, dtype)
    s = torch.zeros(11, dtype=tdt, device="cuda")
    u = torch.zeros((m, 11), dtype=tdt, device="cuda")
    v = torch.zeros((m, 11), dtype=tdt, device="cuda")
    ur, vr = {"gather": (0, 0), "u_only": (0, -2), "v_only": (-2, 0)}[job]
    x = "d" if dtype == np.float64 else "s"
    getattr(lib, f"isvd_{x}PostprocessSymmetric")(
        p, None, 0, None, 0, b"R", ordera.encode(), a_arg.data_ptr(), m, q_arg.data_ptr(), 11,
        s.data_ptr(), u.data_ptr(), 11, v.data_ptr(), 11, ur, vr)
    assert capi.last_error() == ""
    tol = TOL[dtype] * (1 if dtype == np.float64 else golden["s_sym"].max())
    assert np.abs(host(s) - golden["s_sym"][:, 0]).max() < tol
    ptol = TOL[dtype]
    if job in ("gather", "u_only"):
        assert np.abs(proj(host(u)) - proj(golden["u_sym"])).max() < ptol
    if job in ("gather", "v_only"):
        assert np.abs(proj(host(v)) - proj(golden["u_sym"])).max() < ptol
    if job == "u_only":
        assert float(v.abs().max()) == 0.0


@pytest.mark.parametrize("algs", [("GP", "TS", "HR", "TS"), ("GP", "GR", "KN", "SY")])
def test_driver_hr_and_sy_match_oracle(torch_cuda, algs):
    from conftest import lowrank_matrix
    m, k, p, N = 600, 10, 6, 4
    if algs[3] == "SY":
        rng = np.random.default_rng(17)
        vq, _ = np.linalg.qr(rng.standard_normal((m, m)))
        a = (vq * (0.9 ** np.arange(m))[None, :]) @ vq.T
        n = m
    else:
        n = 400
        a = lowrank_matrix(m, n, 0.9 ** np.arange(n), seed=18)
    om = np.random.default_rng(19).standard_normal((n, (k + p) * N))
    prm = capi.create_param(m, n, k, p, N)
    r = pipeline.isvd(prm, dev(torch_cuda, a, np.float64), omega=dev(torch_cuda, om, np.float64),
                      algs=algs)
    y = O.sketch_gaussian_projection(a, om)
    qs = (O.orthogonalize_tall_skinny_qr if algs[1] == "TS" else O.orthogonalize_gramian)(y, N)
    if algs[2] == "HR":
        qbar = O.integrate_hierarchical_reduction(qs, N)
    else:
        qbar, _, _ = O.integrate_kolmogorov_nagumo(qs, N)
    if algs[3] == "SY":
        s_ref, u_ref = O.postprocess_symmetric(a, qbar, k)
    else:
        s_ref, u_ref, _ = O.postprocess_tall_skinny_qr(a, qbar, k)
    assert np.abs(host(r["s"]) - s_ref).max() < 1e-10 * np.abs(s_ref).max()
    from conftest import principal_angles
    assert principal_angles(host(r["u"])[:, :k], u_ref).max() < 1e-8
