import json
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden", "reference_data.npz")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X GPU (runs the HIP library)")


@pytest.fixture(scope="session")
def golden():
    d = np.load(GOLDEN, allow_pickle=False)
    out = {k: d[k] for k in d.files if not k.startswith("_")}
    out["meta"] = json.loads(bytes(d["_meta_json"]).decode())
    return out


def principal_angles(a, b):
    """Principal angles (radians) between span(a) and span(b) (columns)."""
    qa, _ = np.linalg.qr(a)
    qb, _ = np.linalg.qr(b)
    c = np.linalg.svd(qa.T @ qb, compute_uv=False)
    c = np.clip(c, -1.0, 1.0)
    # sin-based formula is accurate for tiny angles
    p = qb - qa @ (qa.T @ qb)
    s = np.linalg.svd(p, compute_uv=False)
    return np.arcsin(np.clip(s, 0.0, 1.0))


def lowrank_matrix(m, n, spectrum, seed, noise=0.0, dtype=np.float64):
    """A = U diag(s) V^T (+ noise), U/V from QR of seeded Gaussians."""
    rng = np.random.default_rng(seed)
    r = len(spectrum)
    u, _ = np.linalg.qr(rng.standard_normal((m, r)))
    v, _ = np.linalg.qr(rng.standard_normal((n, r)))
    a = (u * np.asarray(spectrum)[None, :]) @ v.T
    if noise:
        a = a + noise * rng.standard_normal((m, n))
    return a.astype(dtype)


@pytest.fixture(scope="session")
def torch_cuda():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


@pytest.fixture(scope="session")
def lib():
    from isvd_amd import capi
    return capi.load()
