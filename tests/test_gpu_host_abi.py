"""Host-resident arrays through the C ABI, as the reference's C callers pass them.

  - isvd_gpu_memory_limit (src/lib/libisvd/gpu/env.c:19): with a limit below the size of a
    host-resident A, the sketch and the projection stage A by blocks of columns, as the
    reference's GPU stages do (src/lib/libisvd/gpu/stage/@x@_sketch_gpu.h:63-125,
    @x@_postprocess_gpu.h:73-140); the result matches the unblocked run (the block sums round
    differently: 1e-13 relative on the sketch, the FP64 bar end to end), and a limit below one
    column fails with an error instead of running;
  - a C program compiled with gcc against include/isvd.h runs the reference demo's call
    (demo/cdemo.c:96-105: isvd_Init, isvd_createParam, isvd_dIsvd with "GP_gpu" / "GR" / "KN" /
    "GR_gpu", dista 'R', ordera 'C', seed 0, gathered outputs, isvd_Finalize) on the reference's
    data/a.mtx, with and without a memory limit; the seeded Omega is the Philox stream of
    oracle/philox.py, so the oracle reproduces the run: sigma to 1e-10, subspaces to 1e-8."""
import contextlib
import ctypes
import os
import subprocess

import numpy as np
import pytest

import oracle as O
from conftest import ROOT, lowrank_matrix, principal_angles
from isvd_amd import capi, pipeline
from oracle import philox

pytestmark = pytest.mark.gpu


@contextlib.contextmanager
def memory_limit(nbytes):
    v = ctypes.c_size_t.in_dll(capi.load(), "isvd_gpu_memory_limit")
    old = v.value
    v.value = int(nbytes)
    try:
        yield
    finally:
        v.value = old


@pytest.mark.parametrize("ordera", ["C", "R"])
def test_sketch_stages_host_a_by_column_blocks(torch_cuda, ordera):
    m, n, l, N = 3000, 2000, 16, 4
    Nl = l * N
    rng = np.random.default_rng(1)
    a = rng.standard_normal((m, n))
    om = rng.standard_normal((n, Nl))
    arr = np.ascontiguousarray(a if ordera == "R" else a.T)
    prm = capi.create_param(m, n, l, 0, N)
    ref = pipeline.sketch(prm, torch_cuda.from_numpy(arr).cuda(), omega=om, ordera=ordera)
    ref = ref.cpu().numpy()
    resident = 8 * Nl * (m + n)
    with memory_limit(resident + 300 * 8 * m):  # 300 columns fit: blocks of 256
        y = pipeline.sketch(prm, arr, omega=om, ordera=ordera)
    assert np.abs(y - ref).max() <= 1e-13 * np.abs(ref).max()
    assert np.abs(y - a @ om).max() <= 1e-12 * np.abs(a @ om).max()
    with memory_limit(resident // 2):
        with pytest.raises(RuntimeError, match="not enough GPU memory"):
            pipeline.sketch(prm, arr, omega=om, ordera=ordera)


@pytest.mark.parametrize("alg_p", ["GR", "TS"])
def test_postprocess_stages_host_a_by_column_blocks(torch_cuda, alg_p):
    m, n, k, p = 2500, 1800, 12, 8
    l = k + p
    a = lowrank_matrix(m, n, 0.9 ** np.arange(n), seed=3)
    qbar, _ = np.linalg.qr(a @ np.random.default_rng(4).standard_normal((n, l)))
    prm = capi.create_param(m, n, k, p, 1)
    s0, u0, v0 = pipeline.postprocess(prm, torch_cuda.from_numpy(a).cuda(),
                                      torch_cuda.from_numpy(qbar).cuda(), alg_p)
    with memory_limit(8 * (l * m + l * n + k * (m + n) + 4 * l * l) + 200 * 8 * m):
        s1, u1, v1 = pipeline.postprocess(prm, a, qbar, alg_p)
    s0 = s0.cpu().numpy()
    assert np.abs(s1[:k] - s0[:k]).max() <= 1e-12 * s0[0]
    assert principal_angles(u1[:, :k], u0[:, :k].cpu().numpy()).max() < 1e-9
    assert principal_angles(v1[:, :k], v0[:, :k].cpu().numpy()).max() < 1e-9


def test_driver_stages_host_a_by_column_blocks(torch_cuda):
    m, n, k, p, N = 4000, 3000, 20, 12, 8
    Nl = (k + p) * N
    a = lowrank_matrix(m, n, 0.9 ** np.arange(n), seed=5)
    om = np.random.default_rng(6).standard_normal((n, Nl))
    prm = capi.create_param(m, n, k, p, N)
    algs = ("GP_gpu", "TS", "KN", "TS_gpu")
    g0 = pipeline.isvd(prm, torch_cuda.from_numpy(a).cuda(), omega=om, algs=algs)
    resident = 8 * (Nl * m + (k + p) * (m + n) + k * (m + n) + Nl * n)
    with memory_limit(resident + 500 * 8 * m):  # blocks of 448 columns
        g1 = pipeline.isvd(prm, a, omega=om, algs=algs)
    s0 = g0["s"].cpu().numpy()
    assert g1["iter"] == g0["iter"]
    assert np.abs(g1["s"][:k] - s0[:k]).max() <= 1e-10 * s0[0]
    assert principal_angles(g1["u"][:, :k], g0["u"][:, :k].cpu().numpy()).max() < 1e-8
    assert principal_angles(g1["v"][:, :k], g0["v"][:, :k].cpu().numpy()).max() < 1e-8


CDEMO = r"""
#include <stdio.h>
#include <stdlib.h>
#include "isvd.h"

/* the call of the reference's demo/cdemo.c:75-105 on an A read from a binary file
   (int64 m, int64 n, then m*n doubles, column-major); writes s, ut, vt */
int main(int argc, char **argv) {
  isvd_Init(&argc, &argv, ISVD_COMM_WORLD);
  FILE *f = fopen(argv[1], "rb");
  long long mn[2];
  if (!f || fread(mn, sizeof(long long), 2, f) != 2) return 2;
  const isvd_int_t m = (isvd_int_t)mn[0], n = (isvd_int_t)mn[1];
  double *a = isvd_dMalloc((size_t)m * n);
  if (fread(a, sizeof(double), (size_t)m * n, f) != (size_t)m * n) return 2;
  fclose(f);
  if (argc > 3) isvd_gpu_memory_limit = (size_t)strtoull(argv[3], NULL, 10);
  const isvd_int_t k = 20, p = 12, N = 16;
  const isvd_Param param = isvd_createParam(m, n, k, p, N, 0, ISVD_COMM_WORLD);
  const isvd_int_t l = param.dim_sketch;
  double *s = isvd_dMalloc(l), *ut = isvd_dMalloc((size_t)l * m), *vt = isvd_dMalloc((size_t)l * n);
  double time[4];
  isvd_dIsvd(param, "GP_gpu", "GR", "KN", "GR_gpu", NULL, NULL, NULL, NULL, time, stdout, 'R',
             'C', a, m, s, ut, l, vt, l, 0, 0, 0);
  if (isvd_GetLastError()[0]) {
    fprintf(stderr, "isvd error: %s\n", isvd_GetLastError());
    return 1;
  }
  printf("Average total computing time:   %8.6f seconds.\n", time[0] + time[1] + time[2] + time[3]);
  FILE *o = fopen(argv[2], "wb");
  fwrite(s, sizeof(double), l, o);
  fwrite(ut, sizeof(double), (size_t)l * m, o);
  fwrite(vt, sizeof(double), (size_t)l * n, o);
  fclose(o);
  isvd_Free(a);
  isvd_Free(s);
  isvd_Free(ut);
  isvd_Free(vt);
  isvd_Finalize();
  return 0;
}
"""


@pytest.fixture(scope="module")
def cdemo(tmp_path_factory):
    d = tmp_path_factory.mktemp("cdemo")
    src, exe = d / "cdemo.c", d / "cdemo"
    src.write_text(CDEMO)
    lib = capi._LIB_PATH
    subprocess.run(["gcc", "-std=c99", "-O1", "-I", os.path.join(ROOT, "include"), str(src), lib,
                    "-o", str(exe), f"-Wl,-rpath,{os.path.dirname(lib)}"], check=True)
    return d, exe


@pytest.mark.parametrize("limit", [None, 8 * (101 * 512 + 32 * (101 + 1001) + 20 * 1102 +
                                              512 * 1001) + 300 * 8 * 101])
def test_compiled_c_caller_runs_the_reference_demo(torch_cuda, golden, cdemo, limit):
    d, exe = cdemo
    a = golden["a"]  # data/a.mtx, 101 x 1001
    m, n = a.shape
    k, l, N = 20, 32, 16
    inp, out = d / "a.bin", d / f"out_{limit}.bin"
    with open(inp, "wb") as f:
        f.write(np.array([m, n], dtype=np.int64).tobytes())
        f.write(np.asfortranarray(a).tobytes(order="F"))
    args = [str(exe), str(inp), str(out)] + ([str(limit)] if limit else [])
    r = subprocess.run(args, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "Using Double Precision GPU (MI355X) Gaussian Projection Sketching." in r.stdout
    raw = np.fromfile(out, dtype=np.float64)
    s = raw[:l]
    ut = raw[l:l + l * m].reshape(m, l)      # U^T, l x m column-major = U row-major
    vt = raw[l + l * m:].reshape(n, l)
    ref = O.isvd(a, philox.omega(0, n, l * N), k, N, "GR", "KN", "GR")
    assert np.abs(s[:k] - ref["s"][:k]).max() <= 1e-10 * ref["s"][0]
    assert principal_angles(ut[:, :k], ref["u"]).max() < 1e-8
    assert principal_angles(vt[:, :k], ref["v"]).max() < 1e-8
