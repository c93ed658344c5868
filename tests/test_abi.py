"""C-ABI checks that need no GPU: the library loads, exports every function include/isvd.h
declares, and the host-side logic (parameters, query mode, argument validation, selector
parsing) behaves like the reference's."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

from conftest import ROOT
from isvd_amd import capi

HEADER = os.path.join(ROOT, "include", "isvd.h")


def declared_functions():
    out = subprocess.run(["gcc", "-E", "-P", "-x", "c", HEADER], capture_output=True, text=True,
                         check=True).stdout
    names = set(re.findall(r"\b(isvd_\w+)\s*\(", out))
    # static inline wrappers are defined in the header itself
    inline = set(re.findall(r"static inline \w+ (isvd_\w+)\s*\(", out))
    return names - inline


def exported_functions():
    path = capi._LIB_PATH
    out = subprocess.run(["nm", "-D", "--defined-only", path], capture_output=True, text=True,
                         check=True).stdout
    return {ln.split()[-1] for ln in out.splitlines() if " T " in ln}


def test_library_exports_every_declared_symbol():
    capi.load()
    decl = declared_functions()
    assert len(decl) >= 48
    missing = decl - exported_functions()
    assert not missing, f"declared but not exported: {sorted(missing)}"
    # the memory-limit global of gpu/env.h:19
    assert ctypes.c_size_t.in_dll(capi.load(), "isvd_gpu_memory_limit").value == 0


def test_header_compiles_as_c_and_cxx(tmp_path):
    src = tmp_path / "use.c"
    src.write_text('#include "isvd.h"\nint main(void){ isvd_Param p = isvd_createParam('
                   '10, 5, 2, 1, 3, 0, ISVD_COMM_WORLD); return p.dim_sketch_total == 9 ? 0 : 1; }\n')
    lib = capi._LIB_PATH
    for cc in (["gcc", "-std=c99", "-x", "c"], ["g++", "-x", "c++"]):
        exe = tmp_path / "use"
        subprocess.run([*cc, "-I", os.path.join(ROOT, "include"), str(src), "-x", "none", lib,
                        "-o", str(exe), f"-Wl,-rpath,{os.path.dirname(lib)}"], check=True)
        assert subprocess.run([str(exe)]).returncode == 0


def test_create_param_partition_p1():
    # src/lib/libisvd/core/param.c:46-80 with P = 1; m > n accepted (the reference asserts m <= n)
    p = capi.create_param(2000, 1000, 20, 12, 16)
    assert (p.mpi_size, p.mpi_rank, p.nrow, p.ncol) == (1, 0, 2000, 1000)
    assert (p.nrow_proc, p.ncol_proc, p.nrow_each, p.ncol_each) == (2000, 1000, 2000, 1000)
    assert (p.nrow_total, p.ncol_total) == (2000, 1000)
    assert (p.rowidxbegin, p.rowidxend, p.colidxbegin, p.colidxend) == (0, 2000, 0, 1000)
    assert (p.rank, p.over_rank, p.dim_sketch, p.dim_sketch_total, p.num_sketch) == \
        (20, 12, 32, 512, 16)


@pytest.mark.parametrize("x,ct", [("d", ctypes.c_double), ("s", ctypes.c_float)])
def test_query_mode_returns_defaults(x, ct):
    # @x@_integrate_kolmogorov_nagumo.c:56-61 and @x@_integrate_wen_yin.c:70-82
    lib = capi.load()
    p = capi.create_param(100, 50, 5, 3, 4)
    rv = (ct * 9)()
    getattr(lib, f"isvd_{x}IntegrateKolmogorovNagumo")(p, None, -1, ctypes.addressof(rv), 2,
                                                         None, 1, None, 1)
    assert list(rv)[:2] == pytest.approx([256, 1e-4])
    getattr(lib, f"isvd_{x}IntegrateWenYin")(p, None, -1, ctypes.addressof(rv), 9, None, 1,
                                               None, 1)
    assert list(rv) == pytest.approx([256, 1e-3, 1.0, 1e14, 1e-14, 20, 0.5, 1e-4, 0.85],
                                     rel=1e-6)


def test_invalid_selector_and_char_print_and_return(capfd):
    # src/lib/libisvd/util/arg.h:40-48, :70-78: print to stderr and return without output
    lib = capi.load()
    p = capi.create_param(100, 50, 5, 3, 4)
    a = np.zeros((100, 50))
    s = np.full(8, 7.0)
    pack = capi.ArgPack(ctypes.c_double)
    lib.isvd_dIsvd(p, b"GP", b"XX", b"KN", b"GR", pack.argv, pack.argc, pack.retv, pack.retc,
                   pack.time, None, b"R", b"C", a.ctypes.data, 100, s.ctypes.data, None, 8, None,
                   8, 0, -2, -2)
    err = capfd.readouterr().err
    assert 'ALG_O ("XX") must be one of the following strings: "TS", "GR"' in err.replace(
        "('XX')", '("XX")')
    assert (s == 7.0).all()
    y = np.zeros((100, 32))
    lib.isvd_dSketchGaussianProjection(p, None, 0, None, 0, b"Q", b"C", a.ctypes.data, 100,
                                       y.ctypes.data, 32, 0, 0)
    assert "DISTA ('Q') must be one of the following characters" in capfd.readouterr().err
    assert capi.last_error() == ""


def test_selectors_are_case_insensitive_and_validated_before_gpu(capfd):
    # a lower-case selector is accepted by the parser; the call then fails on ldut (an
    # argument check done before any device work, so this runs without a GPU)
    lib = capi.load()
    p = capi.create_param(100, 50, 5, 3, 4)
    a = np.zeros((100, 50))
    s = np.zeros(8)
    u = np.zeros((100, 8))
    pack = capi.ArgPack(ctypes.c_double)
    lib.isvd_dIsvd(p, b"gp", b"gr", b"kn", b"gr", pack.argv, pack.argc, pack.retv, pack.retc,
                   pack.time, None, b"r", b"c", a.ctypes.data, 100, s.ctypes.data,
                   u.ctypes.data, 7, None, 8, 0, 0, -2)
    assert "ldut (7) must equal l (8)" in capi.last_error()
    assert "must be one of" not in capfd.readouterr().err


@pytest.mark.parametrize("call", ["orth_ld", "kn_ld", "post_ldvt", "wy_beta", "l_too_big"])
def test_argument_validation_errors(call):
    lib = capi.load()
    p = capi.create_param(300, 200, 5, 3, 4)
    y = np.zeros((300, 32))
    if call == "orth_ld":
        lib.isvd_dOrthogonalizeGramian(p, None, 0, None, 0, y.ctypes.data, 31)
        assert "ldyst (31) must be >= Nl (32)" in capi.last_error()
    elif call == "kn_ld":
        q = np.zeros((300, 8))
        lib.isvd_dIntegrateKolmogorovNagumo(p, None, 0, None, 0, y.ctypes.data, 32,
                                             q.ctypes.data, 7)
        assert "ldqt (7) must be >= l" in capi.last_error()
    elif call == "post_ldvt":
        a = np.zeros((300, 200))
        s = np.zeros(8)
        v = np.zeros((200, 9))
        lib.isvd_dPostprocessGramian(p, None, 0, None, 0, b"R", b"C", a.ctypes.data, 300,
                                     y.ctypes.data, 8, s.ctypes.data, None, 8, v.ctypes.data, 9,
                                     -2, 0)
        assert "ldvt (9) must equal l (8)" in capi.last_error()
    elif call == "wy_beta":
        q = np.zeros((300, 8))
        av = (ctypes.c_double * 7)(256, 1e-3, 1, 1e14, 1e-14, 20, 1.5)
        lib.isvd_dIntegrateWenYin(p, ctypes.addressof(av), 7, None, 0, y.ctypes.data, 32,
                                   q.ctypes.data, 8)
        assert "beta, sigma, eta must lie in (0, 1)" in capi.last_error()
    else:
        p2 = capi.create_param(300, 200, 100, 40, 2)
        y2 = np.zeros((300, 280))
        lib.isvd_dOrthogonalizeTallSkinnyQr(p2, None, 0, None, 0, y2.ctypes.data, 280)
        assert "l = k + p must be <= 128" in capi.last_error()


def test_failure_aborts_when_requested():
    # ISVD_ASSERT_ABORT=1 restores the reference's assert() behaviour (src/lib/libisvd/def.h:51-68)
    code = ("import numpy as np\nfrom isvd_amd import capi\nlib = capi.load()\n"
            "p = capi.create_param(300, 200, 5, 3, 4)\ny = np.zeros((300, 32))\n"
            "lib.isvd_dOrthogonalizeGramian(p, None, 0, None, 0, y.ctypes.data, 31)\n")
    env = dict(os.environ, ISVD_ASSERT_ABORT="1", PYTHONPATH=ROOT)
    r = subprocess.run(["python", "-c", code], env=env, capture_output=True, text=True)
    assert r.returncode != 0
    assert "ldyst (31) must be >= Nl (32)" in r.stderr
