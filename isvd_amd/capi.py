"""ctypes binding of libisvd_mi355x.so (the C ABI declared in include/isvd.h).

The functions keep the reference's argument order (emfomy/isvd v1.3.2,
src/include/c/isvd/core/@x@_stage.h:20-77, @x@_driver.h:21-29).  Matrix arguments are raw
addresses: a numpy array (host memory, staged through HBM by the library) or a device
address (e.g. ``torch.Tensor.data_ptr()`` of a CUDA tensor, used in place).

The "b" family (mixed precision) takes A and Omega as BF16 bit patterns (numpy uint16 or
torch.bfloat16) and every other array as FP32.  The "e" family takes A as OCP e4m3 codes
(uint8, row-major, lda8 % 16 == 0) with FP32 power-of-two scales per 128 x 128 block, and a
BF16 copy of A for the postprocessing.

Storage conventions (all column-major, as in the reference):
  a      : A, m x n with lda >= m (ordera 'C'), or the row-major m x n A with lda >= n ('R')
  omegat : Omega^T, Nl x n, ldomegat >= Nl          (= row-major Omega, n x Nl)
  yst    : Y^T,     Nl x m, ldyst >= Nl             (= row-major Y, m x Nl)
  qt     : Qbar^T,  l x m,  ldqt >= l               (= row-major Qbar, m x l)
  ut, vt : U^T (l x m), V^T (l x n), first k rows   (= row-major U, V with l columns)
"""
import ctypes
import os

import numpy as np

_LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib", "libisvd_mi355x.so")

isvd_int_t = ctypes.c_int32
mpi_int_t = ctypes.c_int


class Param(ctypes.Structure):
    """isvd_Param (src/include/c/isvd/core/param.h:27-92): same fields, same order."""
    _fields_ = [("mpi_comm", ctypes.c_int), ("mpi_size", mpi_int_t), ("mpi_root", mpi_int_t),
                ("mpi_rank", mpi_int_t), ("nrow", isvd_int_t), ("ncol", isvd_int_t),
                ("nrow_proc", isvd_int_t), ("ncol_proc", isvd_int_t),
                ("nrow_each", isvd_int_t), ("ncol_each", isvd_int_t),
                ("nrow_total", isvd_int_t), ("ncol_total", isvd_int_t),
                ("rowidxbegin", isvd_int_t), ("rowidxend", isvd_int_t),
                ("colidxbegin", isvd_int_t), ("colidxend", isvd_int_t),
                ("rank", isvd_int_t), ("over_rank", isvd_int_t), ("dim_sketch", isvd_int_t),
                ("dim_sketch_total", isvd_int_t), ("num_sketch", isvd_int_t)]


_lib = None
VP = ctypes.c_void_p
I = isvd_int_t
CH = ctypes.c_char


def _sig(fn, argtypes, restype=None):
    fn.argtypes = argtypes
    fn.restype = restype


def load(path=None):
    """Load the library (raises if it has not been built: there is no fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    path = path or _LIB_PATH
    if not os.path.exists(path):
        raise RuntimeError(f"{path} not found: build it with `python -m isvd_amd.build`")
    lib = ctypes.CDLL(path)
    _sig(lib.isvd_createParam, [I, I, I, I, I, mpi_int_t, ctypes.c_int], Param)
    _sig(lib.isvd_GetLastError, [], ctypes.c_char_p)
    _sig(lib.isvd_GetErrorCount, [], ctypes.c_int)
    _sig(lib.isvd_Init_gpu, [VP, VP, ctypes.c_int])
    _sig(lib.isvd_Finalize_gpu, [])
    _sig(lib.isvd_PrintEnvironment_gpu, [ctypes.c_int])
    for x in "sdb":
        sketch = [Param, VP, I, VP, I, CH, CH, VP, I, VP, I, I, mpi_int_t]
        _sig(getattr(lib, f"isvd_{x}SketchGaussianProjection"), sketch)
        if x != "b":
            _sig(getattr(lib, f"isvd_{x}SketchGaussianProjection_gpu"), sketch)
        _sig(getattr(lib, f"isvd_{x}SketchGaussianProjectionOmega"),
             [Param, CH, CH, VP, I, VP, I, VP, I])
        orth = [Param, VP, I, VP, I, VP, I]
        _sig(getattr(lib, f"isvd_{x}OrthogonalizeTallSkinnyQr"), orth)
        _sig(getattr(lib, f"isvd_{x}OrthogonalizeGramian"), orth)
        integ = [Param, VP, I, VP, I, VP, I, VP, I]
        _sig(getattr(lib, f"isvd_{x}IntegrateKolmogorovNagumo"), integ)
        _sig(getattr(lib, f"isvd_{x}IntegrateWenYin"), integ)
        post = [Param, VP, I, VP, I, CH, CH, VP, I, VP, I, VP, VP, I, VP, I, mpi_int_t, mpi_int_t]
        names = ("PostprocessTallSkinnyQr", "PostprocessGramian")
        if x != "b":
            names += ("PostprocessTallSkinnyQr_gpu", "PostprocessGramian_gpu")
        for nm in names:
            _sig(getattr(lib, f"isvd_{x}{nm}"), post)
        drv = [Param, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p,
               VP, VP, VP, VP, VP, VP, CH, CH, VP, I]
        _sig(getattr(lib, f"isvd_{x}Isvd"), drv + [VP, VP, I, VP, I, I, mpi_int_t, mpi_int_t])
        _sig(getattr(lib, f"isvd_{x}IsvdOmega"),
             drv + [VP, I, VP, VP, I, VP, I, mpi_int_t, mpi_int_t])
    for x in "sd":
        _sig(getattr(lib, f"isvd_{x}IntegrateHierarchicalReduction"),
             [Param, VP, I, VP, I, VP, I, VP, I])
        post = [Param, VP, I, VP, I, CH, CH, VP, I, VP, I, VP, VP, I, VP, I, mpi_int_t, mpi_int_t]
        _sig(getattr(lib, f"isvd_{x}PostprocessSymmetric"), post)
        _sig(getattr(lib, f"isvd_{x}PostprocessSymmetric_gpu"), post)
        _sig(getattr(lib, f"isvd_{x}IsvdBackward"),
             [Param, ctypes.c_char_p, CH, CH, VP, I, VP, I, VP, VP, I, VP, I, VP, I])
    _sig(lib.isvd_dPostprocessGramianNoAlloc,
         [Param, VP, I, VP, I, CH, CH, VP, I, VP, I, VP, VP, I, VP, I, mpi_int_t, mpi_int_t])
    # e family: FP8 (e4m3) sketch
    _sig(lib.isvd_eQuantize, [CH, I, I, VP, I, VP, I, VP, I])
    _sig(lib.isvd_eSketchGaussianProjection,
         [Param, VP, I, VP, I, CH, VP, I, VP, I, VP, I, I, mpi_int_t])
    _sig(lib.isvd_eSketchGaussianProjectionOmega, [Param, CH, VP, I, VP, I, VP, I, VP, I])
    _sig(lib.isvd_eIsvd,
         [Param, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p,
          VP, VP, VP, VP, VP, VP, CH, VP, I, VP, I, CH, VP, I, VP, VP, I, VP, I, I,
          mpi_int_t, mpi_int_t])
    _lib = lib
    return lib


def last_error():
    return load().isvd_GetLastError().decode()


def create_param(m, n, k, p, N, root=0):
    """isvd_createParam (P = 1)."""
    return load().isvd_createParam(m, n, k, p, N, root, 0)


def addr(x):
    """Address of a numpy array (host) or of a torch CUDA tensor (device)."""
    if x is None:
        return None
    if isinstance(x, np.ndarray):
        return x.ctypes.data
    if hasattr(x, "data_ptr"):
        return x.data_ptr()
    if isinstance(x, int):
        return x
    raise TypeError(f"unsupported array type {type(x)}")


def _chr(c):
    return c.encode() if isinstance(c, str) else c


def prefix(dtype):
    """Precision family of the entry points: 'd' (FP64), 's' (FP32) or 'b' (BF16 A and Omega,
    given as uint16 bit patterns or torch.bfloat16; FP32 everywhere else)."""
    name = str(dtype)
    if name in ("torch.bfloat16", "bfloat16") or name.endswith("uint16"):
        return "b"
    if name in ("torch.float64", "float64") or name.endswith("float64'>"):
        return "d"
    try:
        if np.dtype(dtype) == np.float64:
            return "d"
        if np.dtype(dtype) == np.uint16:
            return "b"
    except TypeError:
        pass
    return "s"


class ArgPack:
    """argv/argc/retv/retc arrays for the driver (4 stages)."""

    def __init__(self, ctype, argv=None, retc=(0, 0, 2, 0)):
        self.ctype = ctype
        argv = argv or [None, None, None, None]
        self._argv_arrays = [(ctype * len(a))(*a) if a else None for a in argv]
        self.argv = (VP * 4)(*[ctypes.addressof(a) if a else None for a in self._argv_arrays])
        self.argc = (isvd_int_t * 4)(*[len(a) if a else 0 for a in argv])
        self._retv_arrays = [(ctype * max(c, 1))() for c in retc]
        self.retv = (VP * 4)(*[ctypes.addressof(a) for a in self._retv_arrays])
        self.retc = (isvd_int_t * 4)(*retc)
        self.time = (ctypes.c_double * 4)()

    def ret(self, stage):
        return list(self._retv_arrays[stage])
