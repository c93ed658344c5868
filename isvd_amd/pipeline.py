"""Array-level conveniences over the C ABI (device tensors or host numpy arrays).

Row-major views of the reference's column-major buffers are used throughout: Y is m x Nl,
Qbar m x l, U m x l, V n x l, Omega n x Nl (see capi's module docstring).  A is passed either
row-major m x n (ordera='R', lda = n) or as the row-major n x m array A^T (ordera='C', lda = m).
Every call goes through libisvd_mi355x.so; nothing here computes.
"""
import ctypes

import numpy as np

from . import capi


def _torch():
    import torch
    return torch


def _is_torch(x):
    return hasattr(x, "data_ptr") and not isinstance(x, np.ndarray)


def _family(x, mixed=False):
    """(entry-point prefix, numpy dtype of the FP arrays) for an input array."""
    if mixed:
        return "b", np.float32
    x_ = capi.prefix(x.dtype)
    return x_, (np.float64 if x_ == "d" else np.float32)


def _dtype(x):
    return _family(x)[1]


def _empty_like_kind(ref, shape, dtype):
    if _is_torch(ref):
        torch = _torch()
        tdt = torch.float64 if dtype == np.float64 else torch.float32
        return torch.empty(shape, dtype=tdt, device=ref.device)
    return np.empty(shape, dtype=dtype)


def _ready(*arrays):
    """Wait for the caller's stream to finish producing device inputs: the library runs on its
    own stream, which is ordered after the legacy default stream but not after other streams."""
    for x in arrays:
        if _is_torch(x) and x.is_cuda:
            _torch().cuda.current_stream(x.device).synchronize()
            return


def _check(fn_name):
    err = capi.last_error()
    if err:
        raise RuntimeError(f"{fn_name}: {err}")


def _a_ld(a, ordera):
    # ordera 'R': a is row-major m x n (lda = n); 'C': a is A^T row-major n x m (lda = m)
    return a.shape[1]


def sketch(param, a, omega=None, seed=None, ordera="R", dista="R"):
    """Y = A Omega.  With ``omega`` (n x Nl row-major) uses the caller's draw, else ``seed``."""
    lib = capi.load()
    _ready(a, omega)
    x, dt = _family(a)
    m, Nl = param.nrow, param.dim_sketch_total
    y = _empty_like_kind(a, (m, Nl), dt)
    if omega is not None:
        getattr(lib, f"isvd_{x}SketchGaussianProjectionOmega")(
            param, capi._chr(dista), capi._chr(ordera), capi.addr(a), _a_ld(a, ordera),
            capi.addr(omega), omega.shape[1], capi.addr(y), Nl)
        _check("sketch")
    else:
        getattr(lib, f"isvd_{x}SketchGaussianProjection")(
            param, None, 0, None, 0, capi._chr(dista), capi._chr(ordera), capi.addr(a),
            _a_ld(a, ordera), capi.addr(y), Nl, int(seed or 0), 0)
        _check("sketch")
    return y


def orthogonalize(param, y, alg="GR", mixed=False):
    """In-place orthogonalization of every sketch block of Y (m x Nl); ``mixed`` selects the
    b family (FP32 Y, Gramian twice)."""
    lib = capi.load()
    _ready(y)
    x, _ = _family(y, mixed)
    name = {"GR": "OrthogonalizeGramian", "TS": "OrthogonalizeTallSkinnyQr"}[alg.upper()]
    getattr(lib, f"isvd_{x}{name}")(param, None, 0, None, 0, capi.addr(y), y.shape[1])
    _check(name)
    return y


def integrate(param, qs, alg="KN", argv=None, mixed=False):
    """Returns (Qbar m x l, iter, err_or_mu); ``mixed`` selects the b family (FP32 bases,
    split-BF16 Gram)."""
    lib = capi.load()
    _ready(qs)
    x, dt = _family(qs, mixed)
    ct = ctypes.c_double if dt == np.float64 else ctypes.c_float
    l = param.dim_sketch
    qbar = _empty_like_kind(qs, (param.nrow, l), dt)
    argv = list(argv or [])
    av = (ct * max(len(argv), 1))(*argv)
    rv = (ct * 2)()
    name = {"KN": "IntegrateKolmogorovNagumo", "WY": "IntegrateWenYin",
            "HR": "IntegrateHierarchicalReduction"}[alg.upper()]
    if alg.upper() == "HR":  # destroys its input, as the reference's does
        getattr(lib, f"isvd_{x}{name}")(param, None, 0, None, 0, capi.addr(qs), qs.shape[1],
                                        capi.addr(qbar), l)
        _check(name)
        return qbar, 0, 0.0
    getattr(lib, f"isvd_{x}{name}")(param, ctypes.addressof(av), len(argv),
                                    ctypes.addressof(rv), 2, capi.addr(qs), qs.shape[1],
                                    capi.addr(qbar), l)
    _check(name)
    return qbar, int(rv[0]), float(rv[1])


def postprocess(param, a, qbar, alg="GR", ordera="R", dista="R", want_u=True, want_v=True):
    """Returns (s[l], U m x l or None, V n x l or None); the first k columns of U/V are valid."""
    lib = capi.load()
    _ready(a, qbar)
    x, dt = _family(a)
    l = param.dim_sketch
    s = _empty_like_kind(a, (l,), dt)
    u = _empty_like_kind(a, (param.nrow, l), dt) if want_u else None
    v = _empty_like_kind(a, (param.ncol, l), dt) if want_v else None
    name = {"GR": "PostprocessGramian", "TS": "PostprocessTallSkinnyQr",
            "SY": "PostprocessSymmetric"}[alg.upper()]
    getattr(lib, f"isvd_{x}{name}")(
        param, None, 0, None, 0, capi._chr(dista), capi._chr(ordera), capi.addr(a),
        _a_ld(a, ordera), capi.addr(qbar), qbar.shape[1], capi.addr(s), capi.addr(u), l,
        capi.addr(v), l, 0 if want_u else -2, 0 if want_v else -2)
    _check(name)
    return s, u, v


def isvd(param, a, omega=None, seed=0, algs=("GP", "GR", "KN", "GR"), ordera="R", dista="R",
         argv_i=None, want_u=True, want_v=True):
    """Full driver (isvd_?Isvd / isvd_?IsvdOmega).  Returns dict(s, u, v, iter, err, time).
    A BF16 ``a`` (torch.bfloat16 or uint16 bits) selects the mixed-precision b family."""
    lib = capi.load()
    _ready(a, omega)
    x, dt = _family(a)
    ct = ctypes.c_double if dt == np.float64 else ctypes.c_float
    l = param.dim_sketch
    s = _empty_like_kind(a, (l,), dt)
    u = _empty_like_kind(a, (param.nrow, l), dt) if want_u else None
    v = _empty_like_kind(a, (param.ncol, l), dt) if want_v else None
    pack = capi.ArgPack(ct, [None, None, list(argv_i) if argv_i else None, None])
    common = [param, *(al.encode() for al in algs), pack.argv, pack.argc, pack.retv, pack.retc,
              pack.time, None, capi._chr(dista), capi._chr(ordera), capi.addr(a),
              _a_ld(a, ordera)]
    tail = [capi.addr(s), capi.addr(u), l, capi.addr(v), l]
    roots = [0 if want_u else -2, 0 if want_v else -2]
    if omega is not None:
        getattr(lib, f"isvd_{x}IsvdOmega")(*common, capi.addr(omega), omega.shape[1], *tail,
                                            *roots)
    else:
        getattr(lib, f"isvd_{x}Isvd")(*common, *tail, int(seed), *roots)
    _check("isvd")
    it, err = pack.ret(2)[:2]
    return {"s": s, "u": u, "v": v, "iter": int(it), "err": float(err),
            "time": list(pack.time)}


def isvd_backward(param, a, qbar, gs=None, gu=None, gv=None, alg_p="TS", ordera="R", dista="R"):
    """dL/dA of the rank-k result (isvd_?IsvdBackward).  ``qbar`` is the integration output
    (m x l), ``gs`` (k), ``gu`` (m x k) and ``gv`` (n x k) the gradients of the outputs (None for
    zero).  Returns dL/dA with the shape and storage of ``a``."""
    lib = capi.load()
    x, dt = _family(a)
    if x == "b":
        raise ValueError("isvd_backward: the mixed-precision family has no adjoint")
    k = param.rank
    da = _empty_like_kind(a, tuple(a.shape), dt)

    def cont(t):
        if t is None:
            return None
        return t.contiguous() if _is_torch(t) else np.ascontiguousarray(t)

    # the contiguous copies are queued on the caller's stream: wait for them too
    gs, gu, gv = cont(gs), cont(gu), cont(gv)
    _ready(a, qbar, gs, gu, gv)
    getattr(lib, f"isvd_{x}IsvdBackward")(
        param, alg_p.encode(), capi._chr(dista), capi._chr(ordera), capi.addr(a),
        _a_ld(a, ordera), capi.addr(qbar), qbar.shape[1], capi.addr(gs), capi.addr(gu), k,
        capi.addr(gv), k, capi.addr(da), _a_ld(a, ordera))
    _check("isvd_backward")
    return da


# ---- e family: FP8 (OCP e4m3) sketch ---------------------------------------------------------

def _u8_empty_like_kind(ref, shape):
    if _is_torch(ref):
        return _torch().empty(shape, dtype=_torch().uint8, device=ref.device)
    return np.empty(shape, dtype=np.uint8)


def quantize_e4m3(a, ordera="R"):
    """(a8, scale) of an FP32 matrix (isvd_eQuantize).  ``a`` is row-major m x n (ordera 'R')
    or the row-major n x m array A^T ('C').  a8: m x lda8 uint8 codes, lda8 = roundup(n, 16);
    scale: ceil(m/128) x ceil(n/128) FP32 powers of two."""
    lib = capi.load()
    _ready(a)
    m, n = (a.shape[0], a.shape[1]) if ordera == "R" else (a.shape[1], a.shape[0])
    lda8 = (n + 15) // 16 * 16
    nbi, nbj = (m + 127) // 128, (n + 127) // 128
    a8 = _u8_empty_like_kind(a, (m, lda8))
    scale = _empty_like_kind(a, (nbi, nbj), np.float32)
    lib.isvd_eQuantize(capi._chr(ordera), m, n, capi.addr(a), a.shape[1], capi.addr(a8), lda8,
                       capi.addr(scale), nbj)
    _check("quantize_e4m3")
    return a8, scale


def sketch_e(param, a8, scale, omega8=None, seed=None, dista="R"):
    """Y (m x Nl, FP32) = A Omega from the e4m3 A (a8 m x lda8, scale per 128 x 128 block).
    ``omega8`` (Nl x ldo uint8: row c holds column c of Omega, n codes) or ``seed``."""
    lib = capi.load()
    _ready(a8, scale, omega8)
    m, Nl = param.nrow, param.dim_sketch_total
    y = _empty_like_kind(a8, (m, Nl), np.float32)
    if omega8 is not None:
        lib.isvd_eSketchGaussianProjectionOmega(
            param, capi._chr(dista), capi.addr(a8), a8.shape[1], capi.addr(scale),
            scale.shape[1], capi.addr(omega8), omega8.shape[1], capi.addr(y), Nl)
    else:
        lib.isvd_eSketchGaussianProjection(
            param, None, 0, None, 0, capi._chr(dista), capi.addr(a8), a8.shape[1],
            capi.addr(scale), scale.shape[1], capi.addr(y), Nl, int(seed or 0), 0)
    _check("sketch_e")
    return y


def isvd_e(param, a8, scale, a_bf16, seed=0, algs=("GP", "GR", "KN", "GR"), ordera="R",
           dista="R", argv_i=None, want_u=True, want_v=True):
    """isvd_eIsvd: FP8 sketch from (a8, scale), the other stages on the BF16 copy ``a_bf16``
    (laid out as for ``isvd``).  Returns dict(s, u, v, iter, err, time), FP32."""
    lib = capi.load()
    _ready(a8, scale, a_bf16)
    l = param.dim_sketch
    s = _empty_like_kind(a8, (l,), np.float32)
    u = _empty_like_kind(a8, (param.nrow, l), np.float32) if want_u else None
    v = _empty_like_kind(a8, (param.ncol, l), np.float32) if want_v else None
    pack = capi.ArgPack(ctypes.c_float, [None, None, list(argv_i) if argv_i else None, None])
    lib.isvd_eIsvd(param, *(al.encode() for al in algs), pack.argv, pack.argc, pack.retv,
                   pack.retc, pack.time, None, capi._chr(dista), capi.addr(a8), a8.shape[1],
                   capi.addr(scale), scale.shape[1], capi._chr(ordera), capi.addr(a_bf16),
                   _a_ld(a_bf16, ordera), capi.addr(s), capi.addr(u), l, capi.addr(v), l,
                   int(seed), 0 if want_u else -2, 0 if want_v else -2)
    _check("isvd_e")
    it, err = pack.ret(2)[:2]
    return {"s": s, "u": u, "v": v, "iter": int(it), "err": float(err),
            "time": list(pack.time)}
