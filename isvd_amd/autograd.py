"""torch.autograd.Function over the C ABI.

forward  : the four iSVD stages (isvd_?SketchGaussianProjectionOmega, isvd_?Orthogonalize*,
           isvd_?Integrate*, isvd_?Postprocess*), keeping Qbar from the integration stage;
backward : isvd_?IsvdBackward, which recomputes U, S, V from Qbar with the same postprocessing
           and applies the truncated-SVD adjoint (see include/isvd.h).

A is a row-major m x n CUDA tensor (FP64 or FP32), Omega a row-major n x N*l tensor of the same
dtype.  Outputs: s (k), U (m x k), V (n x k).
"""
import torch

from . import capi, pipeline


class ISVDFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, a, omega, k, p, N, algs):
        m, n = a.shape
        prm = capi.create_param(m, n, k, p, N)
        y = pipeline.sketch(prm, a, omega=omega)
        pipeline.orthogonalize(prm, y, algs[1])
        qbar, _, _ = pipeline.integrate(prm, y, algs[2])
        del y
        s, u, v = pipeline.postprocess(prm, a, qbar, algs[3])
        ctx.save_for_backward(a, qbar)
        ctx.cfg = (k, p, N, algs[3])
        return s[:k].clone(), u[:, :k].contiguous(), v[:, :k].contiguous()

    @staticmethod
    def backward(ctx, gs, gu, gv):
        a, qbar = ctx.saved_tensors
        k, p, N, alg_p = ctx.cfg
        prm = capi.create_param(a.shape[0], a.shape[1], k, p, N)
        da = pipeline.isvd_backward(prm, a, qbar, gs, gu, gv, alg_p)
        return da, None, None, None, None, None


def isvd(a, omega, k, p, N, algs=("GP", "TS", "KN", "TS")):
    """Differentiable rank-k iSVD: returns (s, U, V) with gradients flowing to ``a``."""
    return ISVDFunction.apply(a, omega, k, p, N, tuple(algs))
