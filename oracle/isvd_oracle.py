"""CPU oracle for the iSVD four-stage pipeline -- TEST INFRASTRUCTURE ONLY.

This module restates, stage by stage, the algorithm of the reference library
emfomy/isvd v1.3.2 (a C99 library over BLAS/LAPACK/MPI) for ONE process (P = 1),
so that the HIP implementation in ``isvd_amd`` can be checked against it.
Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import it, and only to check results against; the product path
never calls it.

Arithmetic boundary.  The reference delegates every dense kernel to BLAS/LAPACK
(``dgemm_``, ``dgesvd_``, ``dsyev_``, ``dgelqf_``/``dorglq_``, ``dgetrf_``/``dgetri_``,
src/include/c/isvd/la/blas/blas3.h:47-51, src/lib/libisvd/la/lapack/*.c) and does
not pin a version (cmake/packages.cmake:76-93).  This restatement calls the same
LAPACK drivers through ``scipy.linalg.lapack`` (netlib LAPACK as bundled in
OpenBLAS 0.3.29, scipy 1.15.3) and numpy's ``@`` (OpenBLAS dgemm).  ``gelqf``/
``orglq`` of an l x n matrix are taken as ``geqrf``/``orgqr`` of its transpose,
which is the same Householder sequence (LAPACK's LQ is defined as the QR of the
transpose, with identical ``dlarfg`` sign conventions).

Layout.  The reference keeps every tall matrix transposed in column-major
storage (``yst`` = Y^T, Nl x m; ``qt`` = Qbar^T; ``ut``/``vt``; see
src/lib/libisvd/core/stage/@x@_sketch_gaussian_projection.c:33-39).  Here those
appear as plain numpy arrays in their natural orientation: Y (m x Nl),
Qbar (m x l), U (m x k), V (n x k), A (m x n).

How each function is pinned (checked by tests/test_oracle.py):
  sketch_gaussian_projection  -- golden ys_gp = a * o_gp (data/ys_gp.mtx) at 1e-8
                                 and brute force on tiny inputs.
  orthogonalize_gramian       -- golden projector sum Q_i Q_i^T of data/qs.mtx from
                                 data/ys.mtx at 1e-8; Q_i^T Q_i = I.
  orthogonalize_tall_skinny_qr-- same golden projector; equals numpy QR up to column
                                 signs.
  integrate_kolmogorov_nagumo -- golden data/q_kn.mtx entrywise at 1e-8 and iter == 81;
                                 N = 1 special case returns Q_0 after 1 iteration.
  integrate_wen_yin           -- golden data/q_wy.mtx entrywise at 1e-8 and iter == 21.
  postprocess_gramian         -- golden data/s.mtx (sigma), u.mtx/v.mtx projectors at
                                 1e-8; sigma equals numpy SVD of Qbar^T A.
  postprocess_tall_skinny_qr  -- same goldens; k < l vectors checked against the
                                 textbook SVD of Qbar^T A (see ``reference_swap``).
  isvd (driver)               -- composition of the above; end-to-end sigma vs numpy
                                 dense SVD on a matrix with a decaying spectrum.
"""
from itertools import count

import numpy as np
from scipy.linalg import lapack as _lp

__all__ = [
    "sketch_gaussian_projection", "orthogonalize_gramian", "orthogonalize_tall_skinny_qr",
    "integrate_kolmogorov_nagumo", "integrate_wen_yin", "integrate_hierarchical_reduction",
    "postprocess_gramian", "postprocess_tall_skinny_qr", "postprocess_symmetric", "isvd",
    "KN_DEFAULTS", "WY_DEFAULTS",
]

# Default stage arguments (src/lib/libisvd/core/stage/@x@_integrate_kolmogorov_nagumo.c:16-17,
# src/lib/libisvd/core/stage/@x@_integrate_wen_yin.c:16-24).
KN_DEFAULTS = {"maxit": 256, "tol": 1e-4}
WY_DEFAULTS = {"maxit": 256, "tol": 1e-3, "tau0": 1.0, "taumax": 1e14, "taumin": 1e-14,
               "taumaxit": 20, "beta": 0.5, "sigma": 1e-4, "eta": 0.85}


def _pfx(x):
    return "s" if x.dtype == np.float32 else "d"


def _gesvd_u(w):
    """gesvd(jobu='O', jobvt='N') of a square matrix: returns (U, s).

    Mirrors isvd_?Gesvd (src/lib/libisvd/la/lapack/gesvd.c:47-57)."""
    f = getattr(_lp, _pfx(w) + "gesvd")
    u, s, _vt, info = f(w, compute_uv=1, full_matrices=0)
    if info != 0:
        raise RuntimeError(f"gesvd info={info}")
    return u, s


def _gesvd_full(a):
    f = getattr(_lp, _pfx(a) + "gesvd")
    u, s, vt, info = f(a, compute_uv=1, full_matrices=0)
    if info != 0:
        raise RuntimeError(f"gesvd info={info}")
    return u, s, vt


def _syev_upper(z):
    """syev(jobz='V', uplo='U'): eigenvalues ascending + eigenvectors.

    Only the upper triangle of ``z`` is referenced, as in isvd_?Syev
    (src/lib/libisvd/la/lapack/syev.c:45-54)."""
    f = getattr(_lp, _pfx(z) + "syev")
    w, v, info = f(z, compute_v=1, lower=0)
    if info != 0:
        raise RuntimeError(f"syev info={info}")
    return w, v


def _sym_from_upper(x):
    return np.triu(x) + np.triu(x, 1).T


def _qr_q(y):
    """Explicit Q of the Householder QR of y (m x l, m >= l) == (orglq of gelqf(y^T))^T."""
    pf = _pfx(y)
    qr, tau, _lwork, info = getattr(_lp, pf + "geqrf")(y)
    if info != 0:
        raise RuntimeError(f"geqrf info={info}")
    r = np.triu(qr[: y.shape[1], :])
    q, _lwork, info = getattr(_lp, pf + "orgqr")(qr, tau)
    if info != 0:
        raise RuntimeError(f"orgqr info={info}")
    return q[:, : y.shape[1]], r


def _geinv(c):
    """getrf + getri (isvd_?Geinv, src/lib/libisvd/la/lapack/geinv.c:53-65)."""
    pf = _pfx(c)
    lu, piv, info = getattr(_lp, pf + "getrf")(c)
    if info != 0:
        raise RuntimeError(f"getrf info={info}")
    inv, info = getattr(_lp, pf + "getri")(lu, piv)
    if info != 0:
        raise RuntimeError(f"getri info={info}")
    return inv


# --------------------------------------------------------------------------------------------
# Sketching
# --------------------------------------------------------------------------------------------
def sketch_gaussian_projection(a, omega):
    """Y = A * Omega (Y^T = Omega^T * A^T).

    Follows sketchBlockRow (src/lib/libisvd/core/stage/@x@_sketch.h:92-146), GEMM at
    :138-139.  The reference draws Omega (n x Nl, row-major, one Gaussian stream, :132)
    itself; here Omega is an input so that both sides use the same host-generated draw.
    """
    return a @ omega


# --------------------------------------------------------------------------------------------
# Orthogonalization
# --------------------------------------------------------------------------------------------
def orthogonalize_gramian(y, num_sketch):
    """Gramian orthogonalization, per sketch i (P = 1).

    W_i = Y_i^T Y_i (@x@_orthogonalize_gramian.c:81-84); gesvd('O','N') gives W_i := U_i,
    s_i (:87-90); s_i := sqrt(s_i) (:91); Q_i = Y_i * (U_i / s_i) (:94-98, Dism side 'R'
    scales column j by alpha/s_j, src/lib/libisvd/la/blas/dism.c:26-31).
    """
    m, nl = y.shape
    l = nl // num_sketch
    q = np.empty_like(y)
    for i in range(num_sketch):
        yi = y[:, i * l:(i + 1) * l]
        w = yi.T @ yi
        u, s = _gesvd_u(w)
        s = np.sqrt(s)
        q[:, i * l:(i + 1) * l] = yi @ (u * (1.0 / s)[None, :])
    return q


def orthogonalize_tall_skinny_qr(y, num_sketch):
    """Tall-skinny QR orthogonalization at P = 1.

    gelqf(Y_i^T) then orglq (src/lib/libisvd/core/stage/@x@_orthogonalize_tall_skinny_qr.c:
    127-143); with one process the butterfly (:147-189) and the final multiply (:195-200)
    do not run.  Asserts m >= l (:76).
    """
    m, nl = y.shape
    l = nl // num_sketch
    if m < l:
        raise ValueError("tall-skinny QR needs m >= l")
    q = np.empty_like(y)
    for i in range(num_sketch):
        q[:, i * l:(i + 1) * l], _ = _qr_q(np.ascontiguousarray(y[:, i * l:(i + 1) * l]))
    return q


# --------------------------------------------------------------------------------------------
# Integration
# --------------------------------------------------------------------------------------------
def integrate_kolmogorov_nagumo(qs, num_sketch, maxit=256, tol=1e-4):
    """Kolmogorov-Nagumo integration; returns (Qbar, iter, error).

    Follows src/lib/libisvd/core/stage/@x@_integrate_kolmogorov_nagumo.c:155-276:
      Qc := Q_0; Bc := Qs^T Qc                                           (:156-160)
      loop: Gc := Qs Bc / N; Bgc := Qs^T Gc; Dc := Bc^T Bc / N;          (:174-181)
            Z := upper(Bc^T Bgc / N) - Dc^T Dc                           (:184-190)
            eig(Z) (syev 'V','U', ascending) -> Z, s                    (:193)
            s := sqrt(1/2 + sqrt(1/4 - s)); C := Z s Z^T; iC := Z s^-1 Z^T (:196-212)
            Fc := C - Dc iC; Q+ := Qc Fc + Gc iC; B+ := Bc Fc + Bgc iC  (:218-226)
            err := ||s - 1||_2; stop if err <= tol; iter == maxit -> iter+1 (:230-241)
    """
    qs = np.asarray(qs)
    l = qs.shape[1] // num_sketch
    n_inv = 1.0 / num_sketch
    qc = qs[:, :l].copy()
    bc = qs.T @ qc
    err = -np.inf
    for it in count(1):
        gc = n_inv * (qs @ bc)
        bgc = qs.T @ gc
        dc = n_inv * (bc.T @ bc)
        z = n_inv * (bc.T @ bgc)
        z = z - dc.T @ dc
        w, zv = _syev_upper(z)
        s = np.sqrt(0.5 + np.sqrt(0.25 - w))
        ss = np.sqrt(s)
        zs = zv * ss[None, :]
        zinvs = zv / ss[None, :]
        c = zs @ zs.T
        cinv = _sym_from_upper(zinvs @ zinvs.T)
        fc = c - dc @ cinv
        qp = qc @ fc + gc @ cinv
        bp = bc @ fc + bgc @ cinv
        err = float(np.linalg.norm(s - 1.0))
        if err <= tol:
            break
        if it == maxit:
            it += 1
            break
        qc, bc = qp, bp
    return qp, it, err


def integrate_wen_yin(qs, num_sketch, maxit=256, tol=1e-3, tau0=1.0, taumax=1e14, taumin=1e-14,
                      taumaxit=20, beta=0.5, sigma=1e-4, eta=0.85):
    """Wen-Yin curvilinear-search integration; returns (Qbar, iter, mu).

    Follows src/lib/libisvd/core/stage/@x@_integrate_wen_yin.c:201-384 (see the comments
    below for the line of each step).  Fgc is taken as the code computes it (the -inv(C)_11
    block, :248-259), not as the comment at :247 says.
    """
    qs = np.asarray(qs)
    l = qs.shape[1] // num_sketch
    n_inv = 1.0 / num_sketch
    eye = np.eye(l, dtype=qs.dtype)
    # init (:201-235)
    qc = qs[:, :l].copy()
    bc = qs.T @ qc
    dc = n_inv * (bc.T @ bc)
    gc = n_inv * (qs @ bc)
    bgc = qs.T @ gc
    dgc = n_inv * (bc.T @ bgc)
    xc = gc - qc @ dc
    mu = float(np.sum(np.abs(np.diag(dgc))) - np.sum(dc * dc))
    taug, zeta = tau0, 1.0
    phi = 0.5 * n_inv * float(np.sum(bc * bc))
    for it in count(1):
        # step size search (:243-282)
        tau, phit = taug, phi
        fc = fgc = bp = None
        for _tauiter in range(1, int(taumaxit) + 1):
            c = np.zeros((2 * l, 2 * l), dtype=qs.dtype)
            c[:l, :l] = 0.5 * dc - eye / tau
            c[l:, :l] = -0.5 * dgc
            c[:l, l:] = 0.5 * eye
            c[l:, l:] = -0.5 * dc - eye / tau
            ci = _geinv(c)
            # [c11; c21] := [c12; c22] * Dc - [c11; c21]  (:264)
            cs1 = ci[:, l:] @ dc - ci[:, :l]
            fgc = cs1[:l, :]
            fc = cs1[l:, :] + eye
            bp = bc @ fc + bgc @ fgc
            phit = 0.5 * n_inv * float(np.sum(bp * bp))
            if phit >= phi + tau * sigma * mu:
                break
            tau *= beta
        # update (:287-307)
        phi = (eta * zeta * phi + phit) / (eta * zeta + 1)
        zeta = eta * zeta + 1
        qp = qc @ fc + gc @ fgc
        dc = n_inv * (bp.T @ bp)
        gc = n_inv * (qs @ bp)
        bgc = qs.T @ gc
        dgc = n_inv * (bp.T @ bgc)
        mu = float(np.sum(np.abs(np.diag(dgc))) - np.sum(dc * dc))
        # convergence (:312-319)
        if mu < tol * tol:
            break
        if it == maxit:
            it += 1
            break
        # BB step (:324-348)
        xp = gc - qp @ dc
        d1 = qc - qp
        d2 = xc - xp
        if it % 2:
            t0, t1 = float(np.sum(d1 * d1)), float(np.sum(d1 * d2))
        else:
            t0, t1 = float(np.sum(d2 * d1)), float(np.sum(d2 * d2))
        taug = abs(t0 / t1)
        taug = min(max(taug, taumin), taumax)
        qc, xc, bc = qp, xp, bp
    return qp, it, mu


def integrate_hierarchical_reduction(qs, num_sketch):
    """Hierarchical-reduction integration; returns Qbar (m x l).

    src/lib/libisvd/core/stage/@x@_integrate_hierarchical_reduction.c:87-129: for
    h2 = N, ceil(h2/2), ... while h2 > 1, with h = floor(h2/2) and every i < h:
    B_i = Q_i^T Q_{i+h}; gesvd('O','S'): B_i = W S T^T; Q_i := (Q_i W + Q_{i+h} T) / sqrt(2(I+S))
    (column scaling); Qbar = Q_0.  For odd h2 the block h2 - 1 takes no part in that level and
    the next level reads blocks 0 .. ceil(h2/2) - 1, as the reference's loop does.
    """
    m, nl = qs.shape
    l = nl // num_sketch
    q = [qs[:, i * l:(i + 1) * l].copy() for i in range(num_sketch)]
    h2 = num_sketch
    while h2 > 1:
        h = h2 // 2
        for i in range(h):
            b = q[i].T @ q[i + h]
            w, s, tt = _gesvd_full(b)
            q[i] = (q[i] @ w + q[i + h] @ tt.T) / np.sqrt(2.0 * (1.0 + s))[None, :]
        h2 = (h2 + 1) // 2
    return q[0]


# --------------------------------------------------------------------------------------------
# Postprocessing
# --------------------------------------------------------------------------------------------
def postprocess_gramian(a, qbar, rank, want_u=True, want_v=True):
    """Gramian postprocessing; returns (s[l], U[m,k] or None, V[n,k] or None).

    Z^T = Qbar^T A (projectBlockRow, src/lib/libisvd/core/stage/@x@_postprocess.h:156-157);
    W = Z^T Z (@x@_postprocess_gramian.c:145); gesvd('O','N') -> W, s; s := sqrt(s)
    (:149-151); U = Qbar W[:, :k] (:157-158); V = Z W[:, :k] / s[:k] (:172-174).
    """
    zt = qbar.T @ a
    w = zt @ zt.T
    u_w, s = _gesvd_u(w)
    s = np.sqrt(s)
    u = qbar @ u_w[:, :rank] if want_u else None
    v = zt.T @ (u_w[:, :rank] * (1.0 / s[:rank])[None, :]) if want_v else None
    return s, u, v


def postprocess_tall_skinny_qr(a, qbar, rank, want_u=True, want_v=True, reference_swap=False):
    """Tall-skinny-QR postprocessing at P = 1; returns (s[l], U[m,k], V[n,k]).

    Z^T = Qbar^T A (@x@_postprocess.h:156-157); gelqf(Z^T) = L Q_lq (:199), orglq (:206);
    gesvd(L) with jobu='S' -> T = U_L and jobvt='O' -> W = V_L^T (:252-258).

    With ``reference_swap=False`` (default) the singular vectors are the mathematically
    correct U = Qbar U_L[:, :k], V = Q_lq^T V_L[:, :k].  The reference computes
    U = Qbar V_L[:, :k] (:265) and V^T = (U_L Q_lq)[:k] (:284), which agrees with the
    correct result only in projector form and only when k == l; ``reference_swap=True``
    reproduces that defect exactly.
    """
    zt = qbar.T @ a
    l = zt.shape[0]
    if zt.shape[1] < l:
        raise ValueError("tall-skinny QR postprocessing needs n >= l")
    q_qr, r = _qr_q(np.ascontiguousarray(zt.T))  # Z = Q_qr R  <=>  Z^T = R^T Q_qr^T = L Q_lq
    lmat = r.T
    u_l, s, vt_l = _gesvd_full(lmat)
    if reference_swap:
        u = qbar @ vt_l[:rank, :].T if want_u else None
        v = (u_l @ q_qr.T)[:rank, :].T if want_v else None
    else:
        u = qbar @ u_l[:, :rank] if want_u else None
        v = q_qr @ vt_l[:rank, :].T if want_v else None
    return s, u, v


def postprocess_symmetric(a, qbar, rank, want_u=True):
    """Symmetric postprocessing (A = A^T); returns (s[l] ascending, U[m,k] or None).

    src/lib/libisvd/core/stage/@x@_postprocess_symmetric.c:150-184: Z^T = Qbar^T A; W = Z^T Qbar
    (upper triangle, gemmt); syev(W) = W diag(s) W^T, eigenvalues ascending; U = Qbar W[:, :k].
    V is U.
    """
    zt = qbar.T @ a
    w = zt @ qbar
    s, wv = _syev_upper(np.ascontiguousarray(w))
    u = qbar @ wv[:, :rank] if want_u else None
    return s, u


# --------------------------------------------------------------------------------------------
# Driver
# --------------------------------------------------------------------------------------------
_ORTH = {"GR": orthogonalize_gramian, "TS": orthogonalize_tall_skinny_qr}
_POST = {"GR": postprocess_gramian, "TS": postprocess_tall_skinny_qr}


def isvd(a, omega, rank, num_sketch, alg_o="GR", alg_i="KN", alg_p="GR", argv_i=None):
    """Run sketch -> orthogonalize -> integrate -> postprocess (isvd_?Isvd at P = 1,
    src/lib/libisvd/core/driver/@x@_isvd.c:243-265).  ``omega`` is n x (N*l).

    Returns dict(s, u, v, qbar, iter, err, y, qs)."""
    y = sketch_gaussian_projection(a, omega)
    qs = _ORTH[alg_o.upper()](y, num_sketch)
    argv_i = dict(argv_i or {})
    if alg_i.upper() == "KN":
        qbar, it, err = integrate_kolmogorov_nagumo(qs, num_sketch, **argv_i)
    elif alg_i.upper() == "WY":
        qbar, it, err = integrate_wen_yin(qs, num_sketch, **argv_i)
    else:
        raise ValueError(f"unknown integration {alg_i!r}")
    s, u, v = _POST[alg_p.upper()](a, qbar, rank)
    return {"s": s, "u": u, "v": v, "qbar": qbar, "iter": it, "err": err, "y": y, "qs": qs}
