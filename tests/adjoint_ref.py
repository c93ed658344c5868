"""Reference derivative of the top-k singular triplets of a dense SVD (test infrastructure).

L(U_k, S_k, V_k) = sum_i w_i s_i + <G_U, U_k> + <G_V, V_k>; dense_svd_grad() is the textbook
closed form over ALL n singular triplets of A (m >= n), F_ij = 1 / (s_j^2 - s_i^2), plus the
(I - U_n U_n^T) G_U S^-1 V^T term of the left null space.  It is pinned against FP64 central
differences of dense_loss() (numpy LAPACK) in test_adjoint_reference.py."""
import numpy as np


def dense_svd_grad(torch, u, s, v, w, gu, gv, k):
    """dL/dA for the top-k triplets of A = u diag(s) v^T with ALL n triplets (u m x n, v n x n;
    torch FP64 tensors): u [(F o (J - J^T)) S + S (F o (K - K^T)) + diag(gs)] v^T
    + (I - u u^T) gU S^-1 v_k^T, gradients zero-padded from k to n.  Needs s_k > s_{k+1}."""
    n = s.shape[0]
    J = torch.zeros(n, n, dtype=u.dtype, device=u.device)
    K = torch.zeros_like(J)
    J[:, :k] = u.T @ gu
    K[:, :k] = v.T @ gv
    s2 = s * s
    E = s2[None, :] - s2[:, None]
    # equal singular values occur only among the discarded pairs (j, j' >= k), where
    # J - J^T and K - K^T vanish: F = 0 there
    zero = E == 0
    F = torch.where(zero, torch.zeros_like(E), 1.0 / torch.where(zero, torch.ones_like(E), E))
    mid = (F * (J - J.T)) * s[None, :] + s[:, None] * (F * (K - K.T))
    mid[torch.arange(k), torch.arange(k)] += w
    gperp = gu - u @ (u.T @ gu)
    return u @ (mid @ v.T) + (gperp / s[None, :k]) @ v[:, :k].T


def dense_loss(a, k, w, GU, GV, u_sign_ref):
    U, S, Vt = np.linalg.svd(a, full_matrices=False)
    U, V = U[:, :k], Vt[:k].T
    sg = np.sign(np.sum(U * u_sign_ref, axis=0))  # the forward's sign of each singular pair
    return w @ S[:k] + np.sum(GU * (U * sg)) + np.sum(GV * (V * sg))


def full_rank_problem(m=512, n=256, seed=21):
    rng = np.random.default_rng(seed)
    u, _ = np.linalg.qr(rng.standard_normal((m, n)))
    v, _ = np.linalg.qr(rng.standard_normal((n, n)))
    s = 2 * 0.9 ** np.arange(n) + 0.01
    return (u * s) @ v.T, u, s, v


