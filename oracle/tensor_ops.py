"""Local tensor arithmetic of the BP / Vidal / simple-update path, complex128.

Restates dequantizer/src/tensor_ops.py.  Tensor conventions (as in the
reference): a node tensor has axes (bond_0, ..., bond_{d-1}, phys) with bonds
in the order the node's edges were added; a message M[bra, ket] is contracted
with a node tensor through its axis 1 (the ket side).

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).
"""
from typing import List, Optional, Sequence, Tuple

import numpy as np
from scipy.special import entr

FLT_MAX = float(np.finfo(np.float32).max)


def find_rank(lmbd: np.ndarray, accuracy: Optional[float], rank: Optional[int]) -> int:
    """Bond rank kept by a truncation.  tensor_ops.py:7-17.

    With an accuracy, the rank is n minus the number of tail positions whose
    cumulative (from the smallest value) norm, relative to the full norm, is
    below the accuracy.  Pinned by closed-form cases in the oracle tests."""
    if accuracy is None:
        return lmbd.shape[0] if rank is None else int(rank)
    lmbd = np.asarray(lmbd, dtype=np.float64)
    scale = np.sqrt(np.sum(lmbd**2))
    tail = np.sqrt(np.cumsum(lmbd[::-1] ** 2)) / scale
    return int(lmbd.shape[0] - np.count_nonzero(tail < accuracy))


def safe_inverse(lmbd: np.ndarray, eps: float) -> np.ndarray:
    """1/lambda, with values below eps mapped to 1/(FLT_MAX + lambda).
    tensor_ops.py:20-21."""
    return 1.0 / ((lmbd < eps) * FLT_MAX + lmbd)


def safe_truncate(lmbd: np.ndarray, eps: float) -> np.ndarray:
    """Zero every value that is not above eps.  tensor_ops.py:24-25."""
    return (lmbd > eps) * lmbd


def apply_matrix_on_axis(tensor: np.ndarray, matrix: np.ndarray, axis: int) -> np.ndarray:
    """out[.., a@axis, ..] = sum_b matrix[a, b] * tensor[.., b@axis, ..]."""
    out = np.tensordot(matrix, tensor, axes=([1], [axis]))
    return np.moveaxis(out, 0, axis)


def bond_gram(conj_tensor: np.ndarray, ket: np.ndarray, idx: int) -> np.ndarray:
    """G[k, l] = sum over every axis but ``idx`` of conj_tensor[..k..] * ket[..l..]."""
    others = [a for a in range(ket.ndim) if a != idx]
    return np.tensordot(conj_tensor, ket, axes=(others, others))


def pass_through_node(tensor: np.ndarray, incoming: Sequence[np.ndarray]) -> List[np.ndarray]:
    """BP node update.  tensor_ops.py:36-60.

    For every bond idx: U = T with message j applied (ket side) on every bond
    j != idx; out = Gram(conj T, U) over all axes but idx; out /= trace(out).
    The normalisation is the trace (not the Frobenius norm)."""
    degree = tensor.ndim - 1
    assert degree == len(incoming), f"{degree}, {len(incoming)}"
    conj_tensor = tensor.conj()
    out = []
    for idx in range(degree):
        up = tensor
        for j, msg in enumerate(incoming):
            if j != idx:
                up = apply_matrix_on_axis(up, msg, j)
        m = bond_gram(conj_tensor, up, idx)
        out.append(m / np.trace(m))
    return out


def pass_through_edge(incoming: Sequence[np.ndarray]) -> List[np.ndarray]:
    """Edge update: the two node->edge messages swap roles.  tensor_ops.py:70-76."""
    return [incoming[1], incoming[0]]


def absorb_by_node(tensor: np.ndarray, matrices: Sequence[np.ndarray]) -> np.ndarray:
    """T'[..k_i..] = sum_b T[..b_i..] X_i[b, k] on every bond, then Frobenius
    normalisation.  tensor_ops.py:88-93."""
    for i, x in enumerate(matrices):
        tensor = apply_matrix_on_axis(tensor, x.T, i)
    return tensor / np.linalg.norm(tensor)


def edge_svd(incoming: Sequence[np.ndarray], eps: float) -> Tuple[List[np.ndarray], np.ndarray]:
    """Vidal edge decomposition from the two node->edge messages.
    tensor_ops.py:108-130.

    m^{1/2} and m^{-1/2} from Hermitian eigendecompositions (eigenvalues below
    eps dropped / safely inverted), SVD of m_f^{1/2} (m_b^{1/2})^T, lambda
    normalised then thresholded, U and Vh masked by lambda > eps and divided
    by their own Frobenius norms.  Returns ([m_f^{-1/2} U, m_b^{-1/2} Vh^T], lambda)."""
    if len(incoming) != 2:
        raise NotImplementedError("Edge svd is not yet implemented for edge degrees != 2.")
    lf, uf = np.linalg.eigh(incoming[0])
    lb, ub = np.linalg.eigh(incoming[1])

    def _fn(u, vals):
        return (u * vals[np.newaxis, :]) @ u.conj().T

    m_sq_inv_f = _fn(uf, np.sqrt(safe_inverse(lf, eps)))
    m_sq_inv_b = _fn(ub, np.sqrt(safe_inverse(lb, eps)))
    m_sq_f = _fn(uf, np.sqrt(safe_truncate(lf, eps)))
    m_sq_b = _fn(ub, np.sqrt(safe_truncate(lb, eps)))
    u, lmbd, vh = np.linalg.svd(m_sq_f @ m_sq_b.T, full_matrices=False)
    lmbd = lmbd / np.linalg.norm(lmbd)
    lmbd = safe_truncate(lmbd, eps)
    mask = lmbd > eps
    u = u * mask[np.newaxis, :]
    u = u / np.linalg.norm(u)
    vh = vh * mask[:, np.newaxis]
    vh = vh / np.linalg.norm(vh)
    return [m_sq_inv_f @ u, m_sq_inv_b @ vh.T], lmbd


def _complex_gt(z: np.ndarray, eps: float) -> np.ndarray:
    """Lexicographic ``z > eps`` for complex z (real part first, then imaginary),
    the ordering jax.numpy uses for complex comparisons."""
    re, im = np.real(z), np.imag(z)
    return np.where(re == eps, im > 0.0, re > eps)


def vidal_dist(tensor: np.ndarray, lambdas: Sequence[np.ndarray], eps: float) -> float:
    """Distance of one node tensor to the Vidal gauge.  tensor_ops.py:143-175.

    For every bond idx: R = Gram(conj T, T * prod_{j != idx} lambda_j^2) over
    all axes but idx, R /= trace R; ideal = diag(diag R > eps) / count; the
    contribution is the nuclear norm of R - ideal."""
    degree = tensor.ndim - 1
    assert degree == len(lambdas)
    conj_tensor = tensor.conj()
    dist = 0.0
    for idx in range(degree):
        up = tensor
        for j, lmbd in enumerate(lambdas):
            if j != idx:
                shape = [1] * (degree + 1)
                shape[j] = lmbd.shape[0]
                up = up * (lmbd.reshape(shape) ** 2)
        r = bond_gram(conj_tensor, up, idx)
        r = r / np.trace(r)
        ideal = np.diag(_complex_gt(np.diag(r), eps)).astype(np.complex128)
        ideal = ideal / np.trace(ideal)
        dist += float(np.sum(np.linalg.svd(r - ideal, compute_uv=False)))
    return dist


def canonicalize(tensor: np.ndarray, lambdas: Sequence[np.ndarray], eps: float) -> np.ndarray:
    """Symmetric gauge: T * prod_i sqrt(lambda_i > eps) on bond i, then Frobenius
    normalisation.  tensor_ops.py:189-199."""
    degree = len(lambdas)
    assert degree == tensor.ndim - 1
    for i, lmbd in enumerate(lambdas):
        shape = [1] * (degree + 1)
        shape[i] = lmbd.shape[0]
        tensor = tensor * np.sqrt(safe_truncate(lmbd, eps)).reshape(shape)
    return tensor / np.linalg.norm(tensor)


def decompose_gate(gate: np.ndarray, eps: float) -> Tuple[np.ndarray, np.ndarray]:
    """Operator-Schmidt split of a two-site gate (o1, o2, i1, i2).
    tensor_ops.py:202-218.

    SVD of the (o1 i1) x (o2 i2) matrix; the r singular values above eps are
    kept; the halves are (r, o1, i1) and (r, o2, i2), each scaled by sqrt(s)."""
    assert gate.ndim == 4 and gate.shape[0] == gate.shape[2] and gate.shape[1] == gate.shape[3]
    d1, d2 = gate.shape[0], gate.shape[1]
    mat = np.transpose(gate, (0, 2, 1, 3)).reshape((d1 * d2, -1))
    u, s, vh = np.linalg.svd(mat, full_matrices=False)
    rank = int(np.count_nonzero(s > eps))
    root = np.sqrt(s[:rank])
    first = np.transpose((u[:, :rank] * root[np.newaxis]).reshape((d1, d1, rank)), (2, 0, 1))
    second = (vh[:rank] * root[:, np.newaxis]).reshape((rank, d2, d2))
    return first, second


def apply_half_gate(tensor: np.ndarray, half: np.ndarray, index: int) -> np.ndarray:
    """Contract a gate half (r, out, in) into the physical axis and fuse r into
    bond ``index`` as the fast-running part (fused = b * r + q); then Frobenius
    normalisation.  tensor_ops.py:221-237."""
    shape = list(tensor.shape)
    shape[index] *= half.shape[0]
    t = np.einsum("...i,qoi->...qo", tensor, half)  # (bonds..., r, out)
    t = np.moveaxis(t, -2, index + 1)  # r right after bond `index`
    t = t.reshape(shape)
    return t / np.linalg.norm(t)


def apply_gate_halves(t1, t2, i1, i2, h1, h2):
    """tensor_ops.py:252-262."""
    return apply_half_gate(t1, h1, i1), apply_half_gate(t2, h2, i2)


def _broadcast_on_others(tensor: np.ndarray, lambdas: Sequence[np.ndarray], index: int, fn) -> np.ndarray:
    """Multiply axis k of ``tensor`` by fn(lambda) of the k-th bond other than
    ``index`` (the bonds keep their relative order)."""
    others = [lm for j, lm in enumerate(lambdas) if j != index]
    for k, lm in enumerate(others):
        shape = [1] * tensor.ndim
        shape[k] = lm.shape[0]
        tensor = tensor * fn(lm).reshape(shape)
    return tensor


def su_preprocess(tensor, lambdas, half, eps, index):
    """One side of the simple update.  tensor_ops.py:265-295.

    Scale the other bonds by lambda, apply the gate half, move bond ``index``
    last, reduced QR of the (others.., phys) x (chi r) matrix, and undo the
    lambda scaling on Q with the safe inverse.  Returns (Q tensor, R)."""
    degree = len(lambdas)
    assert tensor.ndim - 1 == degree
    scaled = tensor
    for j, lm in enumerate(lambdas):
        if j != index:
            shape = [1] * (degree + 1)
            shape[j] = lm.shape[0]
            scaled = scaled * lm.reshape(shape)
    t = apply_half_gate(scaled, half, index)
    t = np.moveaxis(t, index, -1)  # (others.., phys, fused)
    lead = t.shape[:-1]
    q, r = np.linalg.qr(t.reshape((-1, t.shape[-1])), mode="reduced")
    q = q.reshape((*lead, -1))
    q = _broadcast_on_others(q, lambdas, index, lambda lm: safe_inverse(lm, eps))
    return q, r


def simple_update(t1, t2, lambdas1, lambdas2, half1, half2, index1, index2, eps, accuracy, max_rank=None):
    """Two-site gate on a Vidal canonical pair with truncation.
    tensor_ops.py:318-389.

    Returns (Gamma_1, Gamma_2, lambda (unnormalised, length <= max_rank),
    truncation error, entanglement entropy)."""
    q1, r1 = su_preprocess(t1, lambdas1, half1, eps, index1)
    q2, r2 = su_preprocess(t2, lambdas2, half2, eps, index2)
    assert np.all(np.abs(lambdas1[index1] - lambdas2[index2]) < 1e-5)
    lam = lambdas1[index1]
    extra = r1.shape[-1] // lam.shape[0]
    lam_rep = np.repeat(lam, extra)  # index b * extra + q
    core = (r1 * lam_rep[np.newaxis, :]) @ r2.T
    u, s, vh = np.linalg.svd(core, full_matrices=False)
    s = safe_truncate(s, eps)
    rank = find_rank(s, accuracy, max_rank)
    s_hat = s / np.linalg.norm(s)
    entropy = float(np.sum(entr(s_hat**2)))
    err = float(np.linalg.norm(s_hat[rank:]))
    s = s[:rank]
    u = u[:, :rank]
    vh = vh[:rank]
    g1 = np.tensordot(q1, u, axes=([q1.ndim - 1], [0]))  # (others.., phys, rank)
    g1 = np.moveaxis(g1, -1, index1)
    g2 = np.tensordot(q2, vh, axes=([q2.ndim - 1], [1]))
    g2 = np.moveaxis(g2, -1, index2)
    return g1, g2, s, err, entropy


def apply1(tensor: np.ndarray, gate: np.ndarray) -> np.ndarray:
    """One-site gate on the physical axis, out = in . gate^T, then Frobenius
    normalisation.  tensor_graph.py:268-275."""
    t = np.tensordot(tensor, gate, axes=([tensor.ndim - 1], [1]))
    return t / np.linalg.norm(t)
