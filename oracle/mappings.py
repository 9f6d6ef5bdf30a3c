"""Whole-graph maps lifted over a traverser.

Restates dequantizer/src/mappings/*.py.  ``two_side_density_matrix`` is new
(the reference exports only the one-site density matrix); it is pinned by a
state-vector contraction on trees and by partial-trace consistency.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).
"""
from typing import Dict, List, Sequence

import numpy as np

from . import tensor_ops as ops
from .graph import Edge, MessageID, Node


def belief_propagation_map(traverser: Sequence, is_synchronous: bool = False):
    """One BP sweep.  mappings/belief_propagation.py:24-64.

    Elements are visited in traverser order.  Asynchronous sweeps read a
    message already updated in this sweep when there is one; synchronous
    sweeps read only the input dict."""

    def bp(tensors: Dict, messages: Dict) -> Dict:
        updated: Dict = {}
        for el in traverser:
            incoming = []
            for nb in el.neighbors:
                mid = MessageID(nb.id, el.id)
                if not is_synchronous and mid in updated:
                    incoming.append(updated[mid])
                else:
                    incoming.append(messages[mid])
            if isinstance(el, Node):
                outs = ops.pass_through_node(tensors[el.id], incoming)
            else:
                outs = ops.pass_through_edge(incoming)
            for nb, m in zip(el.neighbors, outs):
                updated[MessageID(el.id, nb.id)] = m
        return updated

    return bp


def vidal_gauge_fixing_map(traverser: Sequence, eps: float):
    """Converged messages -> Vidal gauge (Gamma, lambda).
    mappings/vidal_gauge.py:21-63."""

    def vg(tensors: Dict, messages: Dict):
        lambdas = {}
        orth = {}
        nodes: List[Node] = []
        for el in traverser:
            if isinstance(el, Edge):
                inc = [messages[MessageID(nb.id, el.id)] for nb in el.neighbors]
                xs, lm = ops.edge_svd(inc, eps)
                lambdas[el.id] = lm
                for x, nb in zip(xs, el.neighbors):
                    orth[MessageID(el.id, nb.id)] = x
            else:
                nodes.append(el)
        gammas = {}
        for node in nodes:
            xs = [orth[MessageID(nb.id, node.id)] for nb in node.neighbors]
            gammas[node.id] = ops.absorb_by_node(tensors[node.id], xs)
        return gammas, lambdas

    return vg


def vidal_gauge_distance_map(traverser: Sequence, eps: float):
    """Sum of per-node Vidal distances over 2|E|.  mappings/vidal_distance.py:17-42."""

    def vd(tensors: Dict, lambdas: Dict) -> float:
        dist = 0.0
        n_edges = 0
        for el in traverser:
            if isinstance(el, Node):
                dist += ops.vidal_dist(tensors[el.id], [lambdas[e.id] for e in el.neighbors], eps)
            else:
                n_edges += 1
        return dist / (2 * n_edges)

    return vd


def symmetric_gauge_fixing_map(traverser: Sequence, eps: float):
    """Vidal -> symmetric gauge.  mappings/symmetric_gauge.py:17-39."""

    def sg(tensors: Dict, lambdas: Dict) -> Dict:
        out = {}
        for el in traverser:
            if isinstance(el, Node):
                out[el.id] = ops.canonicalize(tensors[el.id], [lambdas[e.id] for e in el.neighbors], eps)
        return out

    return sg


def lambdas2messages(lambdas: Dict) -> Dict:
    """diag(lambda) in all four directions of each edge (not trace-normalised).
    mappings/symmetric_gauge.py:50-61."""
    out = {}
    for eid, lm in lambdas.items():
        m = np.diag(np.asarray(lm)).astype(np.complex128)
        for nid in eid:
            out[MessageID(eid, nid)] = m
            out[MessageID(nid, eid)] = m
    return out


def messages_frob_distance(lhs: Dict, rhs: Dict) -> float:
    """Mean Frobenius distance between two message dicts.
    mappings/messages_distance.py:14-22."""
    assert len(lhs) == len(rhs), f"{len(lhs)}, {len(rhs)}"
    return float(sum(np.linalg.norm(v - rhs[k]) for k, v in lhs.items()) / len(lhs))


def one_side_density_matrix(node: Node, tensors: Dict, messages: Dict) -> np.ndarray:
    """rho[s, s'] from the node tensor and its incoming messages, trace 1.
    mappings/density_matrix.py:16-37."""
    t = tensors[node.id]
    assert node.degree == t.ndim - 1
    ket = t
    for i, nb in enumerate(node.neighbors):
        ket = ops.apply_matrix_on_axis(ket, messages[MessageID(nb.id, node.id)], i)
    axes = list(range(t.ndim - 1))
    rho = np.tensordot(ket, t.conj(), axes=(axes, axes))
    return rho / np.trace(rho)


def two_side_density_matrix(edge: Edge, tensors: Dict, messages: Dict) -> np.ndarray:
    """Two-site BP density matrix on an edge (NEW: no reference counterpart).

    rho[(s_a s_b), (s_a' s_b')]: the shared bond is contracted ket-to-ket and
    bra-to-bra, every other bond of a and b carries its incoming message
    (M[bra, ket] on the ket side), then trace normalisation.  On a tree with
    converged messages it equals the exact reduced density matrix."""
    a, b = edge.neighbors
    ta, tb = tensors[a.id], tensors[b.id]
    ia, ib = a.indices[edge.id], b.indices[edge.id]
    ka, kb = ta, tb
    for i, nb in enumerate(a.neighbors):
        if i != ia:
            ka = ops.apply_matrix_on_axis(ka, messages[MessageID(nb.id, a.id)], i)
    for i, nb in enumerate(b.neighbors):
        if i != ib:
            kb = ops.apply_matrix_on_axis(kb, messages[MessageID(nb.id, b.id)], i)
    # theta_ket[oa.., sa, ob.., sb] with the shared bond summed
    ket = np.tensordot(ka, kb, axes=([ia], [ib]))
    bra = np.tensordot(ta.conj(), tb.conj(), axes=([ia], [ib]))
    da, db = ta.ndim - 1, tb.ndim - 1  # axes left per side: da-1 bonds + phys
    # ket axes: a-others (da-1), sa, b-others (db-1), sb
    sa, sb = da - 1, da - 1 + db
    others = [x for x in range(ket.ndim) if x not in (sa, sb)]
    rho = np.tensordot(ket, bra, axes=(others, others))  # (sa, sb, sa', sb')
    rho = rho.reshape(ta.shape[-1] * tb.shape[-1], -1)
    return rho / np.trace(rho)
