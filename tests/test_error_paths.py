"""Error behaviour of the host API (CPU, no GPU): the reference's exception
types for a malformed gate, a gate on a pair of nodes without an edge, a gate
on a hyperedge, duplicate ids and an unknown traversal ordering
(dequantizer/src/graph/tensor_graph.py:62-94,155,305-317,395-407).  Every
check runs before any device work, so these raise the same way with or
without a GPU; the oracle raises the same types."""
import numpy as np
import pytest

import dequantizer_amd as dq
from oracle import graph as og


def _pair(G):
    g = G()
    for _ in range(3):
        g.add_node(2)
    g.add_edge((0, 1), 1)
    g.add_edge((1, 2), 1)
    return g


def _state(g):
    tensors = {n.id: np.ones((*n.bond_shape, 2), dtype=np.complex128) for n in g.nodes.values()}
    lambdas = {e: np.ones(1) for e in g.edges}
    return tensors, lambdas


GOOD = np.eye(4, dtype=np.complex128).reshape(2, 2, 2, 2)
BAD_SHAPES = [np.eye(4, dtype=np.complex128), np.ones((2, 2, 2), dtype=np.complex128),
              np.ones((2, 2, 3, 2), dtype=np.complex128), np.ones((2, 3, 2, 2), dtype=np.complex128)]


@pytest.mark.parametrize("G", [dq.TensorGraph, og.TensorGraph])
@pytest.mark.parametrize("k", range(len(BAD_SHAPES)))
def test_malformed_gate_raises_value_error(G, k):
    g = _pair(G)
    t, lam = _state(g)
    with pytest.raises(ValueError):
        g.apply2_to_vidal_canonical(0, 1, t, lam, BAD_SHAPES[k], 1e-10, None)
    with pytest.raises(ValueError):
        g.apply2(0, 1, t, BAD_SHAPES[k], 1e-10)


@pytest.mark.parametrize("G", [dq.TensorGraph, og.TensorGraph])
def test_gate_without_an_edge_raises_value_error(G):
    g = _pair(G)
    t, lam = _state(g)
    with pytest.raises(ValueError):
        g.apply2_to_vidal_canonical(0, 2, t, lam, GOOD, 1e-10, None)
    with pytest.raises(ValueError):
        g.apply2(2, 0, t, GOOD, 1e-10)


@pytest.mark.parametrize("G", [dq.TensorGraph, og.TensorGraph])
def test_gate_on_a_hyperedge_is_not_found(G):
    """An edge id is the tuple of its nodes, so a hyperedge (0, 1, 2) is not
    found under (0, 1) or (1, 0): ValueError, as in the reference."""
    g = G()
    for _ in range(3):
        g.add_node(2)
    g.add_edge((0, 1, 2), 1)
    t, lam = _state(g)
    with pytest.raises(ValueError):
        g.apply2_to_vidal_canonical(0, 1, t, lam, GOOD, 1e-10, None)


@pytest.mark.parametrize("G", [dq.TensorGraph, og.TensorGraph])
def test_graph_construction_errors_are_key_errors(G):
    g = _pair(G)
    with pytest.raises(KeyError):
        g.add_node(2, 1)  # duplicate node id
    with pytest.raises(KeyError):
        g.add_edge((0, 1), 2)  # duplicate edge id
    with pytest.raises(KeyError):
        g.add_edge((0, 7), 2)  # unknown node
    with pytest.raises(KeyError):  # at the call or at the first step of the iterator
        if G is dq.TensorGraph:
            list(g.get_traversal_iterator(ordering="random"))
        else:
            g.traversal(ordering="random")
