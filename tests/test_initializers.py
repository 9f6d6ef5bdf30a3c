"""Initialisers and graph factories (graph/message_initializer.py:14-26,
graph/tensor_initializers.py:25-138, tensor_graph.py:470-498): the package
gives the oracle's values on identical keys, element for element.  Both
draw from NumPy generators derived from the key (the reference's JAX
threefry values are not reproduced: those stay parity unpinned); the
deterministic ones (standard state, Potts) carry no draw at all.  CPU only."""
import networkx as nx
import numpy as np
import pytest

import dequantizer_amd as dq
from oracle import graph as og
from tests.test_traversal import _copy

KEY = (0, 44)


def _pair(bond):
    pairs = list(nx.random_regular_graph(3, 12, seed=1).edges)
    go = og.graph_from_pairs(12, pairs, bond=bond)
    return _copy(go), go


def _same(a, b):
    assert len(a) == len(b)
    ka, kb = sorted(a, key=str), sorted(b, key=str)
    for x, y in zip(ka, kb):
        assert str(x) == str(y)
        assert np.array_equal(np.asarray(a[x]), np.asarray(b[y])), x


@pytest.mark.parametrize("kind", ["messages", "normal", "ghz"])
def test_keyed_initializers_match_oracle(kind):
    gp, go = _pair(3)
    if kind == "messages":
        _same(gp.init_messages(dq.get_message_random_nonnegative_initializer(KEY)),
              go.init_messages(og.message_random_nonnegative_initializer(KEY)))
    elif kind == "normal":
        _same(gp.init_tensors(dq.get_tensor_random_normal_initializer(KEY)),
              go.init_tensors(og.tensor_random_normal_initializer(KEY)))
    else:
        _same(gp.init_tensors(dq.get_tensor_bloated_ghz_initializer(KEY)),
              go.init_tensors(og.tensor_bloated_ghz_initializer(KEY)))


def test_deterministic_initializers_match_oracle():
    gp, go = _pair(1)
    _same(gp.init_tensors(dq.get_tensor_std_state_initializer()), go.init_tensors(og.tensor_std_state_initializer()))
    gp, go = _pair(2)
    beta = 0.4
    two = np.exp(beta * np.array([[1.0, -1.0], [-1.0, 1.0]]))
    single = np.exp(0.1 * np.array([1.0, -1.0]))
    _same(gp.init_tensors(dq.get_potts_initializer(two, single)), go.init_tensors(og.potts_initializer(two, single)))


def test_random_tree_matches_oracle():
    bonds = [1, 2, 3, 2, 1, 2, 3, 2, 1, 2, 3, 2, 1, 2]
    gp = dq.get_random_tree_tensor_graph(15, 2, bonds, KEY)
    go = og.random_tree(15, 2, bonds, KEY)
    ep = {tuple(eid): e.dimension for eid, e in gp.edges.items()}
    eo = {tuple(eid): e.dimension for eid, e in go.edges.items()}
    assert ep == eo
    assert [(type(el).__name__, el.id) for el in gp.get_traversal_iterator()] == \
        [(type(el).__name__, el.id) for el in go.traversal()]
