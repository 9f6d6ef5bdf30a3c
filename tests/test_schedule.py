"""CPU check of the BP sweep schedule (levels + message source codes).

Executing the product's schedule level by level, with the node update of the
oracle, must reproduce the oracle's sweep exactly -- asynchronous sweeps in
bfs and dfs traversal order and synchronous sweeps -- on lattices and random
regular graphs.  This pins the levelisation of the reference's in-order
message reads (belief_propagation.py:32-61) without a GPU.
"""
import numpy as np
import pytest

import dequantizer_amd as dq
from dequantizer_amd.engine import Topology
from oracle import graph as og
from oracle import mappings as om
from oracle import tensor_ops as ot
from tests.helpers import random_messages, random_tensors, to_oracle_messages


def simulate(topo, sync, tensors, messages):
    """Run one sweep with the schedule arrays: n2e_new from levels, e2n_new from edge codes."""
    n_levels, level_ptr, level_nodes, node_src, edge_src = topo.schedule(sync)
    E = topo.n_edges
    n2e_old = [None] * (2 * E)
    e2n_old = [None] * (2 * E)
    for j, e in enumerate(topo.edges):
        for s, nb in enumerate(e.neighbors):
            n2e_old[2 * j + s] = messages[dq.MessageID(nb.id, e.id)]
            e2n_old[2 * j + s] = messages[dq.MessageID(e.id, nb.id)]
    n2e_new = [None] * (2 * E)
    done = set()
    for lvl in range(n_levels):
        batch = level_nodes[level_ptr[lvl]:level_ptr[lvl + 1]]
        reads = {}
        for u in batch:
            inc = []
            for k in range(topo.adj_ptr[u], topo.adj_ptr[u + 1]):
                e, s = topo.adj_edge[k], topo.adj_side[k]
                code = node_src[k]
                if code == 0:
                    inc.append(e2n_old[2 * e + s])
                elif code == 1:
                    v = topo.edge_node[e][1 - s]
                    assert v in done, "dependency scheduled in the same or a later level"
                    inc.append(n2e_new[2 * e + 1 - s])
                else:
                    inc.append(n2e_old[2 * e + 1 - s])
            reads[u] = inc
        for u in batch:
            outs = ot.pass_through_node(tensors[topo.nodes[u].id], reads[u])
            for k, m in zip(range(topo.adj_ptr[u], topo.adj_ptr[u + 1]), outs):
                e, s = topo.adj_edge[k], topo.adj_side[k]
                n2e_new[2 * e + s] = m
        done.update(int(u) for u in batch)
    out = {}
    for j, e in enumerate(topo.edges):
        for s, nb in enumerate(e.neighbors):
            out[dq.MessageID(nb.id, e.id)] = n2e_new[2 * j + s]
            src = n2e_new if edge_src[2 * j + s] == 0 else n2e_old
            out[dq.MessageID(e.id, nb.id)] = src[2 * j + 1 - s]
    return out


def _regular(n, seed, cls):
    import networkx as nx
    g = cls()
    for _ in range(n):
        g.add_node(2)
    for a, b in nx.random_regular_graph(3, n, seed=seed).edges:
        g.add_edge((a, b), 2)
    return g


BUILDS = {
    "lattice3x3": (lambda: dq.get_nd_lattice([3, 3], 2, 2), lambda: og.nd_lattice([3, 3], 2, 2)),
    "torus232": (lambda: dq.get_nd_lattice([2, 3, 2], 2, 2, False), lambda: og.nd_lattice([2, 3, 2], 2, 2, False)),
    "rr3_16": (lambda: _regular(16, 1, dq.TensorGraph), lambda: _regular(16, 1, og.TensorGraph)),
    "eagle_inf": (dq.get_heavy_hex_ibm_eagle_lattice_infinite, og.heavy_hex_eagle_infinite),
}


@pytest.mark.parametrize("name", list(BUILDS))
@pytest.mark.parametrize("order", ["bfs", "dfs"])
@pytest.mark.parametrize("sync", [True, False])
def test_schedule_reproduces_sweep(name, order, sync):
    g, go = BUILDS[name][0](), BUILDS[name][1]()
    rng = np.random.default_rng(3)
    tensors = random_tensors(g, rng)
    messages = random_messages(g, rng)
    trav = list(g.get_traversal_iterator(ordering=order))
    trav_o = go.traversal(ordering=order)
    assert [el.id for el in trav] == [el.id for el in trav_o]
    topo = Topology([trav])
    got = simulate(topo, sync, tensors, messages)
    ref = om.belief_propagation_map(trav_o, sync)(tensors, to_oracle_messages(messages))
    assert len(got) == len(ref)
    for k, v in ref.items():
        np.testing.assert_array_equal(got[dq.MessageID(k.src, k.dst)], v)


def test_level_counts():
    """Synchronous sweeps are one level; async bfs on a random 3-regular graph
    is shallow, async dfs is a long chain."""
    g = _regular(200, 5, dq.TensorGraph)
    for order, lo, hi in [("bfs", 2, 40), ("dfs", 60, 400)]:
        topo = Topology([list(g.get_traversal_iterator(ordering=order))])
        assert topo.schedule(True)[0] == 1
        assert lo <= topo.schedule(False)[0] <= hi, (order, topo.schedule(False)[0])


def test_gate_rank_bounds():
    """The structural bound used before the device split: at least the SVD
    rank, and equal to it for diagonal and controlled gates."""
    from dequantizer_amd.emulator import gate_rank_bounds
    rng = np.random.default_rng(4)
    cnot = np.eye(4)[[0, 1, 3, 2]].reshape(2, 2, 2, 2)
    diag = np.diag(np.exp(1j * rng.uniform(0, 6, 4))).reshape(2, 2, 2, 2)
    ident = np.eye(4).reshape(2, 2, 2, 2)
    dense = (rng.standard_normal((4, 4)) + 1j * rng.standard_normal((4, 4))).reshape(2, 2, 2, 2)
    gates = np.stack([cnot, diag, ident, dense])
    bounds = gate_rank_bounds(gates)
    mats = np.transpose(gates, (0, 1, 3, 2, 4)).reshape(4, 4, 4)
    ranks = (np.linalg.svd(mats, compute_uv=False) > 1e-10).sum(axis=1)
    assert np.all(bounds >= ranks)
    assert list(bounds) == [2, 2, 2, 4]
    assert list(ranks) == [2, 2, 1, 4]


def test_dims_bookkeeping_matches_shapes():
    """apply_dims: the new bond is min(max_chi, m_a, m_b, chi_e r), checked on
    a 3-node chain against the reference's shape rule worked by hand."""
    from dequantizer_amd.emulator import apply_dims, dim_levels
    # chain 0 -e0- 1 -e1- 2, phys 2; adjacency padded with index 2 (= n_edges)
    adjpad = np.array([[0, 2], [0, 1], [1, 2]])
    dims = np.array([1, 1])
    levels = dim_levels(adjpad, 2, [0, 1], [0, 1], [1, 2], [0, 1])
    apply_dims(dims, levels, np.array([2, 2]), 2, 16)
    # gate on e0: m_0 = 2, m_1 = 2*dims[e1] = 2 -> min(2, 2, 1*2) = 2;
    # gate on e1: m_1 = 2*dims[e0] = 4, m_2 = 2 -> min(4, 2, 1*2) = 2
    assert list(dims) == [2, 2]
    apply_dims(dims, levels, np.array([2, 2]), 2, 16)
    # e0: m_0 = 2 -> 2; e1: m_2 = 2 -> 2 (leaves cap the bonds at phys)
    assert list(dims) == [2, 2]
    apply_dims(dims, levels, np.array([1, 1]), 2, 1)
    assert list(dims) == [1, 1]
