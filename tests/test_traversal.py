"""Graph containers and traversal (tensor_graph.py:61-223, element.py): the
package's TensorGraph visits the same elements in the same order as the
oracle's restatement of get_traversal_iterator (tensor_graph.py:136-170), for
both orderings, on the heavy cube, the small heavy-hex graph, a random
3-regular graph and a chain; traversal order fixes the asynchronous BP
schedule and the exact emulator's qubit order.  CPU only."""
import networkx as nx
import pytest

import dequantizer_amd as dq
from oracle import graph as og


def _copy(src):
    g = dq.TensorGraph()
    for nid in src.nodes:
        g.add_node(2, nid)
    for eid, e in src.edges.items():
        g.add_edge(eid, e.dimension)
    return g


def _graphs():
    rr = list(nx.random_regular_graph(3, 30, seed=5).edges)
    chain = [(i, i + 1) for i in range(11)]
    return {"heavy_cube": og.heavy_cube(), "small_heavy_hex": og.small_heavy_hex(),
            "rr30": og.graph_from_pairs(30, rr), "chain12": og.graph_from_pairs(12, chain)}


@pytest.mark.parametrize("name", ["heavy_cube", "small_heavy_hex", "rr30", "chain12"])
@pytest.mark.parametrize("ordering", ["bfs", "dfs"])
def test_traversal_matches_oracle(name, ordering):
    go = _graphs()[name]
    gp = _copy(go)
    ids_o = [(type(el).__name__, el.id) for el in go.traversal(ordering=ordering)]
    ids_p = [(type(el).__name__, el.id) for el in gp.get_traversal_iterator(ordering=ordering)]
    assert ids_p == ids_o
    assert gp.nodes_number == go.nodes_number and gp.edges_number == go.edges_number
    assert len(ids_p) == gp.nodes_number + gp.edges_number


def test_unknown_ordering_raises():
    gp = _copy(og.heavy_cube())
    with pytest.raises(KeyError):
        gp.get_traversal_iterator(ordering="random")
