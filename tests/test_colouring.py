"""Edge colourings that order the coupling gates into launch levels."""
import numpy as np

from dequantizer_amd.annealing import edge_colored, edge_colouring, random_regular


def _check_proper(n, pairs, col):
    seen = set()
    for (a, b), c in zip(pairs, col):
        for x in (int(a), int(b)):
            assert (x, int(c)) not in seen
            seen.add((x, int(c)))


def test_random_cubic_graphs_get_three_colours():
    for seed in range(4):
        ef = random_regular(400, 3, seed=seed)
        col = edge_colouring(ef.nodes_number, ef.coupled_spin_pairs)
        _check_proper(ef.nodes_number, ef.coupled_spin_pairs, col)
        assert col.max() == 2


def test_petersen_graph_needs_four_colours():
    """The Petersen graph is cubic but not 3-edge-colourable (class 2): the
    max-degree attempts all fail and the max-degree + 1 pass colours it."""
    outer = [(i, (i + 1) % 5) for i in range(5)]
    spokes = [(i, i + 5) for i in range(5)]
    inner = [(5 + i, 5 + (i + 2) % 5) for i in range(5)]
    pairs = outer + spokes + inner
    col = edge_colouring(10, pairs, restarts=50)
    _check_proper(10, pairs, col)
    assert col.max() == 3


def test_complete_graph_k4_and_trees():
    k4 = [(0, 1), (2, 3), (0, 2), (1, 3), (0, 3), (1, 2)]
    col = edge_colouring(4, k4)
    _check_proper(4, k4, col)
    assert col.max() == 2
    star = [(0, i) for i in range(1, 6)]
    col = edge_colouring(6, star)
    _check_proper(6, star, col)
    assert sorted(col.tolist()) == [0, 1, 2, 3, 4]
    assert edge_colouring(3, []).size == 0


def test_edge_colored_keeps_the_instance():
    ef = random_regular(100, 3, seed=5, couplings="pm1", field_std=0.1)
    ec = edge_colored(ef)
    key = {tuple(map(int, p)): j for p, j in zip(ef.coupled_spin_pairs, ef.coupling_amplitudes)}
    assert sorted(map(tuple, ec.coupled_spin_pairs.tolist())) == sorted(key)
    for p, j in zip(ec.coupled_spin_pairs, ec.coupling_amplitudes):
        assert key[tuple(map(int, p))] == j
    assert np.array_equal(ec.fields, ef.fields)
    col = edge_colouring(ec.nodes_number, ec.coupled_spin_pairs)
    assert col.max() == 2
