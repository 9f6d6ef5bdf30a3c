"""Workload drivers on top of the engine.

* ``process_ising_sqrt_partition_network``: the 3D classical Ising
  square-root network (examples/3d_ising_canonicalization.py:89-160): Potts
  node tensors per inverse temperature, BP until a convergence criterion,
  Vidal canonicalisation and the mean magnetisation.
* ``run_eagle_simulation`` / ``run_eagle_simulations`` (several RX angles
  as instances of one batched engine): kicked-Ising dynamics on the heavy-hex Eagle
  lattice (examples/ibm_eagle_processor.py:56-145): RX(theta) on every qubit,
  then RZZ(-pi/2) on every edge, <Z> of the target qubit after each layer.
  ``colour_layers=True`` applies the RZZ gates by edge-colour class with a
  regauge after each class (heavy-hex is bipartite with degree <= 3: three
  classes).
* ``annealing_instances``: the batched annealing instance set (random
  3-regular graphs with N drawn uniformly from 100..1000, +-1 couplings,
  N(0, 0.1^2) fields), one seed per instance.
"""
from dataclasses import dataclass
from typing import Dict, List, Optional

import numpy as np

from .annealing import EnergyFunction, random_regular
from .emulator import BPQuantumEmulator
from .graph import (Edge, Node, TensorGraph, get_heavy_hex_ibm_eagle_lattice, get_heavy_hex_ibm_eagle_lattice_infinite,
                    get_message_random_nonnegative_initializer, get_nd_lattice, get_potts_initializer)
from .mappings import (get_belief_propagation_map, get_one_side_density_matrix, get_vidal_gauge_distance_map,
                       get_vidal_gauge_fixing_map, messages_frob_distance)
from .random import PRNGKey


@dataclass
class CanonicalizationResult:
    """Per inverse temperature: tensors, messages, the convergence history,
    the iteration count and the mean magnetisation (the fields of the
    reference's result class, 3d_ising_canonicalization.py:47-60), plus the
    Vidal gauge (gammas, lambdas) of the converged messages."""
    tensors: List[Dict]
    messages: List[Dict]
    vidal_dist_per_iter: List[np.ndarray]
    number_of_iterations: List[int]
    average_magnetization: List[float]
    gammas: List[Dict]
    lambdas: List[Dict]


def process_ising_sqrt_partition_network(lattice_dim: int, lattice_size: int, betas, coupling: float, field: float,
                                         accuracy: float, open_boundary: bool, synchronous_update: bool = False,
                                         key=PRNGKey(42), traversal_type: str = "dfs", criterion: str = "vidal",
                                         max_iterations: Optional[int] = None) -> CanonicalizationResult:
    """3d_ising_canonicalization.py:89-160.  ``criterion`` "vidal" stops when
    the Vidal distance is below ``accuracy`` (the reference); "residual"
    stops when the mean message change of a sweep is below it, then fixes the
    Vidal gauge once."""
    if criterion not in ("vidal", "residual"):
        raise ValueError(f"unknown criterion {criterion!r}")
    result = CanonicalizationResult([], [], [], [], [], [], [])
    lattice = get_nd_lattice(lattice_dim * [lattice_size], 2, 2, open_boundary)
    traverser = list(lattice.get_traversal_iterator(ordering=traversal_type) or iter([]))
    bp_map = get_belief_propagation_map(traverser, synchronous_update)
    vg_map = get_vidal_gauge_fixing_map(traverser, accuracy)
    vd_map = get_vidal_gauge_distance_map(traverser, accuracy)
    spin = np.array([1.0, -1.0])
    for beta in betas:
        two_spin_factor = np.exp(-coupling * beta * np.tensordot(spin, spin, axes=0))
        one_spin_factor = np.exp(-field * beta * spin)
        tensors = lattice.init_tensors(get_potts_initializer(two_spin_factor, one_spin_factor))
        messages = lattice.init_messages(get_message_random_nonnegative_initializer(key))
        dist = float(np.finfo(np.float64).max)
        history: List[float] = []
        gammas = lambdas = None
        while dist > accuracy and (max_iterations is None or len(history) < max_iterations):
            new_messages = bp_map(tensors, messages)
            if criterion == "vidal":
                messages = new_messages
                gammas, lambdas = vg_map(tensors, messages)
                dist = float(vd_map(gammas, lambdas))
            else:
                dist = float(messages_frob_distance(new_messages, messages))
                messages = new_messages
            history.append(dist)
        if criterion == "residual":
            gammas, lambdas = vg_map(tensors, messages)
        magnetization = 0.0
        for element in traverser:
            if isinstance(element, Node):
                dens = get_one_side_density_matrix(element, tensors, messages)
                magnetization += 2.0 * float(dens[0, 0].real) - 1.0
        result.tensors.append(tensors)
        result.messages.append(messages)
        result.vidal_dist_per_iter.append(np.asarray(history))
        result.number_of_iterations.append(len(history))
        result.average_magnetization.append(magnetization / lattice.nodes_number)
        result.gammas.append(gammas)
        result.lambdas.append(lambdas)
    return result


def edge_colouring(graph: TensorGraph) -> Dict[tuple, int]:
    """Proper edge colouring: each edge in insertion order takes a colour free
    at both ends, after swapping a two-colour alternating path when needed
    (Koenig's theorem: max-degree colours suffice on bipartite graphs; a path
    that closes an odd cycle gets a new colour instead)."""
    at: Dict = {n: {} for n in graph.nodes}  # node -> colour -> neighbour
    pair_edge = {}
    for eid, e in graph.edges.items():
        a, b = (nb.id for nb in e.neighbors)
        pair_edge[(a, b)] = pair_edge[(b, a)] = eid
    colour: Dict[tuple, int] = {}

    def free(n):
        c = 0
        while c in at[n]:
            c += 1
        return c

    for eid, e in graph.edges.items():
        u, v = (nb.id for nb in e.neighbors)
        a, b = free(u), free(v)
        if a not in at[v]:
            c = a
        elif b not in at[u]:
            c = b
        else:
            path, x, cur = [], v, a
            while cur in at[x]:
                y = at[x][cur]
                path.append((x, y, cur))
                x, cur = y, (b if cur == a else a)
            if any(y == u for _, y, _ in path):
                c = max(max(at[u], default=-1), max(at[v], default=-1)) + 1
            else:
                for x, y, cl in path:
                    del at[x][cl]
                    del at[y][cl]
                for x, y, cl in path:
                    nc = b if cl == a else a
                    at[x][nc] = y
                    at[y][nc] = x
                    colour[pair_edge[(x, y)]] = nc
                c = a
        at[u][c] = v
        at[v][c] = u
        colour[eid] = c
    return colour


RZZ = np.diag(np.exp(np.array([1j, -1j, -1j, 1j]) * np.pi / 4)).reshape(2, 2, 2, 2)


def rx(theta: float) -> np.ndarray:
    """exp(-i theta X / 2) (ibm_eagle_processor.py:111-118)."""
    c, s = np.cos(theta / 2), np.sin(theta / 2)
    return np.array([[c, -1j * s], [-1j * s, c]], dtype=np.complex128)


def run_eagle_simulation(max_chi: int, layers_per_regauging: int, trotter_steps: int, accuracy: float, theta: float,
                         is_infinite: bool = True, synchronous_update: bool = False,
                         max_belief_propagation_iterations: Optional[int] = None, key=PRNGKey(42),
                         traversal_type: str = "dfs", *, dtype=None, chi: Optional[int] = None,
                         colour_layers: bool = False) -> List[float]:
    """ibm_eagle_processor.py:56-145: <Z> of the target qubit (2 on the
    infinite lattice, 62 on the 127-qubit one) after 0..trotter_steps layers."""
    import torch
    lattice = get_heavy_hex_ibm_eagle_lattice_infinite() if is_infinite else get_heavy_hex_ibm_eagle_lattice()
    traverser = list(lattice.get_traversal_iterator() or iter([]))
    target = 2 if is_infinite else 62
    per_regauge = lattice.edges_number * layers_per_regauging
    if colour_layers:
        per_regauge = 10 ** 9  # explicit regauges after each colour class
    emu = BPQuantumEmulator(lattice, key, max_chi, per_regauge, accuracy, max_belief_propagation_iterations,
                            synchronous_update, traversal_type, dtype=dtype or torch.complex128, chi=chi)
    mixing = rx(theta)
    edges = [el.id for el in traverser if isinstance(el, Edge)]
    classes: List[List[tuple]] = []
    if colour_layers:
        colours = edge_colouring(lattice)
        for c in range(max(colours.values()) + 1):
            classes.append([e for e in edges if colours[e] == c])
    dynamics = []
    dens = emu.dens_q1(target)
    dynamics.append(float((dens[0, 0] - dens[1, 1]).real))
    for _ in range(trotter_steps):
        for node_id in range(lattice.nodes_number):
            emu.apply_q1(mixing, node_id)
        if colour_layers:
            for cls in classes:
                for a, b in cls:
                    emu.apply_q2(RZZ, a, b)
                emu.regauge()
        else:
            for a, b in edges:
                emu.apply_q2(RZZ, a, b)
        dens = emu.dens_q1(target)
        dynamics.append(float((dens[0, 0] - dens[1, 1]).real))
    return dynamics


def run_eagle_simulations(thetas, max_chi: int, trotter_steps: int, accuracy: float, is_infinite: bool = False,
                          synchronous_update: bool = True, max_belief_propagation_iterations: Optional[int] = 100,
                          key=PRNGKey(42), traversal_type: str = "dfs", *, dtype=None,
                          chi: Optional[int] = None, sweep_events: Optional[list] = None) -> List[List[float]]:
    """The kicked-Ising dynamics of ``run_eagle_simulation`` (three colour
    layers, a regauge after each) for several RX angles at once: one
    instance per angle in one batched engine, every launch covering all
    instances.  Returns <Z> of the target qubit after 0..trotter_steps
    layers, per angle."""
    import torch
    from .emulator import BatchedBPEmulator
    lattice = get_heavy_hex_ibm_eagle_lattice_infinite() if is_infinite else get_heavy_hex_ibm_eagle_lattice()
    target = 2 if is_infinite else 62
    thetas = [float(t) for t in thetas]
    n = len(thetas)
    emu = BatchedBPEmulator([lattice] * n, [key] * n, max_chi, 10 ** 9, accuracy, max_belief_propagation_iterations,
                            synchronous_update, traversal_type, dtype=dtype or torch.complex64, chi=chi)
    # optional timing hook: (start, end, active-instance mask, sweeps) per BP sweep launch
    emu.sweep_events = sweep_events
    traverser = list(lattice.get_traversal_iterator() or iter([]))
    edges = [el.id for el in traverser if isinstance(el, Edge)]
    colours = edge_colouring(lattice)
    classes = [[e for e in edges if colours[e] == c] for c in range(max(colours.values()) + 1)]
    insts = list(range(n))

    def z_target():
        return [float((lambda d: (d[0, 0] - d[1, 1]).real)(emu.rdm1(i, [target])[0])) for i in insts]

    out = [z_target()]
    mixing = [rx(t) for t in thetas]
    for _ in range(trotter_steps):
        for i in insts:
            for node_id in range(lattice.nodes_number):
                emu.queue_q1(i, mixing[i], node_id)
        for cls in classes:
            for i in insts:
                for a, b in cls:
                    emu.apply_q2(i, RZZ, a, b)
            emu._regauge(insts)
        out.append(z_target())
    return [[out[k][i] for k in range(trotter_steps + 1)] for i in insts]


def annealing_instances(n_instances: int = 64, base_seed: int = 1000) -> List[EnergyFunction]:
    """Instance k: rng = default_rng(base_seed + k), N = 2 * rng.integers(50, 501)
    (even, so a 3-regular graph exists), the graph, +-1 couplings and
    N(0, 0.1^2) fields from random_regular seeded with base_seed + k."""
    out = []
    for k in range(n_instances):
        rng = np.random.default_rng(base_seed + k)
        n = 2 * int(rng.integers(50, 501))
        out.append(random_regular(n, 3, seed=base_seed + k, couplings="pm1", field_std=0.1))
    return out
