"""Trotterised transverse-field quantum annealing on the device BP emulator.

Keeps the driver semantics of examples/quantum_annealing/quantum_annealing/
src/{quantum_annealer,scheduler,energy_evaluation}.py: graph from the coupled
pairs with unit bonds (quantum_annealer.py:34-41), Hadamard on every qubit
(:115-116), then per step the coupling gates diag exp(-i tau_c (J zz +
h_a/d_a z1 + h_b/d_b 1z)) on every edge in pair order (:44-69, :121-124)
followed by exp(-i tau_m X) on every qubit (:72-84, :125-126), with
(tau_m, tau_c) = dt (f, 1 - f), f = schedule(i / (steps - 1)) (scheduler.py:7-14).

Several instances run in one batched emulator (one launch per gate level
across all instances).  Two-site density matrices give the BP-estimated
energy -sum h <Z> - sum J <ZZ> (sign convention of energy_evaluation.py:8-16).
"""
from dataclasses import dataclass
from typing import Callable, Iterable, List, Optional, Sequence, Tuple, Union

import numpy as np
import torch

from .emulator import BatchedBPEmulator
from .graph import TensorGraph

HADAMARD = np.sqrt(0.5) * np.array([[1, 1], [1, -1]], dtype=np.complex128)


@dataclass
class EnergyFunction:
    """Ising energy E(s) = -sum_i h_i s_i - sum_(ij) J_ij s_i s_j
    (energy_function.py:18-63)."""

    coupling_amplitudes: np.ndarray
    coupled_spin_pairs: np.ndarray
    fields: np.ndarray

    @property
    def nodes_number(self) -> int:
        return int(self.fields.shape[0])

    @property
    def node_degrees(self) -> List[int]:
        deg = np.zeros(self.nodes_number, dtype=np.int64)
        np.add.at(deg, self.coupled_spin_pairs[:, 0], 1)
        np.add.at(deg, self.coupled_spin_pairs[:, 1], 1)
        return deg.tolist()

    @property
    def edges_number(self) -> int:
        return int(self.coupled_spin_pairs.shape[0])


def random_regular(nodes_number: int, degree: int = 3, seed: int = 0, couplings: str = "pm1",
                   field_std: float = 0.1, coupling_std: float = 1.0) -> EnergyFunction:
    """Random d-regular Ising instance.  Graph from networkx.random_regular_graph
    (as energy_function.py:218-238); couplings +-1 uniform ("pm1") or
    N(0, coupling_std^2) ("normal"); fields N(0, field_std^2)."""
    import networkx as nx

    rng = np.random.default_rng(seed)
    graph = nx.random_regular_graph(degree, nodes_number, seed=int(rng.integers(0, 2**31 - 1)))
    pairs = np.asarray(list(graph.edges), dtype=np.int64).reshape(-1, 2)
    if couplings == "pm1":
        j = rng.choice([-1.0, 1.0], size=len(pairs))
    else:
        j = rng.normal(0.0, coupling_std, size=len(pairs))
    h = rng.normal(0.0, field_std, size=nodes_number)
    return EnergyFunction(j, pairs, h)


def get_scheduler(total_time_step_size: float, schedule: Callable[[float], float],
                  steps_number: int) -> List[Tuple[float, float]]:
    """(mixing time, coupling time) per step (scheduler.py:7-14)."""
    out = []
    for i in range(steps_number):
        f = schedule(i / (steps_number - 1))
        out.append((total_time_step_size * f, total_time_step_size * (1 - f)))
    return out


def linear_schedule(t: float) -> float:
    """The reference configs' schedule string 'lambda t: 1 - t'."""
    return 1.0 - t


def knot_schedule(knots: Sequence[float]) -> Callable[[float], float]:
    """Piecewise-linear s(t) through knots at t_k = k / (K - 1); the mixing
    fraction is 1 - s(t)."""
    knots = np.asarray(knots, dtype=np.float64)
    tk = np.linspace(0.0, 1.0, len(knots))
    return lambda t: 1.0 - float(np.interp(t, tk, knots))


def _kempe_edge_colouring(n_nodes: int, pairs: List[Tuple[int, int]], k: int,
                          rng: np.random.Generator) -> Optional[np.ndarray]:
    """One attempt at a proper edge colouring with k colours: edges in random
    order take a colour free at both ends; when the free colours differ
    (a free at u, b free at v) the a/b alternating path from v is swapped,
    which frees a at v unless the path ends at u.  None when that happens."""
    at = [dict() for _ in range(n_nodes)]  # node -> colour -> edge
    col = np.full(len(pairs), -1, dtype=np.int64)
    for idx in rng.permutation(len(pairs)):
        u, v = pairs[idx]
        fu = [c for c in range(k) if c not in at[u]]
        fv = [c for c in range(k) if c not in at[v]]
        if not fu or not fv:
            return None
        common = [c for c in fu if c in fv]
        if common:
            c = common[rng.integers(len(common))]
        else:
            c = -1
            for a in fu:
                for b in fv:
                    path, x, cur = [], v, a
                    while cur in at[x]:
                        e = at[x][cur]
                        path.append(e)
                        x = pairs[e][1] if pairs[e][0] == x else pairs[e][0]
                        cur = b if cur == a else a
                    if x == u:
                        continue
                    for e in path:
                        del at[pairs[e][0]][col[e]]
                        del at[pairs[e][1]][col[e]]
                    for e in path:
                        col[e] = b if col[e] == a else a
                        at[pairs[e][0]][col[e]] = e
                        at[pairs[e][1]][col[e]] = e
                    c = a
                    break
                if c >= 0:
                    break
            if c < 0:
                return None
        col[idx] = c
        at[u][c] = idx
        at[v][c] = idx
    return col


def edge_colouring(n_nodes: int, pairs, restarts: int = 400, seed: int = 0) -> np.ndarray:
    """Proper edge colouring of a simple graph with few colours: max-degree
    colours (Kempe-chain attempts in seeded random orders; random 3-regular
    graphs are 3-edge-colourable almost surely, and a few percent of attempts
    succeed), else max-degree + 1, else greedy (at most 2 max-degree - 1)."""
    pairs = [(int(a), int(b)) for a, b in pairs]
    if not pairs:
        return np.zeros(0, dtype=np.int64)
    deg = np.bincount(np.asarray(pairs).reshape(-1), minlength=n_nodes)
    rng = np.random.default_rng(seed)
    for k, tries in ((int(deg.max()), restarts), (int(deg.max()) + 1, 20)):
        for _ in range(tries):
            col = _kempe_edge_colouring(n_nodes, pairs, k, rng)
            if col is not None:
                return col
    used = [set() for _ in range(n_nodes)]
    col = np.zeros(len(pairs), dtype=np.int64)
    for i, (a, b) in enumerate(pairs):
        c = 0
        while c in used[a] or c in used[b]:
            c += 1
        col[i] = c
        used[a].add(c)
        used[b].add(c)
    return col


def edge_colored(energy_function: EnergyFunction) -> EnergyFunction:
    """The same instance with its coupled pairs sorted by a proper edge
    colouring with few colours (pairs of one colour share no spin; original
    order kept inside a colour; ``edge_colouring``).  Applying the coupling
    gates in this pair order is the reference driver run on the reordered
    pair list: every colour class is one launch level (3 on random
    3-regular graphs) instead of the ~16 levels the generator's order needs."""
    ef = energy_function
    pairs = ef.coupled_spin_pairs
    colour = edge_colouring(ef.nodes_number, pairs)
    order = np.argsort(colour, kind="stable")
    return EnergyFunction(ef.coupling_amplitudes[order], pairs[order], ef.fields)


def tensor_graph_from(energy_function: EnergyFunction) -> TensorGraph:
    """quantum_annealer.py:34-41."""
    g = TensorGraph()
    for _ in range(energy_function.nodes_number):
        g.add_node()
    for a, b in energy_function.coupled_spin_pairs:
        g.add_edge((int(a), int(b)), 1)
    return g


def coupling_gates(ef: EnergyFunction, tau: float) -> np.ndarray:
    """(E, 2, 2, 2, 2) diagonal coupling gates (quantum_annealer.py:44-69)."""
    deg = np.asarray(ef.node_degrees, dtype=np.float64)
    a, b = ef.coupled_spin_pairs[:, 0], ef.coupled_spin_pairs[:, 1]
    ham = (ef.coupling_amplitudes[:, None] * np.array([1, -1, -1, 1.0])
           + (ef.fields[a] / deg[a])[:, None] * np.array([1, 1, -1, -1.0])
           + (ef.fields[b] / deg[b])[:, None] * np.array([1, -1, 1, -1.0]))
    diag = np.exp(-1j * tau * ham)
    gates = np.zeros((len(a), 4, 4), dtype=np.complex128)
    gates[:, np.arange(4), np.arange(4)] = diag
    return gates.reshape(-1, 2, 2, 2, 2)


def mixing_gate(tau: float) -> np.ndarray:
    """exp(-i tau X) (quantum_annealer.py:72-84)."""
    c, s = np.cos(tau), np.sin(tau)
    return np.array([[c, -1j * s], [-1j * s, c]], dtype=np.complex128)


def eval_energy(energy_function: EnergyFunction, config) -> float:
    """energy_evaluation.py:8-16."""
    ef = energy_function
    config = np.asarray(config, dtype=np.float64)
    pairs = ef.coupled_spin_pairs
    return float(-np.dot(ef.fields, config) - np.dot(config[pairs[:, 0]] * config[pairs[:, 1]],
                                                       ef.coupling_amplitudes))


@dataclass
class QuantumAnnealingResults:
    """quantum_annealer.py:23-31 plus the two-site matrices and BP energy."""

    configuration: Optional[np.ndarray]
    density_matrices: np.ndarray
    density_matrices_history: Optional[List[np.ndarray]]
    vidal_distances_after_regauging: np.ndarray
    truncation_affected_vidal_distances: np.ndarray
    truncation_errors: np.ndarray
    entropies: np.ndarray
    two_site_density_matrices: Optional[np.ndarray] = None
    bp_energy: Optional[float] = None
    sign_energy: Optional[float] = None
    bp_iterations: Optional[List[int]] = None


def bp_energy(ef: EnergyFunction, rho1: np.ndarray, rho2: np.ndarray) -> float:
    """-sum h <Z> - sum J <ZZ> from one- and two-site density matrices."""
    z = (rho1[:, 0, 0] - rho1[:, 1, 1]).real
    zz = (rho2[:, 0, 0] - rho2[:, 1, 1] - rho2[:, 2, 2] + rho2[:, 3, 3]).real
    return float(-np.dot(ef.fields, z) - np.dot(ef.coupling_amplitudes, zz))


class Annealer:
    """Batched annealing runner: one BatchedBPEmulator over all instances and
    a launch plan (gate levels) computed once and reused every step."""

    def __init__(self, energy_functions: Sequence[EnergyFunction], keys, max_chi: int,
                 layers_per_regauging: int = 1, accuracy: float = 1e-5, max_belief_propagation_iterations: int = 100,
                 synchronous_update: bool = False, traversal_type: str = "dfs",
                 dtype: torch.dtype = torch.complex64, chi: Optional[int] = None,
                 fixed_sweeps: Optional[int] = None):
        self.efs = list(energy_functions)
        self.graphs = [tensor_graph_from(ef) for ef in self.efs]
        per = [ef.edges_number * layers_per_regauging for ef in self.efs]
        self.emu = BatchedBPEmulator(self.graphs, keys, max_chi, per, accuracy, max_belief_propagation_iterations,
                                     synchronous_update, traversal_type, dtype=dtype,
                                     chi=chi if chi is not None else max_chi, fixed_sweeps=fixed_sweeps)
        for inst in range(len(self.efs)):
            for i in range(self.efs[inst].nodes_number):
                self.emu.queue_q1(inst, HADAMARD, i)
        self.steps_done = 0
        # per-step op pattern: coupling gates in pair order, then mixing gates
        two = [(inst, int(a), int(b)) for inst, ef in enumerate(self.efs) for a, b in ef.coupled_spin_pairs]
        one = [(inst, i) for inst, ef in enumerate(self.efs) for i in range(ef.nodes_number)]
        self._aligned = all(ef.edges_number * layers_per_regauging == per_i
                            for ef, per_i in zip(self.efs, per)) and layers_per_regauging == 1
        self.plan = self.emu.make_plan(two, one) if self._aligned else None

    def step(self, mixing_time: float, coupling_time: float):
        """One Trotter step on every instance."""
        if self.plan is not None:
            # regauge points of the reference fall on the first gate of the step
            due = [i for i in range(len(self.efs)) if self.emu.gates_since[i] >= self.emu.per_regauge[i]]
            self.emu._regauge(due)
            gates = np.concatenate([coupling_gates(ef, coupling_time) for ef in self.efs])
            # coupling gates are diagonal: operator-Schmidt rank <= phys = 2
            self.emu.run_plan(self.plan, gates, mixing_gate(mixing_time), rank_bound=2)
        else:
            layer = []
            for ef in self.efs:
                gates = coupling_gates(ef, coupling_time)
                layer.append([(gates[k], int(a), int(b)) for k, (a, b) in enumerate(ef.coupled_spin_pairs)])
            self.emu.apply_layer(layer)
            mg = mixing_gate(mixing_time)
            for inst, ef in enumerate(self.efs):
                for i in range(ef.nodes_number):
                    self.emu.queue_q1(inst, mg, i)
        self.steps_done += 1

    def finish(self, two_site: bool = True) -> List[QuantumAnnealingResults]:
        d1 = self.emu.all_rdm1()
        d2 = self.emu.all_rdm2() if two_site else None
        rho1 = d1.cpu().numpy()
        rho2 = d2.cpu().numpy() if two_site else None
        topo = self.emu.topo
        out = []
        for inst, ef in enumerate(self.efs):
            # packing order of nodes follows the traversal; map back to node ids
            idx = np.asarray([topo.node_index[inst][i] for i in range(ef.nodes_number)])
            r1 = rho1[idx]
            r2 = None
            if rho2 is not None:
                g = self.graphs[inst]
                eidx = np.asarray([topo.edge_index[inst][(int(a), int(b))] for a, b in ef.coupled_spin_pairs])
                r2 = rho2[eidx]
            config = 2 * (r1[:, 0, 0].real > 0.5) - 1
            out.append(QuantumAnnealingResults(
                configuration=None, density_matrices=r1, density_matrices_history=None,
                vidal_distances_after_regauging=self.emu.after_regauging_vidal_distances(inst),
                truncation_affected_vidal_distances=self.emu.truncated_affected_vidal_distances(inst),
                truncation_errors=np.asarray(self.emu.truncation_errors(inst)),
                entropies=np.asarray(self.emu.entropies(inst)),
                two_site_density_matrices=r2,
                bp_energy=bp_energy(ef, r1, r2) if r2 is not None else None,
                sign_energy=eval_energy(ef, config),
                bp_iterations=list(self.emu.bp_iterations[inst])))
        return out


def run_quantum_annealer(energy_function: EnergyFunction, scheduler: Iterable[Tuple[float, float]], key,
                         max_chi: int, gates_number_per_regauging: int, accuracy: float,
                         max_belief_propagation_iteration: int, synchronous_update: bool, traversal_type: str,
                         sample_measurements: bool = False, record_history: bool = False,
                         dtype: torch.dtype = torch.complex128) -> QuantumAnnealingResults:
    """Single-instance driver with the reference signature
    (quantum_annealer.py:90-152)."""
    ef = energy_function
    if gates_number_per_regauging % max(1, ef.edges_number):
        raise NotImplementedError("gates_number_per_regauging must be a multiple of the edge count")
    lpr = max(1, gates_number_per_regauging // max(1, ef.edges_number))
    ann = Annealer([ef], [key], max_chi, lpr, accuracy, max_belief_propagation_iteration, synchronous_update,
                   traversal_type, dtype=dtype, chi=None)
    history = [] if record_history else None
    for mix, cpl in scheduler:
        ann.step(mix, cpl)
        if record_history:
            history.append(ann.emu.rdm1(0, list(range(ef.nodes_number))).cpu().numpy())
    res = ann.finish(two_site=True)[0]
    res.density_matrices_history = history
    if sample_measurements:
        results = [ann.emu.measure(0, i) for i in range(ef.nodes_number)]
        res.configuration = 1 - 2 * np.asarray(results)
    return res
