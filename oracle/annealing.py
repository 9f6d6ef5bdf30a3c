"""Trotterised transverse-field annealing on the BP emulator.

Restates examples/quantum_annealing/quantum_annealing/src/{quantum_annealer,
scheduler,energy_evaluation}.py.  The instance (pairs, couplings, fields) is an
input here; it is generated by the caller.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).
"""
from typing import Callable, Iterable, List, Sequence, Tuple

import numpy as np

from .emulator import BPQuantumEmulator
from .graph import graph_from_pairs

HADAMARD = np.sqrt(0.5) * np.array([[1, 1], [1, -1]], dtype=np.complex128)


def scheduler(dt: float, schedule: Callable[[float], float], steps: int) -> List[Tuple[float, float]]:
    """scheduler.py:7-14: (mixing time, coupling time) = dt * (f, 1 - f), f = schedule(i / (steps - 1))."""
    out = []
    for i in range(steps):
        f = schedule(i / (steps - 1))
        out.append((dt * f, dt * (1 - f)))
    return out


def coupling_gate(j: float, h_a: float, h_b: float, deg_a: int, deg_b: int, tau: float) -> np.ndarray:
    """quantum_annealer.py:44-69: diag exp(-i tau (J zz + h_a/d_a z1 + h_b/d_b 1z))."""
    ham = (j * np.array([1, -1, -1, 1], dtype=np.complex128)
           + (h_a / deg_a) * np.array([1, 1, -1, -1], dtype=np.complex128)
           + (h_b / deg_b) * np.array([1, -1, 1, -1], dtype=np.complex128))
    return np.diag(np.exp(-1j * tau * ham)).reshape((2, 2, 2, 2))


def mixing_gate(tau: float) -> np.ndarray:
    """quantum_annealer.py:72-84: exp(-i tau X)."""
    c, s = np.cos(tau), np.sin(tau)
    return np.array([[c, -1j * s], [-1j * s, c]], dtype=np.complex128)


def eval_energy(fields, pairs, couplings, config) -> float:
    """energy_evaluation.py:8-16: E = -h.s - sum J s_a s_b."""
    config = np.asarray(config, dtype=np.float64)
    pairs = np.asarray(pairs)
    return float(-np.dot(fields, config) - np.dot(config[pairs[:, 0]] * config[pairs[:, 1]], couplings))


def run_quantum_annealer(pairs, couplings, fields, schedule: Iterable[Tuple[float, float]], key, max_chi: int,
                         gates_number_per_regauging: int, accuracy: float, max_bp_iterations: int,
                         synchronous_update: bool, traversal_type: str, initial_messages=None,
                         two_site: bool = False):
    """quantum_annealer.py:90-152 (no sampling, no history).

    Returns (emulator, rho1 (n, 2, 2), rho2 (|E|, 4, 4) or None)."""
    n = len(fields)
    graph = graph_from_pairs(n, pairs)
    emu = BPQuantumEmulator(graph, key, max_chi, gates_number_per_regauging, accuracy, max_bp_iterations,
                            synchronous_update, traversal_type, initial_messages=initial_messages)
    degrees = [graph.nodes[i].degree for i in range(n)]
    for i in range(n):
        emu.apply_q1(HADAMARD, i)
    for mix_t, cpl_t in schedule:
        for (a, b), j in zip(pairs, couplings):
            a, b = int(a), int(b)
            emu.apply_q2(coupling_gate(j, fields[a], fields[b], degrees[a], degrees[b], cpl_t), a, b)
        mg = mixing_gate(mix_t)
        for i in range(n):
            emu.apply_q1(mg, i)
    rho1 = np.array([emu.dens_q1(i) for i in range(n)])
    rho2 = np.array([emu.dens_q2(int(a), int(b)) for a, b in pairs]) if two_site else None
    return emu, rho1, rho2


def bp_energy(rho1: np.ndarray, rho2: np.ndarray, fields, couplings) -> float:
    """BP-estimated energy -sum h <Z> - sum J <ZZ> from one- and two-site
    density matrices (sign convention of energy_evaluation.py:8-16)."""
    z = (rho1[:, 0, 0] - rho1[:, 1, 1]).real
    zz = (rho2[:, 0, 0] - rho2[:, 1, 1] - rho2[:, 2, 2] + rho2[:, 3, 3]).real
    return float(-np.dot(fields, z) - np.dot(couplings, zz))


def sign_energy(rho1: np.ndarray, pairs, couplings, fields) -> float:
    """qbp.py:125-128: energy of the configuration s_i = sign(rho_00 - 1/2)."""
    config = 2 * (rho1[:, 0, 0].real > 0.5) - 1
    return eval_energy(fields, pairs, couplings, config)
