"""Exact state-vector emulation on the GPU.

``ExactQuantumEmulator`` mirrors
dequantizer/src/quantum_emulator/exact_quantum_emulator.py:11-147 (same
constructor, methods, qubit order, measurement rule and key handling).
``run_exact_quantum_annealer`` mirrors
examples/quantum_annealing/quantum_annealing/src/exact_quantum_annealer.py:74-122,
whose state object comes from the external Rust package ``qem`` (not part of
the reference tree): its measurement follows the exact emulator's rule
(exact_quantum_emulator.py:126-147) with the uniform sample passed in.

The state is one device complex128 vector of 2^n amplitudes; every update is
a libdq ``dq_sv_*`` launch.  A coupling layer (any number of diagonal
two-qubit gates) is one pass over the vector, a layer of single-qubit gates
one pass per 12 qubits; an annealing layer (up to 64 couplings, then the
mixing gates) rides the coupling phases on the first of those passes.
"""
from dataclasses import dataclass
from typing import Callable, Dict, Iterable, List, Optional, Sequence, Tuple

import numpy as np
import torch

from . import _lib
from . import random as dqr
from .annealing import HADAMARD, EnergyFunction, coupling_gates, mixing_gate
from .emulator import QuantumEmulator
from .graph import Node, TensorGraph

DEVICE = "cuda"
MAX_QUBITS = 40


def _dev(arr: np.ndarray) -> torch.Tensor:
    return torch.from_numpy(np.array(arr, dtype=np.complex128, order="C", copy=True)).to(DEVICE)


class StateVector:
    """2^n complex128 amplitudes on the GPU.  Qubit q (q = 0 is the most
    significant, the order of the reference's reshape to n axes of 2) is
    amplitude-index bit n - 1 - q."""

    def __init__(self, n_qubits: int):
        n = int(n_qubits)
        if n < 1 or n > MAX_QUBITS:
            raise ValueError(f"n_qubits = {n} outside [1, {MAX_QUBITS}]")
        L = _lib.lib()
        need = 16 << n
        free, _ = torch.cuda.mem_get_info()
        if need > free:
            raise MemoryError(f"a {n}-qubit state needs {need / 2**30:.1f} GiB; {free / 2**30:.1f} GiB free")
        self.n = n
        self.psi = torch.empty(1 << n, dtype=torch.complex128, device=DEVICE)
        self.work = torch.empty(int(L.dq_sv_work_bytes()), dtype=torch.uint8, device=DEVICE)
        self._rho = torch.zeros((n, 2, 2), dtype=torch.complex128, device=DEVICE)
        self.set_basis(0)

    def _bit(self, q: int) -> int:
        q = int(q)
        if not 0 <= q < self.n:
            raise ValueError(f"qubit {q} outside [0, {self.n})")
        return self.n - 1 - q

    def _mask(self, qubits: Optional[Iterable[int]]) -> int:
        if qubits is None:
            return (1 << self.n) - 1
        m = 0
        for q in qubits:
            m |= 1 << self._bit(q)
        return m

    def set_basis(self, index: int):
        _lib.check(_lib.lib().dq_sv_set_basis(_lib.ptr(self.psi), self.n, int(index), _lib.stream_handle()))

    def apply1(self, q: int, gate):
        g = _dev(np.asarray(gate).reshape(2, 2))
        _lib.check(_lib.lib().dq_sv_apply1(_lib.ptr(self.psi), self.n, self._bit(q), _lib.ptr(g),
                                           _lib.stream_handle()))

    def apply2(self, q1: int, q2: int, gate):
        """gate[o1][o2][i1][i2]; index 1 on qubit q1, index 2 on qubit q2."""
        g = _dev(np.asarray(gate).reshape(2, 2, 2, 2))
        _lib.check(_lib.lib().dq_sv_apply2(_lib.ptr(self.psi), self.n, self._bit(q1), self._bit(q2), _lib.ptr(g),
                                           _lib.stream_handle()))

    def _terms(self, pairs, diagonals):
        pairs = np.asarray(pairs, dtype=np.int64).reshape(-1, 2)
        diagonals = np.asarray(diagonals, dtype=np.complex128).reshape(-1, 4)
        if pairs.shape[0] != diagonals.shape[0]:
            raise ValueError("one diagonal of 4 entries per pair")
        if np.any(pairs[:, 0] == pairs[:, 1]):
            raise ValueError("a two-qubit term needs two distinct qubits")
        bits = np.array([[self._bit(a), self._bit(b)] for a, b in pairs], dtype=np.int32).reshape(-1)
        return int(pairs.shape[0]), torch.from_numpy(bits).to(DEVICE), _dev(diagonals)

    def _gates(self, gates) -> torch.Tensor:
        gates = np.asarray(gates, dtype=np.complex128)
        if gates.shape == (2, 2):
            gates = np.broadcast_to(gates, (self.n, 2, 2))
        if gates.shape != (self.n, 2, 2):
            raise ValueError(f"gates must be (2, 2) or ({self.n}, 2, 2)")
        return _dev(gates[::-1])  # bit b = qubit n - 1 - b

    def apply_diagonal(self, pairs, diagonals):
        """Diagonal two-qubit gates: term k multiplies the amplitude by
        diagonals[k][2 x1 + x2], x1 = value of qubit pairs[k][0], x2 of
        pairs[k][1].  One pass over the vector for all terms."""
        k, tb, td = self._terms(pairs, diagonals)
        if k == 0:
            return
        _lib.check(_lib.lib().dq_sv_apply_diagonal(_lib.ptr(self.psi), self.n, k, _lib.ptr(tb), _lib.ptr(td),
                                                   _lib.ptr(self.work), _lib.stream_handle()))

    def apply_diagonal_layer1(self, pairs, diagonals, gates, qubits: Optional[Sequence[int]] = None):
        """apply_diagonal, then apply1_layer: one annealing layer (with at
        least 12 qubits and at most 64 terms, in one pass fewer)."""
        k, tb, td = self._terms(pairs, diagonals)
        g = self._gates(gates)
        _lib.check(_lib.lib().dq_sv_apply_diagonal_layer1(_lib.ptr(self.psi), self.n, k, _lib.ptr(tb), _lib.ptr(td),
                                                          self._mask(qubits), _lib.ptr(g), _lib.ptr(self.work),
                                                          _lib.stream_handle()))

    def apply1_layer(self, gates, qubits: Optional[Sequence[int]] = None):
        """gates: one 2x2 gate for all listed qubits, or an (n, 2, 2) array
        (gate of qubit q at [q]); qubits: None = all."""
        g = self._gates(gates)
        _lib.check(_lib.lib().dq_sv_apply1_layer(_lib.ptr(self.psi), self.n, self._mask(qubits), _lib.ptr(g),
                                                 _lib.stream_handle()))

    def rdm1(self, qubits: Optional[Sequence[int]] = None) -> torch.Tensor:
        """One-qubit density matrices rho[s][s'] (device, (len, 2, 2))."""
        qs = list(range(self.n)) if qubits is None else [int(q) for q in qubits]
        _lib.check(_lib.lib().dq_sv_rdm1(_lib.ptr(self.psi), self.n, self._mask(qs), _lib.ptr(self._rho),
                                         _lib.ptr(self.work), _lib.stream_handle()))
        idx = torch.tensor([self._bit(q) for q in qs], dtype=torch.long, device=DEVICE)
        return self._rho.index_select(0, idx)

    def measure(self, q: int, uniform_sample: float) -> int:
        """Projective Z measurement of qubit q with a given uniform sample
        (exact_quantum_emulator.py:126-147): 0 if rho[0][0] > sample; the
        state is projected and renormalised."""
        dens = self.rdm1([q])[0].cpu().numpy()
        p = float(dens[0, 0].real)
        result, norm = (0, p) if p > float(uniform_sample) else (1, 1.0 - p)
        proj = np.zeros((2, 2), dtype=np.complex128)
        proj[result, result] = 1.0 / np.sqrt(norm)
        self.apply1(q, proj)
        return result


class ExactQuantumEmulator(QuantumEmulator):
    """exact_quantum_emulator.py:11-147: qubit order = node order of the
    graph's default traversal (:24-31)."""

    def __init__(self, tensor_graph: TensorGraph, key):
        super().__init__(tensor_graph, dqr.as_key(key))
        node2pos: Dict = {}
        for el in tensor_graph.get_traversal_iterator() or iter([]):
            if isinstance(el, Node):
                node2pos[el.id] = len(node2pos)
        self._node2pos = node2pos
        self._sv = StateVector(tensor_graph.nodes_number)

    @property
    def state(self) -> torch.Tensor:
        """The device amplitudes (2^n complex128, reference index order)."""
        return self._sv.psi

    def set_to_product_state(self, state: Dict):
        """exact_quantum_emulator.py:37-49."""
        n = self._sv.n
        shift = 0
        for node_id, pos in self._node2pos.items():
            if state[node_id] == 1:
                shift += 2 ** (n - pos - 1)
        self._sv.set_basis(shift)

    def apply_q2(self, gate, controlling_node_id, controlled_node_id):
        """exact_quantum_emulator.py:51-85: gate (o1, o2, i1, i2)."""
        self._sv.apply2(self._node2pos[controlling_node_id], self._node2pos[controlled_node_id], gate)

    def apply_q1(self, gate, node_id):
        """exact_quantum_emulator.py:87-103: gate (out, in)."""
        self._sv.apply1(self._node2pos[node_id], gate)

    def dens_q1(self, node_id) -> np.ndarray:
        """exact_quantum_emulator.py:105-124."""
        dens = self._sv.rdm1([self._node2pos[node_id]])[0].cpu().numpy()
        tr = np.trace(dens)
        if not np.isclose(tr, 1.0, 1e-8):
            raise AssertionError(f"trace of the density matrix is {tr}")
        return dens

    def measure(self, node_id) -> int:
        """exact_quantum_emulator.py:126-147 (the key is split twice per
        measurement, :135-136 and :145-146)."""
        key, sub = dqr.split(self.prng_key)
        self.prng_key = key
        sample = dqr.generator(sub).uniform()
        result = self._sv.measure(self._node2pos[node_id], sample)
        key, _ = dqr.split(self.prng_key)
        self.prng_key = key
        return result


@dataclass
class ExactQuantumAnnealingResults:
    """exact_quantum_annealer.py:24-28."""

    configuration: Optional[np.ndarray]
    density_matrices: np.ndarray
    density_matrices_history: Optional[List[np.ndarray]]


def run_exact_quantum_annealer(energy_function: EnergyFunction, scheduler: Iterable[Tuple[float, float]], key,
                               sample_measurements: bool, record_history: bool,
                               state_callback: Optional[Callable[[StateVector], None]] = None
                               ) -> ExactQuantumAnnealingResults:
    """exact_quantum_annealer.py:74-122 on the GPU state vector: Hadamards on
    every spin, then per layer the coupling gates (one diagonal pass) and the
    mixing gates (one single-qubit layer); qubit i = spin i."""
    ef = energy_function
    n = ef.nodes_number
    sv = StateVector(n)
    sv.apply1_layer(HADAMARD)
    history: Optional[List[np.ndarray]] = [] if record_history else None
    pairs = np.asarray(ef.coupled_spin_pairs, dtype=np.int64).reshape(-1, 2)
    for mixing_time, coupling_time in scheduler:
        diag = np.diagonal(coupling_gates(ef, coupling_time).reshape(-1, 4, 4), axis1=1, axis2=2) if pairs.shape[0] \
            else np.zeros((0, 4), dtype=np.complex128)
        sv.apply_diagonal_layer1(pairs, diag, mixing_gate(mixing_time))
        if state_callback is not None:
            state_callback(sv)
        if history is not None:
            history.append(sv.rdm1().cpu().numpy())
    dens = sv.rdm1().cpu().numpy()
    config = None
    if sample_measurements:
        results = []
        for q in range(n):
            key, sub = dqr.split(key)
            results.append(sv.measure(q, dqr.generator(sub).uniform()))
        config = 1 - 2 * np.asarray(results)
    return ExactQuantumAnnealingResults(config, dens, history)
