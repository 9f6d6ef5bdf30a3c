"""Gradient of the BP-estimated annealing energy by a hand-written adjoint.

The forward run is the annealing flow of
examples/quantum_annealing/quantum_annealing/src/quantum_annealer.py:104-141
(Hadamards, then per step: regauge, coupling gates in pair order, mixing
gates; final regauge and E = -sum h <Z> - sum J <ZZ>, energy_evaluation.py
sign convention) with the schedule s(t) piecewise linear through ``knots``
(coupling time dt s(t_k), mixing time dt (1 - s(t_k))).  It runs in
complex128 with synchronous BP.  Each regauge either runs a fixed number of
sweeps from the diag(lambda) messages, or iterates BP to its fixed point
(message residual below ``bp_tol``), the loop of
dequantizer/src/quantum_emulator/bp_quantum_emulator.py:283-296 driven to
convergence.  Either way the energy is a smooth function of (knots, J).

Reverse mode, per step from the last: the forward kernels record what their
adjoints need (the simple-update and edge-SVD tapes, the sweep inputs or
the fixed point, the states before every gate level), then the adjoint
kernels run in reverse order: apply1 -> simple-update levels -> Vidal gauge
-> BP -> symmetric gauge.  Through BP, fixed sweeps are walked back one by
one and then into the diag(lambda) messages they started from; at a fixed
point m* = F(T, m*) the adjoint is the implicit one, the reverse iteration
mbar <- F_m^H mbar with every step's F_T^H mbar summed into the tensors'
gradient, run at m* until mbar vanishes (nothing reaches the starting
messages, which the fixed point does not depend on).  Records are kept for
one segment of steps at a time; earlier segments are recomputed from
checkpoints (``value_and_grad``).

Coupling gates are diagonal, G = sum_q |q><q| (x) diag(g_q0, g_q1): this
exact rank-2 split replaces the SVD split of _decompose_gate
(tensor_ops.py:202-218); the simple update depends on the gate only through
the gate itself, so results agree up to the singular-vector gauge and the
scale of the unnormalised lambdas (which the next regauge removes), and the
gradient of g_{q s} is the (s, s) entry of the second half's gradient.
"""
import warnings
from typing import Optional

import numpy as np
import torch

from . import _lib
from .annealing import EnergyFunction, tensor_graph_from
from .emulator import EPS, BatchedBPEmulator

DEVICE = "cuda"
HADAMARD = np.array([[1.0, 1.0], [1.0, -1.0]], dtype=np.complex128) / np.sqrt(2.0)
_ZZ = np.array([1.0, -1.0, -1.0, 1.0])
_Z1 = np.array([1.0, 1.0, -1.0, -1.0])
_Z2 = np.array([1.0, -1.0, 1.0, -1.0])


def _wide(x: torch.Tensor) -> torch.Tensor:
    """A forward record in complex128 for the adjoint kernels (a copy when
    the forward ran in complex64)."""
    return x if x.dtype != torch.complex64 else x.to(torch.complex128)


def knot_weights(n_knots: int, n_steps: int) -> np.ndarray:
    """(n_steps, n_knots) linear-interpolation weights: s(t_k) = W @ knots at
    t_k = k / (n_steps - 1), knots at linspace(0, 1, n_knots)."""
    tk = np.linspace(0.0, 1.0, n_knots)
    W = np.zeros((n_steps, n_knots))
    for k in range(n_steps):
        t = k / max(1, n_steps - 1)
        j = min(int(np.searchsorted(tk, t, side="right")) - 1, n_knots - 2)
        j = max(j, 0)
        w = (t - tk[j]) / (tk[j + 1] - tk[j])
        W[k, j] += 1.0 - w
        W[k, j + 1] += w
    return W


class AnnealingGradient:
    """value_and_grad(knots, J) -> (E, dE/dknots, dE/dJ) for one instance
    (``ef`` an EnergyFunction, ``key`` a PRNGKey) or for several run together
    in one engine under the same schedule (``ef`` and ``key`` sequences).

    ``cut``: backward cutoff of the simple-update and Vidal-gauge adjoints.
    Singular values, Cholesky pivots and message eigenvalues below ``cut``
    times the largest on their bond are frozen (no gradient flows through
    them); 0 gives the exact adjoint of the forward.  Through such
    directions the exact adjoint is a difference of terms that scale like
    1/sigma and w^-3/2 and cancel only in exact arithmetic, so on states
    with many near-zero Schmidt values (the first steps after the product
    state) rounding grows by orders of magnitude per step.  On the
    1000-spin, chi=16, 200-step run the usable window is about 3e-7..1e-6:
    at 1e-8 and below early-step gradients are rounding noise, at 1e-4 and
    above the frozen directions include the growing Schmidt values that carry
    the early-schedule gradient.

    ``fixed_sweeps``: BP sweeps per regauge; None iterates each regauge to
    the fixed point (mean Frobenius distance between successive message sets
    below ``bp_tol``) and differentiates it implicitly, the adjoint iteration
    stopping when its cotangent has fallen by ``bp_tol`` (both at most
    ``max_sweeps``; hitting the cap warns).  ``forward_sweeps`` and
    ``adjoint_iterations`` collect the counts of every regauge since they
    were last cleared.

    ``forward_dtype``: the precision of the forward's node tensors and
    messages (torch.complex128, or torch.complex64 for the complex64 engine
    of the annealing benchmark).  The adjoint always runs in complex128: the
    records of a complex64 forward (states, sweep inputs, fixed points, in
    complex64; the decompositions' tapes, in FP64 as always) are widened when
    the backward reads them, so the gradient is that of the FP64 operations
    evaluated along the complex64 trajectory."""

    def __init__(self, ef, key, chi: int, n_steps: int, dt: float = 0.1, n_knots: int = 32,
                 fixed_sweeps: Optional[int] = 10, traversal: str = "bfs", cut: float = 1e-6,
                 bp_tol: float = 1e-12, max_sweeps: int = 1000, forward_dtype: torch.dtype = torch.complex128):
        if not 0.0 <= cut < 1.0:
            raise ValueError(f"cut {cut} outside [0, 1)")
        self.batched = not isinstance(ef, EnergyFunction)
        efs = list(ef) if self.batched else [ef]
        keys = list(key) if self.batched else [key]
        if not efs or len(keys) != len(efs):
            raise ValueError(f"{len(efs)} energy functions and {len(keys)} keys")
        self.cut = float(cut)
        if forward_dtype not in (torch.complex64, torch.complex128):
            raise ValueError(f"forward_dtype {forward_dtype} is neither complex64 nor complex128")
        self.fdt = forward_dtype
        self.fcode = _lib.dtype_code(forward_dtype)
        self.ef, self.efs = efs[0], efs
        self.n_inst = len(efs)
        self.n_steps, self.dt, self.n_knots = int(n_steps), float(dt), int(n_knots)
        self.K = None if fixed_sweeps is None else int(fixed_sweeps)
        self.bp_tol, self.max_sweeps = float(bp_tol), int(max_sweeps)
        self.forward_sweeps, self.adjoint_iterations = [], []
        self.W = knot_weights(self.n_knots, self.n_steps)
        graphs = [tensor_graph_from(e) for e in efs]
        emu = BatchedBPEmulator(graphs, keys, chi, max(e.edges_number for e in efs), 1e-5, 100, True, traversal,
                                dtype=forward_dtype, chi=chi, fixed_sweeps=self.K)
        for inst, e in enumerate(efs):
            for i in range(e.nodes_number):
                emu.queue_q1(inst, HADAMARD, i)
        emu._flush()
        torch.cuda.synchronize()
        self.emu = emu
        self.h = emu.hs
        self.chi = self.h.chi
        self.topo = emu.topo
        self.E, self.N = self.topo.n_edges, self.topo.n_nodes
        self.gam0 = emu.gam.clone()
        self.lam0 = emu.lam.clone()
        # coupled pairs of every instance, concatenated in instance order
        plan = emu.make_plan([(inst, int(a), int(b)) for inst, e in enumerate(efs) for a, b in e.coupled_spin_pairs],
                             [])
        self.launches = [(o2, c2) for (o2, c2, _o1, _c1) in plan.launches if c2]
        self.order2 = plan.order2  # launch position -> pair index
        meta = plan.meta.copy()
        meta[:, 2] = 2
        self.d_meta = torch.from_numpy(meta).to(DEVICE)
        self.pairs = np.concatenate([np.asarray(e.coupled_spin_pairs, dtype=np.int64) for e in efs])
        self.pair_inst = np.concatenate([np.full(e.edges_number, i) for i, e in enumerate(efs)])
        self.spin_inst = np.concatenate([np.full(e.nodes_number, i) for i, e in enumerate(efs)])
        self.fields = np.concatenate([np.asarray(e.fields, dtype=np.float64) for e in efs])
        ha, hb = [], []
        for e in efs:
            pr = np.asarray(e.coupled_spin_pairs, dtype=np.int64)
            deg = np.asarray(e.node_degrees, dtype=np.float64)
            ha.append(e.fields[pr[:, 0]] / deg[pr[:, 0]])
            hb.append(e.fields[pr[:, 1]] / deg[pr[:, 1]])
        self.ha, self.hb = np.concatenate(ha), np.concatenate(hb)
        self.d_nodes_all = torch.arange(self.N, dtype=torch.int32, device=DEVICE)
        self.node_inst = np.asarray(self.topo.node_inst)  # instance of each engine node (mixing-gate order)
        self.rdm_nodes = torch.tensor([self.topo.node_index[inst][i] for inst, e in enumerate(efs)
                                       for i in range(e.nodes_number)], dtype=torch.int32, device=DEVICE)
        self.rdm_edges = torch.tensor([self.topo.edge_index[inst][(int(a), int(b))] for inst, e in enumerate(efs)
                                       for a, b in e.coupled_spin_pairs], dtype=torch.int32, device=DEVICE)
        self.n_spins, self.n_pairs = len(self.spin_inst), len(self.pair_inst)
        L = _lib.lib()
        self.su_tape_bytes = int(L.dq_su_tape_bytes(self.h.ptr, 2))
        self.svd_tape_bytes = int(L.dq_edge_svd_tape_bytes(self.h.ptr)) * self.E
        self.cc = self.chi * self.chi

    # ------------------------------------------------------------ gates
    def _diag_phases(self, tau_c: float, J: np.ndarray) -> np.ndarray:
        """(E, 4) diagonal entries g_c of the coupling gates, c = 2 s_a + s_b."""
        theta = J[:, None] * _ZZ + self.ha[:, None] * _Z1 + self.hb[:, None] * _Z2
        return np.exp(-1j * tau_c * theta), theta

    def _halves(self, g: np.ndarray) -> torch.Tensor:
        """Structured halves in launch order: h1_q = |q><q|, h2_q = diag(g_q0, g_q1)."""
        n = len(g)
        halves = np.zeros((n, 2, 16), dtype=np.complex128)
        for q in range(2):
            halves[:, 0, q * 4 + q * 2 + q] = 1.0
            halves[:, 1, q * 4 + 0] = g[:, 2 * q + 0]
            halves[:, 1, q * 4 + 3] = g[:, 2 * q + 1]
        return torch.from_numpy(halves[self.order2]).to(DEVICE)

    @staticmethod
    def _mixing(tau: float) -> np.ndarray:
        c, s = np.cos(tau), np.sin(tau)
        return np.array([[c, -1j * s], [-1j * s, c]], dtype=np.complex128)

    # ---------------------------------------------------------- forward
    def _regauge(self, gam, lam, record: bool):
        L, h, st = _lib.lib(), self.h, _lib.stream_handle()
        fc = self.fcode
        ts = torch.empty_like(gam)
        _lib.check(L.dq_symmetric_gauge(h.ptr, fc, _lib.ptr(gam), _lib.ptr(lam), EPS, _lib.ptr(ts), st))
        n2e = torch.zeros(2 * self.E * self.cc, dtype=self.fdt, device=DEVICE)
        e2n = torch.zeros_like(n2e)
        _lib.check(L.dq_lambdas_to_messages(h.ptr, fc, _lib.ptr(lam), _lib.ptr(n2e), _lib.ptr(e2n), st))
        inputs = []
        if self.K is not None:
            for _ in range(self.K):
                inputs.append(e2n)
                nn, ne = torch.empty_like(n2e), torch.empty_like(n2e)
                _lib.check(L.dq_bp_sweep(h.ptr, fc, _lib.ptr(ts), _lib.ptr(n2e), _lib.ptr(e2n),
                                         _lib.ptr(nn), _lib.ptr(ne), None, st))
                n2e, e2n = nn, ne
        else:
            n2e, e2n, fixed_point, sweeps = self._converge(ts, n2e, e2n)
            inputs = [fixed_point]
            self.forward_sweeps.append(sweeps)
        gam_new = torch.empty_like(gam)
        lam_new = torch.empty_like(lam)
        X = torch.empty(2 * self.E * self.cc, dtype=torch.complex128, device=DEVICE)
        tape = torch.empty(self.svd_tape_bytes, dtype=torch.uint8, device=DEVICE)
        _lib.check(L.dq_vidal_gauge_tape(h.ptr, fc, _lib.ptr(ts), _lib.ptr(n2e), EPS, _lib.ptr(gam_new),
                                         _lib.ptr(lam_new), _lib.ptr(X), _lib.ptr(tape), st))
        rec = (gam, lam, ts, inputs, X, tape) if record else None
        return gam_new, lam_new, rec

    def _converge(self, ts, n2e, e2n):
        """Synchronous sweeps until the mean Frobenius distance between
        successive message sets is below bp_tol; returns the last messages,
        the last sweep's input e2n (the fixed point the adjoint linearises
        at) and the sweep count."""
        L, h, st = _lib.lib(), self.h, _lib.stream_handle()
        dist = torch.empty(1, dtype=torch.float64, device=DEVICE)
        for it in range(1, self.max_sweeps + 1):
            nn, ne = torch.empty_like(n2e), torch.empty_like(n2e)
            _lib.check(L.dq_bp_sweep(h.ptr, self.fcode, _lib.ptr(ts), _lib.ptr(n2e), _lib.ptr(e2n), _lib.ptr(nn),
                                     _lib.ptr(ne), None, st))
            _lib.check(L.dq_messages_distance(h.ptr, self.fcode, _lib.ptr(nn), _lib.ptr(ne), _lib.ptr(n2e),
                                              _lib.ptr(e2n), _lib.ptr(dist), st))
            prev, n2e, e2n = e2n, nn, ne
            if float(dist) < self.bp_tol:
                return n2e, e2n, prev, it
        warnings.warn(f"BP did not reach the residual {self.bp_tol:g} in {self.max_sweeps} sweeps "
                      f"(last {float(dist):.3g})")
        return n2e, e2n, prev, self.max_sweeps

    def _step(self, gam, lam, k: int, knots, J, record: bool):
        L, h, st = _lib.lib(), self.h, _lib.stream_handle()
        s = float(self.W[k] @ knots)
        tau_c, tau_m = self.dt * s, self.dt * (1.0 - s)
        rec = {"regauge": None, "levels": [], "tau": (tau_c, tau_m)}
        if k >= 1:
            gam, lam, rec["regauge"] = self._regauge(gam, lam, record)
        g, _ = self._diag_phases(tau_c, J)
        halves = self._halves(g)
        err = torch.zeros(self.E, dtype=torch.float64, device=DEVICE)
        ent = torch.zeros_like(err)
        for o2, c2 in self.launches:
            tape = torch.empty(c2 * self.su_tape_bytes, dtype=torch.uint8, device=DEVICE)
            pre = (gam.clone(), lam.clone()) if record else None
            _lib.check(L.dq_simple_update_tape(h.ptr, self.fcode, _lib.ptr(gam), _lib.ptr(lam),
                                               _lib.ptr(self.d_meta[o2:]), c2, _lib.ptr(halves[o2:]), 2, EPS,
                                               self.chi, _lib.ptr(err[o2:]), _lib.ptr(ent[o2:]), _lib.ptr(tape), st))
            if record:
                rec["levels"].append((o2, c2, pre, tape))
        if record:
            rec["halves"] = halves
            rec["pre_mix"] = gam.clone()
        mg = torch.from_numpy(np.broadcast_to(self._mixing(tau_m), (self.N, 2, 2)).copy()).to(DEVICE)
        _lib.check(L.dq_apply1(h.ptr, self.fcode, _lib.ptr(gam), _lib.ptr(self.d_nodes_all), self.N, _lib.ptr(mg),
                               st))
        if record:
            rec["mix"] = mg
        return gam, lam, rec

    def _observables(self, gam, lam):
        L, h, st = _lib.lib(), self.h, _lib.stream_handle()
        fc = self.fcode
        gam, lam, rec = self._regauge(gam, lam, True)
        ts = torch.empty_like(gam)
        _lib.check(L.dq_symmetric_gauge(h.ptr, fc, _lib.ptr(gam), _lib.ptr(lam), EPS, _lib.ptr(ts), st))
        n_d = torch.zeros(2 * self.E * self.cc, dtype=self.fdt, device=DEVICE)
        e_d = torch.zeros_like(n_d)
        _lib.check(L.dq_lambdas_to_messages(h.ptr, fc, _lib.ptr(lam), _lib.ptr(n_d), _lib.ptr(e_d), st))
        n1, n2 = self.n_spins, self.n_pairs
        r1 = torch.empty((n1, 2, 2), dtype=torch.complex128, device=DEVICE)
        r2 = torch.empty((n2, 4, 4), dtype=torch.complex128, device=DEVICE)
        _lib.check(L.dq_rdm1(h.ptr, fc, _lib.ptr(ts), _lib.ptr(e_d), _lib.ptr(self.rdm_nodes), n1,
                             _lib.ptr(r1), st))
        _lib.check(L.dq_rdm2(h.ptr, fc, _lib.ptr(ts), _lib.ptr(e_d), _lib.ptr(self.rdm_edges), n2,
                             _lib.ptr(r2), st))
        return (gam, lam, ts, rec), r1, r2

    def _energy(self, r1, r2, J):
        """Energy of every instance (a float for a single one) and the <ZZ>
        of every coupled pair."""
        z = (r1[:, 0, 0] - r1[:, 1, 1]).real
        zz = (r2[:, 0, 0] - r2[:, 1, 1] - r2[:, 2, 2] + r2[:, 3, 3]).real
        e = (-np.bincount(self.spin_inst, self.fields * z, minlength=self.n_inst)
             - np.bincount(self.pair_inst, J * zz, minlength=self.n_inst))
        return (e if self.batched else float(e[0])), zz

    def _couplings(self, J) -> np.ndarray:
        """Couplings of every pair, concatenated over instances (J: None for
        the energy functions' own, an array, or one array per instance)."""
        if J is None:
            return np.concatenate([np.asarray(e.coupling_amplitudes, dtype=np.float64) for e in self.efs])
        if self.batched:
            J = np.concatenate([np.asarray(j, dtype=np.float64).reshape(-1) for j in J])
        J = np.asarray(J, dtype=np.float64).reshape(-1)
        if J.shape[0] != self.n_pairs:
            raise ValueError(f"{J.shape[0]} couplings for {self.n_pairs} coupled pairs")
        return J

    def value(self, knots, J=None):
        """E(knots, J): a float, or one energy per instance when batched."""
        knots = np.asarray(knots, dtype=np.float64)
        J = self._couplings(J)
        gam, lam = self.gam0.clone(), self.lam0.clone()
        for k in range(self.n_steps):
            gam, lam, _ = self._step(gam, lam, k, knots, J, False)
        _, r1, r2 = self._observables(gam, lam)
        return self._energy(r1.cpu().numpy(), r2.cpu().numpy(), J)[0]

    # --------------------------------------------------------- backward
    def _regauge_vjp(self, rec, gbar, lbar):
        """Gradient of the state before a regauge from that after it."""
        L, h, st = _lib.lib(), self.h, _lib.stream_handle()
        gam, lam, ts, inputs, X, tape = rec
        gam, ts, inputs = _wide(gam), _wide(ts), [_wide(x) for x in inputs]
        ts_bar = torch.zeros_like(ts)
        n_bar = torch.zeros(2 * self.E * self.cc, dtype=torch.complex128, device=DEVICE)
        X_bar = torch.empty_like(X)
        _lib.check(L.dq_vidal_gauge_vjp(h.ptr, _lib.DQ_C128, _lib.ptr(ts), _lib.ptr(X), _lib.ptr(tape), EPS, self.cut,
                                        _lib.ptr(gbar), _lib.ptr(lbar), _lib.ptr(X_bar), _lib.ptr(ts_bar),
                                        _lib.ptr(n_bar), st))
        # synchronous sweep i (1..K): n2e_i = F(T, e2n_{i-1}) and, the edge
        # update reading the sweep's input, e2n_i = swap(n2e_{i-1}); walk both
        # gradients back to the diag(lambda) messages of sweep 0
        e_bar = None
        if self.K is not None:
            for i in range(self.K, 0, -1):
                new_e = None
                if n_bar is not None:
                    new_e = torch.empty_like(X)
                    _lib.check(L.dq_bp_sweep_vjp(h.ptr, _lib.DQ_C128, _lib.ptr(ts), _lib.ptr(inputs[i - 1]),
                                                 _lib.ptr(n_bar), _lib.ptr(new_e), _lib.ptr(ts_bar), st))
                new_n = None if e_bar is None else e_bar.view(self.E, 2, self.cc).flip(1).reshape(-1).contiguous()
                n_bar, e_bar = new_n, new_e
        else:
            # implicit adjoint at the fixed point: n2e = F(T, e2n), e2n =
            # swap(n2e), so the cotangent of n2e maps to itself through
            # swap(F_e2n^H .) while F_T^H . accumulates into ts_bar
            fp = inputs[0]
            ref = float(torch.linalg.vector_norm(n_bar))
            its = 0
            while its < self.max_sweeps and ref > 0.0:
                e_in = torch.empty_like(X)
                _lib.check(L.dq_bp_sweep_vjp(h.ptr, _lib.DQ_C128, _lib.ptr(ts), _lib.ptr(fp), _lib.ptr(n_bar),
                                             _lib.ptr(e_in), _lib.ptr(ts_bar), st))
                n_bar = e_in.view(self.E, 2, self.cc).flip(1).reshape(-1).contiguous()
                its += 1
                if float(torch.linalg.vector_norm(n_bar)) <= self.bp_tol * ref:
                    break
            else:
                if ref > 0.0:
                    warnings.warn(f"BP adjoint did not fall by {self.bp_tol:g} in {self.max_sweeps} iterations")
            self.adjoint_iterations.append(its)
            n_bar = None
        gam_bar = torch.zeros_like(gam)
        lam_bar = torch.zeros_like(lam)
        if n_bar is not None or e_bar is not None:
            _lib.check(L.dq_lambdas_to_messages_vjp(h.ptr, _lib.DQ_C128,
                                                    None if n_bar is None else _lib.ptr(n_bar),
                                                    None if e_bar is None else _lib.ptr(e_bar), _lib.ptr(lam_bar),
                                                    st))
        _lib.check(L.dq_symmetric_gauge_vjp(h.ptr, _lib.DQ_C128, _lib.ptr(gam), _lib.ptr(lam), EPS, _lib.ptr(ts_bar),
                                            _lib.ptr(gam_bar), _lib.ptr(lam_bar), st))
        return gam_bar, lam_bar

    def record_bytes(self) -> int:
        """Device bytes of one step's forward record: the state before the
        regauge, the symmetric tensors, the K sweep inputs, the gauge's X and
        tape, the state before every gate level with the level's tape, the
        state before the mixing gates and the gate halves."""
        state = self.gam0.numel() * self.gam0.element_size() + self.lam0.numel() * 8
        msg = 2 * self.E * self.cc * self.gam0.element_size()
        levels = len(self.launches) * state + sum(c2 for _, c2 in self.launches) * self.su_tape_bytes
        sweeps = 2 if self.K is None else self.K + 1
        return 3 * state + sweeps * msg + self.svd_tape_bytes + levels + self.E * 2 * 16 * 16

    def segment_length(self, memory_gb: float) -> int:
        """Steps per recomputation segment: the largest s for which the
        checkpoints (one state per segment) and the records of one segment
        fit in ``memory_gb``."""
        state = self.gam0.numel() * self.gam0.element_size() + self.lam0.numel() * 8
        rec, budget, n = self.record_bytes(), memory_gb * 2 ** 30, self.n_steps
        best = 0
        for s in range(1, n + 1):
            if -(-n // s) * state + s * rec <= budget:
                best = s
        if best == 0:
            need = (n * state + rec) / 2 ** 30
            raise ValueError(f"memory_gb={memory_gb} holds neither the checkpoints nor one step's records "
                             f"({need:.1f} GB at one step per segment)")
        return best

    def value_and_grad(self, knots, J=None, trace: Optional[list] = None, memory_gb: float = 24.0,
                       segment_steps: Optional[int] = None):
        """(E, dE/dknots, dE/dJ).  Batched: E has one entry per instance,
        dE/dknots is (instances, knots) (each instance's energy against the
        shared schedule) and dE/dJ one array per instance, all from one
        backward.  trace (optional list) receives (step, |Gamma-bar|,
        |lambda-bar|) after each step's adjoint.

        Memory is bounded by ``memory_gb`` (records and checkpoints; the
        engine's own buffers come on top): the run is cut into segments of
        ``segment_length(memory_gb)`` steps.  The forward keeps the state at
        the start of each segment and the records of the last segment
        (~1 GB per step at 1000 spins, chi=16); the backward recomputes each
        earlier segment from its checkpoint with records on, then walks it
        back.  Cost: one forward, one recomputation of all but the last
        segment, one adjoint.  A budget that holds every step's records gives
        one segment and no recomputation.  ``segment_steps`` sets the segment
        length directly instead."""
        knots = np.asarray(knots, dtype=np.float64)
        J = self._couplings(J)
        L, h, st = _lib.lib(), self.h, _lib.stream_handle()
        n = self.n_steps
        seg = self.segment_length(memory_gb) if segment_steps is None else int(segment_steps)
        if not 1 <= seg <= n:
            raise ValueError(f"segment_steps {seg} outside [1, {n}]")
        last_start = ((n - 1) // seg) * seg
        self.segment_steps = seg
        checkpoints = {}
        records = [None] * n
        gam, lam = self.gam0.clone(), self.lam0.clone()
        for k in range(n):
            if k % seg == 0 and k < last_start:
                checkpoints[k] = (gam, lam)
            keep = k >= last_start
            gam, lam, rec = self._step(gam.clone(), lam.clone(), k, knots, J, keep)
            records[k] = rec if keep else None
        (gf, lf, tsf, reg_f), r1, r2 = self._observables(gam, lam)
        gf, tsf = _wide(gf), _wide(tsf)
        r1h, r2h = r1.cpu().numpy(), r2.cpu().numpy()
        energy, zz = self._energy(r1h, r2h, J)
        # energy -> density matrices -> final symmetric tensors and lambdas
        rho1_bar = np.zeros_like(r1h)
        rho1_bar[:, 0, 0] = -self.fields
        rho1_bar[:, 1, 1] = self.fields
        rho2_bar = np.zeros_like(r2h)
        for c, sgn in enumerate(_ZZ):
            rho2_bar[:, c, c] = -J * sgn
        d1, d2 = torch.from_numpy(rho1_bar).to(DEVICE), torch.from_numpy(rho2_bar).to(DEVICE)
        ts_bar = torch.zeros_like(tsf)
        lam_bar = torch.zeros_like(lf)
        _lib.check(L.dq_rdm1_vjp(h.ptr, _lib.DQ_C128, _lib.ptr(tsf), _lib.ptr(lf), _lib.ptr(self.rdm_nodes),
                                 self.n_spins, _lib.ptr(d1), _lib.ptr(ts_bar), _lib.ptr(lam_bar), st))
        _lib.check(L.dq_rdm2_vjp(h.ptr, _lib.DQ_C128, _lib.ptr(tsf), _lib.ptr(lf), _lib.ptr(self.rdm_edges),
                                 self.n_pairs, _lib.ptr(d2), _lib.ptr(ts_bar), _lib.ptr(lam_bar), st))
        gam_bar = torch.zeros_like(gf)
        _lib.check(L.dq_symmetric_gauge_vjp(h.ptr, _lib.DQ_C128, _lib.ptr(gf), _lib.ptr(lf), EPS, _lib.ptr(ts_bar),
                                            _lib.ptr(gam_bar), _lib.ptr(lam_bar), st))
        gam_bar, lam_bar = self._regauge_vjp(reg_f, gam_bar, lam_bar)
        dJ = -zz.copy()  # explicit -J <ZZ> term
        # per instance: the instances share the schedule but not the state
        dtau_c = np.zeros((self.n_inst, self.n_steps))
        dtau_m = np.zeros((self.n_inst, self.n_steps))
        for k in range(n - 1, -1, -1):
            if records[k] is None:
                # recompute the segment ending at step k from its checkpoint
                start = (k // seg) * seg
                g, l = checkpoints.pop(start)
                for q in range(start, k + 1):
                    g, l, records[q] = self._step(g.clone(), l.clone(), q, knots, J, True)
                del g, l
            rec = records[k]
            records[k] = None
            tau_c, tau_m = rec["tau"]
            # mixing gates
            mg_bar = torch.empty_like(rec["mix"])
            pre_mix = _wide(rec["pre_mix"])
            _lib.check(L.dq_apply1_vjp(h.ptr, _lib.DQ_C128, _lib.ptr(pre_mix), _lib.ptr(self.d_nodes_all),
                                       self.N, _lib.ptr(rec["mix"]), _lib.ptr(gam_bar), _lib.ptr(mg_bar), st))
            c, s = np.cos(tau_m), np.sin(tau_m)
            dG = np.array([[-s, -1j * c], [-1j * c, -s]])
            mg = mg_bar.cpu().numpy()
            per_node = np.real(np.einsum("nij,ij->n", np.conj(mg), dG))
            dtau_m[:, k] = np.bincount(self.node_inst, per_node, minlength=self.n_inst)
            # coupling gates, levels in reverse
            hb_all = torch.zeros((self.n_pairs, 2, 16), dtype=torch.complex128, device=DEVICE)
            for o2, c2, (pg, pl), tape in reversed(rec["levels"]):
                pg = _wide(pg)
                _lib.check(L.dq_simple_update_vjp(h.ptr, _lib.DQ_C128, _lib.ptr(pg), _lib.ptr(pl),
                                                  _lib.ptr(self.d_meta[o2:]), c2, _lib.ptr(rec["halves"][o2:]), 2,
                                                  EPS, self.cut, _lib.ptr(tape), _lib.ptr(gam_bar),
                                                  _lib.ptr(lam_bar),
                                                  _lib.ptr(hb_all[o2:]), st))
            hb = np.empty((self.n_pairs, 2, 16), dtype=np.complex128)
            hb[self.order2] = hb_all.cpu().numpy()
            gbar = np.stack([hb[:, 1, 0], hb[:, 1, 3], hb[:, 1, 4], hb[:, 1, 7]], axis=1)  # (E, 4): g_00, g_01, g_10, g_11
            g, theta = self._diag_phases(tau_c, J)
            # g_c = exp(-i tau theta_c): dg/dtau = -i theta g, dg/dJ = -i tau zz_c g
            per_pair = np.real(np.sum(np.conj(gbar) * (-1j * theta * g), axis=1))
            dtau_c[:, k] = np.bincount(self.pair_inst, per_pair, minlength=self.n_inst)
            dJ += np.real(np.sum(np.conj(gbar) * (-1j * tau_c * _ZZ[None, :] * g), axis=1))
            if rec["regauge"] is not None:
                gam_bar, lam_bar = self._regauge_vjp(rec["regauge"], gam_bar, lam_bar)
            if trace is not None:
                trace.append((k, float(torch.linalg.vector_norm(gam_bar)), float(torch.linalg.vector_norm(lam_bar))))
        dknots = (self.dt * dtau_c - self.dt * dtau_m) @ self.W
        if not self.batched:
            return energy, dknots[0], dJ
        cuts = np.cumsum([e.edges_number for e in self.efs])[:-1]
        return energy, dknots, np.split(dJ, cuts)


class AnnealingEnergy(torch.autograd.Function):
    """torch.autograd wrapper: E = AnnealingEnergy.apply(runner, knots, J) with
    knots (K,) and J (E,) float64 tensors; backward uses the hand-written
    adjoint of ``runner`` (an AnnealingGradient)."""

    @staticmethod
    def forward(ctx, runner: AnnealingGradient, knots: torch.Tensor, J: torch.Tensor):
        if runner.batched:
            raise ValueError("AnnealingEnergy takes a single-instance runner")
        e, dk, dj = runner.value_and_grad(knots.detach().cpu().numpy(), J.detach().cpu().numpy())
        ctx.save_for_backward(torch.from_numpy(dk).to(knots), torch.from_numpy(dj).to(J))
        return torch.tensor(e, dtype=torch.float64, device=knots.device)

    @staticmethod
    def backward(ctx, grad_out):
        dk, dj = ctx.saved_tensors
        return None, grad_out * dk, grad_out * dj
