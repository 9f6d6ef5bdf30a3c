"""Gradient of the BP-estimated annealing energy by a hand-written adjoint.

The forward run is the annealing flow of
examples/quantum_annealing/quantum_annealing/src/quantum_annealer.py:104-141
(Hadamards, then per step: regauge, coupling gates in pair order, mixing
gates; final regauge and E = -sum h <Z> - sum J <ZZ>, energy_evaluation.py
sign convention) with the schedule s(t) piecewise linear through ``knots``
(coupling time dt s(t_k), mixing time dt (1 - s(t_k))).  It runs in
complex128 with synchronous BP and a fixed number of sweeps per regauge, so
the energy is a smooth function of (knots, J) and the adjoint is exact.

Reverse mode, per step from the last: the step is recomputed from its
checkpoint with the forward kernels recording what their adjoints need (the
simple-update and edge-SVD tapes, the sweep inputs, the states before every
gate level), then the adjoint kernels run in reverse order:
apply1 -> simple-update levels -> Vidal gauge -> BP sweeps -> diag(lambda)
messages -> symmetric gauge.  Checkpoints are kept in device memory (one
state per step).

Coupling gates are diagonal, G = sum_q |q><q| (x) diag(g_q0, g_q1): this
exact rank-2 split replaces the SVD split of _decompose_gate
(tensor_ops.py:202-218); the simple update depends on the gate only through
the gate itself, so results agree up to the singular-vector gauge and the
scale of the unnormalised lambdas (which the next regauge removes), and the
gradient of g_{q s} is the (s, s) entry of the second half's gradient.
"""
from typing import Optional, Sequence, Tuple

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
    """value_and_grad(knots, J) -> (E, dE/dknots, dE/dJ) for one instance.

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
    the early-schedule gradient."""

    def __init__(self, ef: EnergyFunction, key, chi: int, n_steps: int, dt: float = 0.1, n_knots: int = 32,
                 fixed_sweeps: int = 10, traversal: str = "bfs", cut: float = 1e-6):
        if not 0.0 <= cut < 1.0:
            raise ValueError(f"cut {cut} outside [0, 1)")
        self.cut = float(cut)
        self.ef = ef
        self.n_steps, self.dt, self.n_knots, self.K = int(n_steps), float(dt), int(n_knots), int(fixed_sweeps)
        self.W = knot_weights(self.n_knots, self.n_steps)
        graph = tensor_graph_from(ef)
        emu = BatchedBPEmulator([graph], [key], chi, ef.edges_number, 1e-5, 100, True, traversal,
                                dtype=torch.complex128, chi=chi, fixed_sweeps=self.K)
        for i in range(ef.nodes_number):
            emu.queue_q1(0, HADAMARD, i)
        emu._flush()
        torch.cuda.synchronize()
        self.emu = emu
        self.h = emu.hs
        self.chi = self.h.chi
        self.topo = emu.topo
        self.E, self.N = self.topo.n_edges, self.topo.n_nodes
        self.gam0 = emu.gam.clone()
        self.lam0 = emu.lam.clone()
        plan = emu.make_plan([(0, int(a), int(b)) for a, b in ef.coupled_spin_pairs], [])
        self.launches = [(o2, c2) for (o2, c2, _o1, _c1) in plan.launches if c2]
        self.order2 = plan.order2  # launch position -> pair index
        meta = plan.meta.copy()
        meta[:, 2] = 2
        self.d_meta = torch.from_numpy(meta).to(DEVICE)
        self.pairs = np.asarray(ef.coupled_spin_pairs, dtype=np.int64)
        deg = np.asarray(ef.node_degrees, dtype=np.float64)
        self.ha = ef.fields[self.pairs[:, 0]] / deg[self.pairs[:, 0]]
        self.hb = ef.fields[self.pairs[:, 1]] / deg[self.pairs[:, 1]]
        nidx = self.topo.node_index[0]
        self.d_nodes_all = torch.arange(self.N, dtype=torch.int32, device=DEVICE)
        self.rdm_nodes = torch.tensor([nidx[i] for i in range(ef.nodes_number)], dtype=torch.int32, device=DEVICE)
        eidx = self.topo.edge_index[0]
        self.rdm_edges = torch.tensor([eidx[(int(a), int(b))] for a, b in self.pairs], dtype=torch.int32,
                                      device=DEVICE)
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
        ts = torch.empty_like(gam)
        _lib.check(L.dq_symmetric_gauge(h.ptr, _lib.DQ_C128, _lib.ptr(gam), _lib.ptr(lam), EPS, _lib.ptr(ts), st))
        n2e = torch.zeros(2 * self.E * self.cc, dtype=torch.complex128, device=DEVICE)
        e2n = torch.zeros_like(n2e)
        _lib.check(L.dq_lambdas_to_messages(h.ptr, _lib.DQ_C128, _lib.ptr(lam), _lib.ptr(n2e), _lib.ptr(e2n), st))
        inputs = []
        for _ in range(self.K):
            inputs.append(e2n)
            nn, ne = torch.empty_like(n2e), torch.empty_like(n2e)
            _lib.check(L.dq_bp_sweep(h.ptr, _lib.DQ_C128, _lib.ptr(ts), _lib.ptr(n2e), _lib.ptr(e2n), _lib.ptr(nn),
                                     _lib.ptr(ne), None, st))
            n2e, e2n = nn, ne
        gam_new = torch.empty_like(gam)
        lam_new = torch.empty_like(lam)
        X = torch.empty(2 * self.E * self.cc, dtype=torch.complex128, device=DEVICE)
        tape = torch.empty(self.svd_tape_bytes, dtype=torch.uint8, device=DEVICE)
        _lib.check(L.dq_vidal_gauge_tape(h.ptr, _lib.DQ_C128, _lib.ptr(ts), _lib.ptr(n2e), EPS, _lib.ptr(gam_new),
                                         _lib.ptr(lam_new), _lib.ptr(X), _lib.ptr(tape), st))
        rec = (gam, lam, ts, inputs, X, tape) if record else None
        return gam_new, lam_new, rec

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
            _lib.check(L.dq_simple_update_tape(h.ptr, _lib.DQ_C128, _lib.ptr(gam), _lib.ptr(lam),
                                               _lib.ptr(self.d_meta[o2:]), c2, _lib.ptr(halves[o2:]), 2, EPS,
                                               self.chi, _lib.ptr(err[o2:]), _lib.ptr(ent[o2:]), _lib.ptr(tape), st))
            if record:
                rec["levels"].append((o2, c2, pre, tape))
        if record:
            rec["halves"] = halves
            rec["pre_mix"] = gam.clone()
        mg = torch.from_numpy(np.broadcast_to(self._mixing(tau_m), (self.N, 2, 2)).copy()).to(DEVICE)
        _lib.check(L.dq_apply1(h.ptr, _lib.DQ_C128, _lib.ptr(gam), _lib.ptr(self.d_nodes_all), self.N, _lib.ptr(mg),
                               st))
        if record:
            rec["mix"] = mg
        return gam, lam, rec

    def _observables(self, gam, lam):
        L, h, st = _lib.lib(), self.h, _lib.stream_handle()
        gam, lam, rec = self._regauge(gam, lam, True)
        ts = torch.empty_like(gam)
        _lib.check(L.dq_symmetric_gauge(h.ptr, _lib.DQ_C128, _lib.ptr(gam), _lib.ptr(lam), EPS, _lib.ptr(ts), st))
        n_d = torch.zeros(2 * self.E * self.cc, dtype=torch.complex128, device=DEVICE)
        e_d = torch.zeros_like(n_d)
        _lib.check(L.dq_lambdas_to_messages(h.ptr, _lib.DQ_C128, _lib.ptr(lam), _lib.ptr(n_d), _lib.ptr(e_d), st))
        n1, n2 = self.ef.nodes_number, self.ef.edges_number
        r1 = torch.empty((n1, 2, 2), dtype=torch.complex128, device=DEVICE)
        r2 = torch.empty((n2, 4, 4), dtype=torch.complex128, device=DEVICE)
        _lib.check(L.dq_rdm1(h.ptr, _lib.DQ_C128, _lib.ptr(ts), _lib.ptr(e_d), _lib.ptr(self.rdm_nodes), n1,
                             _lib.ptr(r1), st))
        _lib.check(L.dq_rdm2(h.ptr, _lib.DQ_C128, _lib.ptr(ts), _lib.ptr(e_d), _lib.ptr(self.rdm_edges), n2,
                             _lib.ptr(r2), st))
        return (gam, lam, ts, rec), r1, r2

    def _energy(self, r1, r2, J):
        z = (r1[:, 0, 0] - r1[:, 1, 1]).real
        zz = (r2[:, 0, 0] - r2[:, 1, 1] - r2[:, 2, 2] + r2[:, 3, 3]).real
        return float(-np.dot(self.ef.fields, z) - np.dot(J, zz)), zz

    def value(self, knots, J=None) -> float:
        knots = np.asarray(knots, dtype=np.float64)
        J = np.asarray(self.ef.coupling_amplitudes if J is None else J, dtype=np.float64)
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
        for i in range(self.K, 0, -1):
            new_e = None
            if n_bar is not None:
                new_e = torch.empty_like(X)
                _lib.check(L.dq_bp_sweep_vjp(h.ptr, _lib.DQ_C128, _lib.ptr(ts), _lib.ptr(inputs[i - 1]),
                                             _lib.ptr(n_bar), _lib.ptr(new_e), _lib.ptr(ts_bar), st))
            new_n = None if e_bar is None else e_bar.view(self.E, 2, self.cc).flip(1).reshape(-1).contiguous()
            n_bar, e_bar = new_n, new_e
        gam_bar = torch.zeros_like(gam)
        lam_bar = torch.zeros_like(lam)
        _lib.check(L.dq_lambdas_to_messages_vjp(h.ptr, _lib.DQ_C128, None if n_bar is None else _lib.ptr(n_bar),
                                                None if e_bar is None else _lib.ptr(e_bar), _lib.ptr(lam_bar), st))
        _lib.check(L.dq_symmetric_gauge_vjp(h.ptr, _lib.DQ_C128, _lib.ptr(gam), _lib.ptr(lam), EPS, _lib.ptr(ts_bar),
                                            _lib.ptr(gam_bar), _lib.ptr(lam_bar), st))
        return gam_bar, lam_bar

    def value_and_grad(self, knots, J=None, trace: Optional[list] = None,
                       record_fraction: float = 0.6) -> Tuple[float, np.ndarray, np.ndarray]:
        """trace (optional list) receives (step, |Gamma-bar|, |lambda-bar|) after
        each step's adjoint.  The forward keeps the records of as many steps
        (tapes, sweep inputs, states before every gate level; ~0.9 GB per step
        at 1000 spins, chi=16) as fit in ``record_fraction`` of the free device
        memory; the backward recomputes the other steps from their
        checkpoints."""
        knots = np.asarray(knots, dtype=np.float64)
        J = np.asarray(self.ef.coupling_amplitudes if J is None else J, dtype=np.float64)
        L, h, st = _lib.lib(), self.h, _lib.stream_handle()
        # forward with one checkpoint per step, and the records that fit
        checkpoints = []
        records = [None] * self.n_steps
        # free device memory plus what torch's allocator holds cached but unused
        free = torch.cuda.mem_get_info()[0] + torch.cuda.memory_reserved() - torch.cuda.memory_allocated()
        budget = record_fraction * free
        used, per_step = 0, None
        gam, lam = self.gam0.clone(), self.lam0.clone()
        for k in range(self.n_steps):
            checkpoints.append((gam, lam))
            keep = per_step is None or used + per_step <= budget
            before = torch.cuda.memory_allocated()
            gam, lam, rec = self._step(gam.clone(), lam.clone(), k, knots, J, keep)
            if keep:
                grown = torch.cuda.memory_allocated() - before
                if k >= 1:  # a step with a regauge
                    per_step = max(per_step or 0, grown)
                used += grown
                records[k] = rec
        (gf, lf, tsf, reg_f), r1, r2 = self._observables(gam, lam)
        r1h, r2h = r1.cpu().numpy(), r2.cpu().numpy()
        energy, zz = self._energy(r1h, r2h, J)
        # energy -> density matrices -> final symmetric tensors and lambdas
        rho1_bar = np.zeros_like(r1h)
        rho1_bar[:, 0, 0] = -self.ef.fields
        rho1_bar[:, 1, 1] = self.ef.fields
        rho2_bar = np.zeros_like(r2h)
        for c, sgn in enumerate(_ZZ):
            rho2_bar[:, c, c] = -J * sgn
        d1, d2 = torch.from_numpy(rho1_bar).to(DEVICE), torch.from_numpy(rho2_bar).to(DEVICE)
        ts_bar = torch.zeros_like(tsf)
        lam_bar = torch.zeros_like(lf)
        _lib.check(L.dq_rdm1_vjp(h.ptr, _lib.DQ_C128, _lib.ptr(tsf), _lib.ptr(lf), _lib.ptr(self.rdm_nodes),
                                 self.ef.nodes_number, _lib.ptr(d1), _lib.ptr(ts_bar), _lib.ptr(lam_bar), st))
        _lib.check(L.dq_rdm2_vjp(h.ptr, _lib.DQ_C128, _lib.ptr(tsf), _lib.ptr(lf), _lib.ptr(self.rdm_edges),
                                 self.ef.edges_number, _lib.ptr(d2), _lib.ptr(ts_bar), _lib.ptr(lam_bar), st))
        gam_bar = torch.zeros_like(gf)
        _lib.check(L.dq_symmetric_gauge_vjp(h.ptr, _lib.DQ_C128, _lib.ptr(gf), _lib.ptr(lf), EPS, _lib.ptr(ts_bar),
                                            _lib.ptr(gam_bar), _lib.ptr(lam_bar), st))
        gam_bar, lam_bar = self._regauge_vjp(reg_f, gam_bar, lam_bar)
        dJ = -zz.copy()  # explicit -J <ZZ> term
        dtau_c = np.zeros(self.n_steps)
        dtau_m = np.zeros(self.n_steps)
        for k in range(self.n_steps - 1, -1, -1):
            rec = records[k]
            records[k] = None
            if rec is None:
                gam_k, lam_k = checkpoints[k]
                _, _, rec = self._step(gam_k.clone(), lam_k.clone(), k, knots, J, True)
            tau_c, tau_m = rec["tau"]
            # mixing gates
            mg_bar = torch.empty_like(rec["mix"])
            _lib.check(L.dq_apply1_vjp(h.ptr, _lib.DQ_C128, _lib.ptr(rec["pre_mix"]), _lib.ptr(self.d_nodes_all),
                                       self.N, _lib.ptr(rec["mix"]), _lib.ptr(gam_bar), _lib.ptr(mg_bar), st))
            c, s = np.cos(tau_m), np.sin(tau_m)
            dG = np.array([[-s, -1j * c], [-1j * c, -s]])
            dtau_m[k] = float(np.real(np.vdot(mg_bar.sum(dim=0).cpu().numpy(), dG)))
            # coupling gates, levels in reverse
            hb_all = torch.zeros((self.E, 2, 16), dtype=torch.complex128, device=DEVICE)
            for o2, c2, (pg, pl), tape in reversed(rec["levels"]):
                _lib.check(L.dq_simple_update_vjp(h.ptr, _lib.DQ_C128, _lib.ptr(pg), _lib.ptr(pl),
                                                  _lib.ptr(self.d_meta[o2:]), c2, _lib.ptr(rec["halves"][o2:]), 2,
                                                  EPS, self.cut, _lib.ptr(tape), _lib.ptr(gam_bar),
                                                  _lib.ptr(lam_bar),
                                                  _lib.ptr(hb_all[o2:]), st))
            hb = np.empty((self.E, 2, 16), dtype=np.complex128)
            hb[self.order2] = hb_all.cpu().numpy()
            gbar = np.stack([hb[:, 1, 0], hb[:, 1, 3], hb[:, 1, 4], hb[:, 1, 7]], axis=1)  # (E, 4): g_00, g_01, g_10, g_11
            g, theta = self._diag_phases(tau_c, J)
            # g_c = exp(-i tau theta_c): dg/dtau = -i theta g, dg/dJ = -i tau zz_c g
            dtau_c[k] = float(np.real(np.sum(np.conj(gbar) * (-1j * theta * g))))
            dJ += np.real(np.sum(np.conj(gbar) * (-1j * tau_c * _ZZ[None, :] * g), axis=1))
            if rec["regauge"] is not None:
                gam_bar, lam_bar = self._regauge_vjp(rec["regauge"], gam_bar, lam_bar)
            if trace is not None:
                trace.append((k, float(torch.linalg.vector_norm(gam_bar)), float(torch.linalg.vector_norm(lam_bar))))
        dknots = self.W.T @ (self.dt * dtau_c - self.dt * dtau_m)
        return energy, dknots, dJ


class AnnealingEnergy(torch.autograd.Function):
    """torch.autograd wrapper: E = AnnealingEnergy.apply(runner, knots, J) with
    knots (K,) and J (E,) float64 tensors; backward uses the hand-written
    adjoint of ``runner`` (an AnnealingGradient)."""

    @staticmethod
    def forward(ctx, runner: AnnealingGradient, knots: torch.Tensor, J: torch.Tensor):
        e, dk, dj = runner.value_and_grad(knots.detach().cpu().numpy(), J.detach().cpu().numpy())
        ctx.save_for_backward(torch.from_numpy(dk).to(knots), torch.from_numpy(dj).to(J))
        return torch.tensor(e, dtype=torch.float64, device=knots.device)

    @staticmethod
    def backward(ctx, grad_out):
        dk, dj = ctx.saved_tensors
        return None, grad_out * dk, grad_out * dj
