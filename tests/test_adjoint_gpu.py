"""GPU checks of the hand-written adjoints against central finite differences
of the forward kernels (complex128).

Losses are chosen invariant under the simple update's singular-vector phases
(the gauge a network state does not see): the contracted two-site tensor,
the norms of the new bond's columns and the new lambdas.
"""
import numpy as np
import pytest

from tests.conftest import gpu_available

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not gpu_available(), reason="needs a ROCm GPU")]

import torch  # noqa: E402

from dequantizer_amd import _lib, local  # noqa: E402

DEV = "cuda"


def _cg(rng, shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


class _SuCase:
    """One simple update on a local handle: node 0 (controlling, bond slot i1)
    and node 1 (slot i2), one degree-1 stub per outer bond."""

    def __init__(self, d1, d2, i1, i2, chi, rank, dims1, dims2, seed, p=2):
        self.h = local._local_handle(d1, d2, i1, i2, chi, p)
        self.chi, self.p, self.r = chi, p, rank
        self.d = (d1, d2)
        self.slots = (i1, i2)
        rng = np.random.default_rng(seed)
        h = self.h
        T = np.zeros(h.tensor_elems, dtype=np.complex128)
        self.active = np.zeros(h.tensor_elems, dtype=bool)
        for node, dims in ((0, dims1), (1, dims2)):
            off = int(h.offsets[node])
            view = T[off: off + h.node_size(node)].reshape(*((chi,) * len(dims)), p)
            act = self.active[off: off + h.node_size(node)].reshape(view.shape)
            sl = tuple(slice(0, x) for x in dims)
            view[sl] = _cg(rng, tuple(dims) + (p,))
            act[sl] = True
        self.T = T
        lam = np.zeros((h.topo.n_edges, chi))
        self.lam_active = np.zeros_like(lam, dtype=bool)
        for node, dims in ((0, dims1), (1, dims2)):
            for e, dm in zip(h.slot_edges[node], dims):
                if not self.lam_active[e].any():
                    lam[e, :dm] = np.sort(rng.uniform(0.2, 1.0, dm))[::-1]
                    self.lam_active[e, :dm] = True
        self.lam = lam
        halves = np.zeros((1, 2, p ** 4), dtype=np.complex128)
        halves[0, 0, : rank * p * p] = _cg(rng, rank * p * p) / 2
        halves[0, 1, : rank * p * p] = _cg(rng, rank * p * p) / 2
        self.halves = halves
        self.meta = np.array([[0, 0, rank, 0]], dtype=np.int32)
        self.tape_bytes = int(_lib.lib().dq_su_tape_bytes(h.ptr, rank))

    def forward(self, T, lam, halves, tape=None):
        L = _lib.lib()
        dT = torch.from_numpy(T.copy()).to(DEV)
        dl = torch.from_numpy(lam.reshape(-1).copy()).to(DEV)
        dm = torch.from_numpy(self.meta).to(DEV)
        dh = torch.from_numpy(halves.copy()).to(DEV)
        err = torch.zeros(1, dtype=torch.float64, device=DEV)
        ent = torch.zeros(1, dtype=torch.float64, device=DEV)
        st = _lib.stream_handle()
        if tape is None:
            tape = torch.zeros(self.tape_bytes, dtype=torch.uint8, device=DEV)
        _lib.check(L.dq_simple_update_tape(self.h.ptr, _lib.DQ_C128, _lib.ptr(dT), _lib.ptr(dl), _lib.ptr(dm), 1,
                                           _lib.ptr(dh), self.r, 1e-10, self.chi, _lib.ptr(err), _lib.ptr(ent),
                                           _lib.ptr(tape), st))
        return dT.cpu().numpy(), dl.cpu().numpy().reshape(lam.shape), tape

    def node(self, T, k):
        off = int(self.h.offsets[k])
        return T[off: off + self.h.node_size(k)].reshape(*((self.chi,) * self.d[k]), self.p)

    def loss_and_grads(self, Tout, lamout, W, wl, wc):
        """L = Re<W, Theta'> + sum wl lambda' + sum wc |column|^2 and its
        gradients with respect to the outputs."""
        A = np.moveaxis(self.node(Tout, 0), self.slots[0], -1)   # (others.., p, k)
        B = np.moveaxis(self.node(Tout, 1), self.slots[1], -1)
        lm = lamout[0]
        Af, Bf = A.reshape(-1, self.chi), B.reshape(-1, self.chi)
        theta = (Af * lm) @ Bf.T
        colsq = (np.abs(Af) ** 2).sum(axis=0)
        L = float(np.real(np.vdot(W, theta)) + wl @ lm + wc @ colsq)
        Abar = (W @ np.conj(Bf)) * lm + 2 * wc * Af
        Bbar = (W.T @ np.conj(Af)) * lm
        lbar = wl + np.real(np.einsum("xy,xj,yj->j", np.conj(W), Af, Bf))
        gT = np.zeros_like(Tout)
        for k, Xb in ((0, Abar), (1, Bbar)):
            g = np.moveaxis(Xb.reshape(np.moveaxis(self.node(Tout, k), self.slots[k], -1).shape), -1, self.slots[k])
            off = int(self.h.offsets[k])
            gT[off: off + self.h.node_size(k)] = g.reshape(-1)
        glam = np.zeros_like(lamout)
        glam[0] = lbar
        return L, gT, glam

    def vjp(self, T, lam, halves, tape, gT, glam):
        L = _lib.lib()
        dT = torch.from_numpy(T.copy()).to(DEV)
        dl = torch.from_numpy(lam.reshape(-1).copy()).to(DEV)
        dm = torch.from_numpy(self.meta).to(DEV)
        dh = torch.from_numpy(halves.copy()).to(DEV)
        gT_d = torch.from_numpy(gT.copy()).to(DEV)
        gl_d = torch.from_numpy(glam.reshape(-1).copy()).to(DEV)
        hb = torch.zeros_like(dh)
        _lib.check(L.dq_simple_update_vjp(self.h.ptr, _lib.DQ_C128, _lib.ptr(dT), _lib.ptr(dl), _lib.ptr(dm), 1,
                                          _lib.ptr(dh), self.r, 1e-10, 0.0, _lib.ptr(tape), _lib.ptr(gT_d), _lib.ptr(gl_d),
                                          _lib.ptr(hb), _lib.stream_handle()))
        return gT_d.cpu().numpy(), gl_d.cpu().numpy().reshape(lam.shape), hb.cpu().numpy()


CASES = [
    # d1, d2, i1, i2, chi, gate rank, dims node 0, dims node 1
    (3, 3, 1, 0, 4, 4, (4, 4, 4), (4, 4, 4)),
    (3, 2, 2, 1, 4, 2, (4, 3, 4), (2, 4)),
    (2, 3, 0, 2, 4, 2, (2, 3), (3, 3, 2)),
    (1, 3, 0, 1, 4, 2, (4,), (4, 4, 4)),
    (3, 3, 0, 2, 16, 2, (16, 16, 16), (16, 16, 16)),
]


@pytest.mark.parametrize("case", range(len(CASES)))
def test_simple_update_vjp_matches_central_differences(case):
    d1, d2, i1, i2, chi, rank, dims1, dims2 = CASES[case]
    c = _SuCase(d1, d2, i1, i2, chi, rank, dims1, dims2, seed=100 + case)
    rng = np.random.default_rng(7 + case)
    Tout, lamout, tape = c.forward(c.T, c.lam, c.halves)
    n_out = chi
    shapeA = np.moveaxis(c.node(Tout, 0), i1, -1).reshape(-1, chi).shape[0]
    shapeB = np.moveaxis(c.node(Tout, 1), i2, -1).reshape(-1, chi).shape[0]
    W = _cg(rng, (shapeA, shapeB)) / np.sqrt(shapeA * shapeB)
    wl = rng.standard_normal(n_out)
    wc = rng.standard_normal(n_out)
    L0, gT, glam = c.loss_and_grads(Tout, lamout, W, wl, wc)
    Tb, lb, hb = c.vjp(c.T, c.lam, c.halves, tape, gT, glam)

    def L_of(T, lam, halves):
        To, lo, _ = c.forward(T, lam, halves)
        return c.loss_and_grads(To, lo, W, wl, wc)[0]

    h = 1e-6
    for trial in range(3):
        dT = np.where(c.active, _cg(rng, c.T.shape), 0.0)
        dl = np.where(c.lam_active, rng.standard_normal(c.lam.shape), 0.0)
        dh = np.zeros_like(c.halves)
        dh[0, :, : rank * 4] = _cg(rng, (2, rank * 4))
        fd = (L_of(c.T + h * dT, c.lam + h * dl, c.halves + h * dh)
              - L_of(c.T - h * dT, c.lam - h * dl, c.halves - h * dh)) / (2 * h)
        an = (np.real(np.vdot(Tb[c.active], dT[c.active])) + float(np.sum(lb[c.lam_active] * dl[c.lam_active]))
              + np.real(np.vdot(hb, dh)))
        assert abs(fd - an) < 1e-5 * max(1.0, abs(fd)), (trial, fd, an)
        # each input separately
        for name, args, val in (("T", (c.T + h * dT, c.lam, c.halves, c.T - h * dT, c.lam, c.halves),
                                 np.real(np.vdot(Tb[c.active], dT[c.active]))),
                                ("lam", (c.T, c.lam + h * dl, c.halves, c.T, c.lam - h * dl, c.halves),
                                 float(np.sum(lb[c.lam_active] * dl[c.lam_active]))),
                                ("halves", (c.T, c.lam, c.halves + h * dh, c.T, c.lam, c.halves - h * dh),
                                 np.real(np.vdot(hb, dh)))):
            fd1 = (L_of(*args[:3]) - L_of(*args[3:])) / (2 * h)
            assert abs(fd1 - val) < 1e-5 * max(1.0, abs(fd1)), (name, trial, fd1, val)


# --------------------------------------------------------------------------
# whole-graph adjoints on a small random 3-regular graph


def _graph_handle(n, chi, seed):
    import networkx as nx

    import dequantizer_amd as dq
    from dequantizer_amd.engine import Topology
    g = dq.TensorGraph()
    for _ in range(n):
        g.add_node(2)
    for a, b in nx.random_regular_graph(3, n, seed=seed).edges:
        g.add_edge((a, b), chi)
    topo = Topology([list(g.get_traversal_iterator())])
    return topo, topo.handle(chi, 2, True)


def _psd_messages(rng, count, chi):
    a = _cg(rng, (count, chi, chi))
    m = a @ np.conj(np.transpose(a, (0, 2, 1)))
    return m / np.trace(m, axis1=1, axis2=2)[:, None, None]


def _d(x):
    return torch.from_numpy(np.ascontiguousarray(x)).to(DEV)


def _fd_check(L_of, grads_dot, dirs, h=1e-6, tol=1e-5):
    for d in dirs:
        fd = (L_of(*[x + h * dx for x, dx in d]) - L_of(*[x - h * dx for x, dx in d])) / (2 * h)
        an = grads_dot([dx for _, dx in d])
        assert abs(fd - an) < tol * max(1.0, abs(fd)), (fd, an)


@pytest.mark.parametrize("chi", [3, 16])
def test_bp_sweep_vjp_matches_central_differences(chi):
    """chi=16 on a 3-regular graph runs the matrix-core adjoint of the sweep."""
    topo, h = _graph_handle(8, chi, 1)
    E = topo.n_edges
    rng = np.random.default_rng(3)
    T0 = _cg(rng, h.tensor_elems)
    e2n0 = _psd_messages(rng, 2 * E, chi)
    W = _cg(rng, (2 * E, chi, chi))
    L = _lib.lib()
    st = _lib.stream_handle()

    def fwd(T, e2n):
        dT, de = _d(T), _d(e2n)
        n_out = torch.zeros_like(de)
        e_out = torch.zeros_like(de)
        _lib.check(L.dq_bp_sweep(h.ptr, _lib.DQ_C128, _lib.ptr(dT), _lib.ptr(de), _lib.ptr(de), _lib.ptr(n_out),
                                 _lib.ptr(e_out), None, st))
        return n_out.cpu().numpy()

    def L_of(T, e2n):
        return float(np.real(np.vdot(W, fwd(T, e2n))))

    dT, de = _d(T0), _d(e2n0)
    ebar = torch.zeros_like(de)
    Tbar = torch.zeros_like(dT)
    dW = _d(W)
    _lib.check(L.dq_bp_sweep_vjp(h.ptr, _lib.DQ_C128, _lib.ptr(dT), _lib.ptr(de), _lib.ptr(dW), _lib.ptr(ebar),
                                 _lib.ptr(Tbar), st))
    Tb, eb = Tbar.cpu().numpy(), ebar.cpu().numpy()
    dirs = [[(T0, _cg(rng, T0.shape)), (e2n0, _cg(rng, e2n0.shape))] for _ in range(3)]
    _fd_check(L_of, lambda ds: float(np.real(np.vdot(Tb, ds[0]) + np.vdot(eb, ds[1]))), dirs)


@pytest.mark.parametrize("chi", [3, 16])
def test_vidal_gauge_to_observables_vjp_matches_central_differences(chi):
    """messages -> Vidal gauge (edge SVD + absorb) -> symmetric gauge ->
    one- and two-site density matrices with diag(lambda) messages; the loss
    sees only gauge-invariant quantities.  chi=16 on a 3-regular graph runs
    the matrix-core absorb and its adjoint."""
    topo, h = _graph_handle(8, chi, 2)
    E, N = topo.n_edges, topo.n_nodes
    rng = np.random.default_rng(4)
    T0 = _cg(rng, h.tensor_elems)
    n2e0 = _psd_messages(rng, 2 * E, chi)
    W1 = _cg(rng, (N, 2, 2))
    W2 = _cg(rng, (E, 4, 4))
    wl = rng.standard_normal((E, chi))
    L = _lib.lib()
    st = _lib.stream_handle()
    nodes = torch.arange(N, dtype=torch.int32, device=DEV)
    edges = torch.arange(E, dtype=torch.int32, device=DEV)
    tape_bytes = int(L.dq_edge_svd_tape_bytes(h.ptr)) * E

    def fwd(T, n2e, keep=False):
        dT, dn = _d(T), _d(n2e)
        gam = torch.zeros_like(dT)
        lam = torch.zeros(E * chi, dtype=torch.float64, device=DEV)
        X = torch.zeros(2 * E * chi * chi, dtype=torch.complex128, device=DEV)
        tape = torch.zeros(tape_bytes, dtype=torch.uint8, device=DEV)
        _lib.check(L.dq_vidal_gauge_tape(h.ptr, _lib.DQ_C128, _lib.ptr(dT), _lib.ptr(dn), 1e-10, _lib.ptr(gam),
                                         _lib.ptr(lam), _lib.ptr(X), _lib.ptr(tape), st))
        ts = torch.zeros_like(gam)
        _lib.check(L.dq_symmetric_gauge(h.ptr, _lib.DQ_C128, _lib.ptr(gam), _lib.ptr(lam), 1e-10, _lib.ptr(ts), st))
        n_d = torch.zeros(2 * E * chi * chi, dtype=torch.complex128, device=DEV)
        e_d = torch.zeros_like(n_d)
        _lib.check(L.dq_lambdas_to_messages(h.ptr, _lib.DQ_C128, _lib.ptr(lam), _lib.ptr(n_d), _lib.ptr(e_d), st))
        r1 = torch.zeros((N, 2, 2), dtype=torch.complex128, device=DEV)
        r2 = torch.zeros((E, 4, 4), dtype=torch.complex128, device=DEV)
        _lib.check(L.dq_rdm1(h.ptr, _lib.DQ_C128, _lib.ptr(ts), _lib.ptr(e_d), _lib.ptr(nodes), N, _lib.ptr(r1), st))
        _lib.check(L.dq_rdm2(h.ptr, _lib.DQ_C128, _lib.ptr(ts), _lib.ptr(e_d), _lib.ptr(edges), E, _lib.ptr(r2), st))
        out = (r1.cpu().numpy(), r2.cpu().numpy(), lam.cpu().numpy().reshape(E, chi))
        return (out, (dT, gam, lam, X, tape, ts)) if keep else out

    def L_of(T, n2e):
        r1, r2, lm = fwd(T, n2e)
        return float(np.real(np.vdot(W1, r1) + np.vdot(W2, r2)) + np.sum(wl * lm))

    _, (dT, gam, lam, X, tape, ts) = fwd(T0, n2e0, keep=True)
    ts_bar = torch.zeros_like(ts)
    lam_bar = _d(wl.reshape(-1).copy())
    dW1, dW2 = _d(W1), _d(W2)
    _lib.check(L.dq_rdm1_vjp(h.ptr, _lib.DQ_C128, _lib.ptr(ts), _lib.ptr(lam), _lib.ptr(nodes), N, _lib.ptr(dW1),
                             _lib.ptr(ts_bar), _lib.ptr(lam_bar), st))
    _lib.check(L.dq_rdm2_vjp(h.ptr, _lib.DQ_C128, _lib.ptr(ts), _lib.ptr(lam), _lib.ptr(edges), E, _lib.ptr(dW2),
                             _lib.ptr(ts_bar), _lib.ptr(lam_bar), st))
    gam_bar = torch.zeros_like(gam)
    _lib.check(L.dq_symmetric_gauge_vjp(h.ptr, _lib.DQ_C128, _lib.ptr(gam), _lib.ptr(lam), 1e-10, _lib.ptr(ts_bar),
                                        _lib.ptr(gam_bar), _lib.ptr(lam_bar), st))
    Xbar = torch.zeros_like(X)
    Tbar = torch.zeros_like(dT)
    nbar = torch.zeros(2 * E * chi * chi, dtype=torch.complex128, device=DEV)
    _lib.check(L.dq_vidal_gauge_vjp(h.ptr, _lib.DQ_C128, _lib.ptr(dT), _lib.ptr(X), _lib.ptr(tape), 1e-10, 0.0,
                                    _lib.ptr(gam_bar), _lib.ptr(lam_bar), _lib.ptr(Xbar), _lib.ptr(Tbar),
                                    _lib.ptr(nbar), st))
    Tb, nb = Tbar.cpu().numpy(), nbar.cpu().numpy().reshape(n2e0.shape)

    def herm(a):
        return 0.5 * (a + np.conj(np.transpose(a, (0, 2, 1))))

    dirs = [[(T0, _cg(rng, T0.shape)), (n2e0, herm(_cg(rng, n2e0.shape)) * 0.1)] for _ in range(3)]
    _fd_check(L_of, lambda ds: float(np.real(np.vdot(Tb, ds[0]) + np.vdot(nb, ds[1]))), dirs)


def test_apply1_vjp_matches_central_differences():
    topo, h = _graph_handle(8, 3, 5)
    N = topo.n_nodes
    rng = np.random.default_rng(6)
    T0 = _cg(rng, h.tensor_elems)
    G0 = _cg(rng, (N, 2, 2))
    W = _cg(rng, T0.shape)
    L = _lib.lib()
    st = _lib.stream_handle()
    nodes = torch.arange(N, dtype=torch.int32, device=DEV)

    def L_of(T, G):
        dT, dG = _d(T), _d(G)  # keep every device buffer alive across the call
        _lib.check(L.dq_apply1(h.ptr, _lib.DQ_C128, _lib.ptr(dT), _lib.ptr(nodes), N, _lib.ptr(dG), st))
        return float(np.real(np.vdot(W, dT.cpu().numpy())))

    Tbar = _d(W.copy())
    gbar = torch.zeros((N, 2, 2), dtype=torch.complex128, device=DEV)
    dT0, dG0 = _d(T0), _d(G0)
    _lib.check(L.dq_apply1_vjp(h.ptr, _lib.DQ_C128, _lib.ptr(dT0), _lib.ptr(nodes), N, _lib.ptr(dG0),
                               _lib.ptr(Tbar), _lib.ptr(gbar), st))
    Tb, gb = Tbar.cpu().numpy(), gbar.cpu().numpy()
    dirs = [[(T0, _cg(rng, T0.shape)), (G0, _cg(rng, G0.shape))] for _ in range(3)]
    _fd_check(L_of, lambda ds: float(np.real(np.vdot(Tb, ds[0]) + np.vdot(gb, ds[1]))), dirs)


# --------------------------------------------------------------------------
# end to end: dE/d(knots, J) of the annealing run


def test_annealing_gradient_matches_central_differences():
    import dequantizer_amd as dq
    from dequantizer_amd.adjoint import AnnealingGradient
    from dequantizer_amd.annealing import Annealer, get_scheduler, knot_schedule, random_regular
    ef = random_regular(8, 3, seed=3)
    steps, chi, sweeps = 5, 4, 6
    runner = AnnealingGradient(ef, dq.PRNGKey(5), chi=chi, n_steps=steps, dt=0.3, n_knots=4, fixed_sweeps=sweeps)
    knots = np.array([0.1, 0.4, 0.7, 0.95])
    E, dk, dJ = runner.value_and_grad(knots)
    assert abs(E - runner.value(knots)) < 1e-10
    # the same run through the ordinary annealer (SVD gate split) gives the same energy
    ann = Annealer([ef], [dq.PRNGKey(5)], chi, 1, 1e-5, 100, True, "bfs", dtype=torch.complex128, chi=chi,
                   fixed_sweeps=sweeps)
    for mix, cpl in get_scheduler(0.3, knot_schedule(knots), steps):
        ann.step(mix, cpl)
    assert abs(ann.finish(two_site=True)[0].bp_energy - E) < 1e-9
    h = 1e-4
    for j in range(len(knots)):
        d = np.zeros_like(knots)
        d[j] = h
        fd = (runner.value(knots + d) - runner.value(knots - d)) / (2 * h)
        assert abs(fd - dk[j]) < 1e-5 * max(1.0, abs(fd)), ("knot", j, fd, dk[j])
    J0 = np.asarray(ef.coupling_amplitudes, dtype=np.float64)
    for e in (0, 3, 7, 11):
        d = np.zeros_like(J0)
        d[e] = h
        fd = (runner.value(knots, J0 + d) - runner.value(knots, J0 - d)) / (2 * h)
        assert abs(fd - dJ[e]) < 1e-5 * max(1.0, abs(fd)), ("J", e, fd, dJ[e])


def test_converged_bp_gradient_matches_central_differences():
    """Regauges iterated to the BP fixed point (message residual 1e-13) and
    differentiated implicitly (reverse iteration at the fixed point): the
    knot and coupling gradients match central differences of the forward to
    1e-5, and the energy equals the emulator's run under the reference's
    Vidal-distance stopping rule driven to 1e-12 (the same fixed points)."""
    import dequantizer_amd as dq
    from dequantizer_amd.adjoint import AnnealingGradient
    from dequantizer_amd.annealing import Annealer, get_scheduler, knot_schedule, random_regular
    ef = random_regular(8, 3, seed=3)
    steps, chi = 5, 4
    runner = AnnealingGradient(ef, dq.PRNGKey(5), chi=chi, n_steps=steps, dt=0.3, n_knots=4, fixed_sweeps=None,
                               bp_tol=1e-13)
    knots = np.array([0.1, 0.4, 0.7, 0.95])
    E, dk, dJ = runner.value_and_grad(knots)
    # every regauge converged, and every adjoint iteration ran at least once
    assert len(runner.adjoint_iterations) == steps and min(runner.adjoint_iterations) >= 1
    assert max(runner.forward_sweeps) < runner.max_sweeps
    ann = Annealer([ef], [dq.PRNGKey(5)], chi, 1, 1e-12, 1000, True, "bfs", dtype=torch.complex128, chi=chi)
    for mix, cpl in get_scheduler(0.3, knot_schedule(knots), steps):
        ann.step(mix, cpl)
    assert abs(ann.finish(two_site=True)[0].bp_energy - E) < 1e-9
    h = 1e-4
    for j in range(len(knots)):
        d = np.zeros_like(knots)
        d[j] = h
        fd = (runner.value(knots + d) - runner.value(knots - d)) / (2 * h)
        assert abs(fd - dk[j]) < 1e-5 * max(1.0, abs(fd)), ("knot", j, fd, dk[j])
    J0 = np.asarray(ef.coupling_amplitudes, dtype=np.float64)
    for e in (0, 5, 11):
        d = np.zeros_like(J0)
        d[e] = h
        fd = (runner.value(knots, J0 + d) - runner.value(knots, J0 - d)) / (2 * h)
        assert abs(fd - dJ[e]) < 1e-5 * max(1.0, abs(fd)), ("J", e, fd, dJ[e])
    # the fixed-sweep gradient differs: it is the gradient of another function
    fixed = AnnealingGradient(ef, dq.PRNGKey(5), chi=chi, n_steps=steps, dt=0.3, n_knots=4, fixed_sweeps=1)
    assert np.abs(fixed.value_and_grad(knots)[1] - dk).max() > 1e-6


@pytest.mark.parametrize("fixed_sweeps", [6, None])
def test_batched_gradient_matches_single_instance_runs(fixed_sweeps):
    """Three instances of different sizes in one engine under one schedule:
    each instance's energy, knot gradient and coupling gradient equal those
    of the instance run alone (one backward serves all three), with fixed
    sweeps and with converged BP."""
    import dequantizer_amd as dq
    from dequantizer_amd.adjoint import AnnealingGradient
    from dequantizer_amd.annealing import random_regular
    efs = [random_regular(n, 3, seed=20 + n) for n in (6, 10, 8)]
    keys = [dq.PRNGKey(40 + i) for i in range(3)]
    kw = dict(chi=4, n_steps=4, dt=0.3, n_knots=3, fixed_sweeps=fixed_sweeps, bp_tol=1e-13)
    knots = np.array([0.15, 0.55, 0.9])
    batched = AnnealingGradient(efs, keys, **kw)
    Jb = [np.asarray(ef.coupling_amplitudes, dtype=np.float64) * (1.0 + 0.1 * i) for i, ef in enumerate(efs)]
    E, dk, dJ = batched.value_and_grad(knots, Jb)
    assert E.shape == (3,) and dk.shape == (3, 3) and [len(x) for x in dJ] == [ef.edges_number for ef in efs]
    assert np.allclose(batched.value(knots, Jb), E, rtol=0, atol=1e-12)
    for i, ef in enumerate(efs):
        single = AnnealingGradient(ef, keys[i], **kw)
        e1, dk1, dJ1 = single.value_and_grad(knots, Jb[i])
        assert abs(E[i] - e1) < 1e-11 * max(1.0, abs(e1)), (i, E[i], e1)
        assert np.allclose(dk[i], dk1, rtol=1e-9, atol=1e-11), (i, dk[i], dk1)
        assert np.allclose(dJ[i], dJ1, rtol=1e-9, atol=1e-11), (i, np.abs(dJ[i] - dJ1).max())


@pytest.mark.parametrize("fixed_sweeps", [6, None])
def test_complex64_forward_gradient_matches_complex128(fixed_sweeps):
    """The forward in complex64 (the benchmark engine's precision) with the
    complex128 adjoint reading its widened records: energy within the
    complex64 rounding of the complex128 run and knot and coupling gradients
    within 1e-4 relative of it, with fixed sweeps and with converged BP (the
    complex64 fixed point at residual 1e-7)."""
    import dequantizer_amd as dq
    from dequantizer_amd.adjoint import AnnealingGradient
    from dequantizer_amd.annealing import random_regular
    ef = random_regular(12, 3, seed=9)
    kw = dict(chi=8, n_steps=6, dt=0.3, n_knots=4, fixed_sweeps=fixed_sweeps)
    knots = np.array([0.1, 0.35, 0.7, 0.95])
    r128 = AnnealingGradient(ef, dq.PRNGKey(8), bp_tol=1e-13, **kw)
    r64 = AnnealingGradient(ef, dq.PRNGKey(8), bp_tol=1e-7, forward_dtype=torch.complex64, **kw)
    assert r64.gam0.dtype == torch.complex64
    e1, dk1, dj1 = r128.value_and_grad(knots)
    e2, dk2, dj2 = r64.value_and_grad(knots)
    print("c64 forward: |dE|", abs(e2 - e1), "knots", np.abs(dk2 - dk1).max() / np.abs(dk1).max(),
          "couplings", np.abs(dj2 - dj1).max() / np.abs(dj1).max(), flush=True)
    assert abs(e2 - e1) < 1e-5 * ef.nodes_number, (e1, e2)
    assert abs(e2 - r64.value(knots)) < 1e-10 * max(1.0, abs(e2))
    assert np.abs(dk2 - dk1).max() < 1e-4 * np.abs(dk1).max(), (dk1, dk2)
    assert np.abs(dj2 - dj1).max() < 1e-4 * np.abs(dj1).max(), np.abs(dj2 - dj1).max()


def test_segmented_recomputation_matches_one_segment():
    """The backward recomputing segments of 1, 2 and 3 steps from their
    checkpoints returns the same energy, bit for bit, and the same gradients
    to rounding (the adjoint kernels accumulate with atomics) as the one that
    keeps every step's records; a memory budget below one step's records is
    refused."""
    import dequantizer_amd as dq
    from dequantizer_amd.adjoint import AnnealingGradient
    from dequantizer_amd.annealing import random_regular
    ef = random_regular(10, 3, seed=4)
    runner = AnnealingGradient(ef, dq.PRNGKey(3), chi=4, n_steps=7, dt=0.3, n_knots=4, fixed_sweeps=4)
    knots = np.array([0.05, 0.3, 0.6, 0.9])
    ref = runner.value_and_grad(knots, segment_steps=7)
    assert runner.segment_length(1e6) == 7
    for seg in (1, 2, 3):
        E, dk, dJ = runner.value_and_grad(knots, segment_steps=seg)
        assert runner.segment_steps == seg
        assert E == ref[0], seg
        assert np.allclose(dk, ref[1], rtol=1e-11, atol=1e-13), (seg, np.abs(dk - ref[1]).max())
        assert np.allclose(dJ, ref[2], rtol=1e-11, atol=1e-13), (seg, np.abs(dJ - ref[2]).max())
    # the budget holding 3 checkpoints and 3 steps' records gives segments of 3
    state = runner.gam0.numel() * 16 + runner.lam0.numel() * 8
    gb = (3 * state + 3 * runner.record_bytes()) / 2 ** 30
    assert runner.segment_length(gb * 1.0001) == 3
    with pytest.raises(ValueError):
        runner.segment_length(runner.record_bytes() / 2 ** 31)


def test_annealing_gradient_taylor_remainder():
    """E(x + h v) - E(x) - h <grad, v> = O(h^2) along a random direction v of
    (knots, J): halving h divides the remainder by about 4.  Three steps at
    chi=8 keep every bond below the cap (a gate at most doubles its bond), so
    no truncation kink lies on the path."""
    import dequantizer_amd as dq
    from dequantizer_amd.adjoint import AnnealingGradient
    from dequantizer_amd.annealing import random_regular
    ef = random_regular(10, 3, seed=11)
    runner = AnnealingGradient(ef, dq.PRNGKey(2), chi=8, n_steps=3, dt=0.3, n_knots=3, fixed_sweeps=6)
    knots = np.array([0.2, 0.5, 0.9])
    J0 = np.asarray(ef.coupling_amplitudes, dtype=np.float64)
    E, dk, dJ = runner.value_and_grad(knots)
    rng = np.random.default_rng(0)
    vk, vJ = rng.standard_normal(knots.shape), rng.standard_normal(J0.shape)
    slope = float(dk @ vk + dJ @ vJ)
    rem = [abs(runner.value(knots + h * vk, J0 + h * vJ) - E - h * slope) for h in (2e-2, 1e-2, 5e-3)]
    assert rem[0] > 1e-9, rem  # the second-order term is above rounding
    for a, b in zip(rem, rem[1:]):
        assert 3.0 < a / b < 5.0, rem
    # a gradient off by 10% leaves an O(h) remainder instead
    bad = [abs(runner.value(knots + h * vk, J0 + h * vJ) - E - h * 1.1 * slope) for h in (1e-3, 5e-4)]
    assert bad[0] / bad[1] < 3.0, bad


def test_annealing_gradient_full_size():
    """1000-spin random 3-regular instance (seed 7), chi=16, 200 steps of
    dt=0.1, 32 knots at s_k = t_k, records held to the default 24 GiB (the
    backward recomputes segments from checkpoints): gradients against
    central differences of the forward at h = 3e-6, where they have
    converged in h (E has large third derivatives along the early knots:
    at h = 1e-4 the difference quotient of knot 8 is off by 2e-3).

    Eight knots from 10 on and sixteen couplings hold the 1e-4 target
    (measured on MI355X: knots <= 5e-7 apart from FD noise, couplings
    <= 4.5e-5).  The knots of the first ~60 steps, where many Schmidt values
    grow from zero and the backward freezes those below cut = 1e-6 of the
    largest on their bond, are held to what that cutoff gives: knots 1, 3,
    5, 8 measured at 1.4e-3, 1.1e-4, 3.2e-3, 7.0e-4."""
    import dequantizer_amd as dq
    from dequantizer_amd.adjoint import AnnealingGradient
    from dequantizer_amd.annealing import edge_colored, random_regular
    ef = edge_colored(random_regular(1000, 3, seed=7, couplings="pm1", field_std=0.1))
    runner = AnnealingGradient(ef, dq.PRNGKey(7), chi=16, n_steps=200, dt=0.1, n_knots=32, fixed_sweeps=10)
    knots = np.linspace(0.0, 1.0, 32)
    J0 = np.asarray(ef.coupling_amplitudes, dtype=np.float64)
    E, dk, dJ = runner.value_and_grad(knots)
    assert runner.segment_steps < 200  # recomputation is on
    assert np.all(np.isfinite(dk)) and np.all(np.isfinite(dJ))
    h = 3e-6

    def fd_knot(j):
        d = np.zeros(32)
        d[j] = h
        return (runner.value(knots + d) - runner.value(knots - d)) / (2 * h)

    for j in (10, 12, 14, 16, 20, 24, 28, 30):
        fd = fd_knot(j)
        print(f"knot {j}: adjoint {dk[j]:.7f} fd {fd:.7f} rel {abs(fd - dk[j]) / abs(fd):.1e}", flush=True)
        assert abs(fd - dk[j]) < 1e-4 * max(1.0, abs(fd)), ("knot", j, fd, dk[j])
    for j, tol in ((1, 5e-3), (3, 5e-3), (5, 1e-2), (8, 5e-3)):
        fd = fd_knot(j)
        print(f"knot {j}: adjoint {dk[j]:.7f} fd {fd:.7f} rel {abs(fd - dk[j]) / abs(fd):.1e}", flush=True)
        assert abs(fd - dk[j]) < tol * abs(fd), ("knot", j, fd, dk[j])
    rng = np.random.default_rng(0)
    worst = 0.0
    for c in sorted(rng.choice(len(J0), 16, replace=False).tolist()):
        d = np.zeros_like(J0)
        d[c] = h
        fd = (runner.value(knots, J0 + d) - runner.value(knots, J0 - d)) / (2 * h)
        worst = max(worst, abs(fd - dJ[c]) / max(1.0, abs(fd)))
        assert abs(fd - dJ[c]) < 1e-4 * max(1.0, abs(fd)), ("J", c, fd, dJ[c])
    print(f"16 couplings: worst rel {worst:.1e}", flush=True)
