"""GPU parity of the exact state-vector kernels (dq_sv_*, through the C ABI).

Against the oracle's state-vector emulator (itself pinned against explicit
bit-by-bit operators in test_oracle.py): single one- and two-qubit gates,
layers of single-qubit gates and fused diagonal layers, one-qubit density
matrices and measurements, at sizes below, at and above the 12-bit LDS tile
(1, 2, 5, 12, 13, 17 qubits).  The exact annealer is checked against the same
sequence of gates on the oracle and, on a tree, against the BP annealer
(exact on trees).  At 28 qubits (4 GiB): a layer followed by its inverse
restores the state, and every density matrix has unit trace.
Floating point (complex128): tolerance 1e-12 absolute on unit-norm states.
"""
import numpy as np
import pytest

from tests.conftest import gpu_available

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not gpu_available(), reason="needs a ROCm GPU")]

import torch  # noqa: E402

import dequantizer_amd as dq  # noqa: E402
from dequantizer_amd import annealing as da  # noqa: E402
from dequantizer_amd import exact as dx  # noqa: E402
from oracle import emulator as oe  # noqa: E402
from oracle import graph as og  # noqa: E402

TOL = 1e-12


def _oracle(n, psi):
    ex = oe.ExactQuantumEmulator.__new__(oe.ExactQuantumEmulator)
    ex.n, ex.node2pos, ex.state = n, {i: i for i in range(n)}, psi.copy()
    return ex


def _state(rng, n):
    psi = rng.standard_normal(2**n) + 1j * rng.standard_normal(2**n)
    return psi / np.linalg.norm(psi)


def _unitary(rng, k):
    q, _ = np.linalg.qr(rng.standard_normal((k, k)) + 1j * rng.standard_normal((k, k)))
    return q


def _load(sv, psi):
    sv.psi.copy_(torch.from_numpy(psi))


def _diff(sv, ex):
    return float(np.abs(sv.psi.cpu().numpy() - ex.state).max())


@pytest.mark.parametrize("n", [1, 2, 5, 12, 13, 17])
def test_kernels_vs_oracle(n):
    print(f"exact kernels n={n}", flush=True)
    rng = np.random.default_rng(100 + n)
    psi = _state(rng, n)
    sv = dx.StateVector(n)
    _load(sv, psi)
    ex = _oracle(n, psi)
    # single-qubit gates (generic, not unitary) on the first, last and a middle qubit
    for q in sorted({0, n - 1, n // 2}):
        g = rng.standard_normal((2, 2)) + 1j * rng.standard_normal((2, 2))
        g /= np.linalg.norm(g, 2)
        sv.apply1(q, g)
        ex.apply_q1(g, q)
    assert _diff(sv, ex) < TOL
    if n >= 2:
        pairs = [(0, n - 1), (n - 1, 0), (n // 2, 0) if n // 2 != 0 else (1, 0)]
        for a, b in pairs:
            g = _unitary(rng, 4).reshape(2, 2, 2, 2)
            sv.apply2(a, b, g)
            ex.apply_q2(g, a, b)
        assert _diff(sv, ex) < TOL
    # single-qubit layer on a subset and on all qubits
    gates = np.array([_unitary(rng, 2) for _ in range(n)])
    subset = sorted(rng.choice(n, size=max(1, n // 2), replace=False).tolist())
    sv.apply1_layer(gates, subset)
    for q in subset:
        ex.apply_q1(gates[q], q)
    assert _diff(sv, ex) < TOL
    sv.apply1_layer(gates)
    for q in range(n):
        ex.apply_q1(gates[q], q)
    assert _diff(sv, ex) < TOL
    # fused diagonal layer: terms on low-low, low-high and high-high bits and
    # random pairs, unit-modulus entries plus one generic term
    if n >= 2:
        terms = [(n - 1, n - 2), (0, 1), (0, n - 1), (n - 1, 0)]
        for _ in range(3 * n):
            a, b = rng.choice(n, size=2, replace=False)
            terms.append((int(a), int(b)))
        diags = np.exp(1j * rng.uniform(-np.pi, np.pi, (len(terms), 4)))
        diags[1] *= np.array([1.0, 0.5, 0.8, 1.1])
        sv.apply_diagonal(terms, diags)
        for (a, b), d in zip(terms, diags):
            ex.apply_q2(np.diag(d).reshape(2, 2, 2, 2), a, b)
        assert _diff(sv, ex) < TOL
    # density matrices of all qubits and of a subset
    norm2 = float(np.vdot(ex.state, ex.state).real)
    rho = sv.rdm1().cpu().numpy()
    for q in range(n):
        assert np.abs(rho[q] - ex.dens_q1(q)).max() < TOL * max(1.0, norm2)
    sub = [n - 1, 0]
    rs = sv.rdm1(sub).cpu().numpy()
    for k, q in enumerate(sub):
        assert np.abs(rs[k] - ex.dens_q1(q)).max() < TOL * max(1.0, norm2)


def test_diagonal_layer_longer_than_one_launch():
    """300 diagonal terms: the library splits them over two launches of at
    most 256 terms."""
    n = 14
    rng = np.random.default_rng(14)
    psi = _state(rng, n)
    sv = dx.StateVector(n)
    _load(sv, psi)
    ex = _oracle(n, psi)
    terms = [(int(a), int(b)) for a, b in (rng.choice(n, size=2, replace=False) for _ in range(300))]
    diags = np.exp(1j * rng.uniform(-np.pi, np.pi, (len(terms), 4)))
    sv.apply_diagonal(terms, diags)
    for (a, b), d in zip(terms, diags):
        ex.apply_q2(np.diag(d).reshape(2, 2, 2, 2), a, b)
    assert _diff(sv, ex) < TOL


@pytest.mark.parametrize("n,k", [(5, 10), (12, 40), (13, 64), (13, 0), (17, 45), (17, 65)])
def test_diagonal_then_layer_vs_oracle(n, k):
    """dq_sv_apply_diagonal_layer1: the diagonal terms ride on the low tile
    pass (12 or more qubits, at most 64 terms) or take their own pass (fewer
    qubits, more terms); subset and full layers."""
    print(f"diagonal + layer n={n} terms={k}", flush=True)
    rng = np.random.default_rng(1000 + 10 * n + k)
    psi = _state(rng, n)
    sv = dx.StateVector(n)
    _load(sv, psi)
    ex = _oracle(n, psi)
    for subset in (sorted(rng.choice(n, size=max(1, n // 3), replace=False).tolist()), None):
        terms = [(int(a), int(b)) for a, b in (rng.choice(n, size=2, replace=False) for _ in range(k))]
        diags = np.exp(1j * rng.uniform(-np.pi, np.pi, (k, 4)))
        gates = np.array([_unitary(rng, 2) for _ in range(n)])
        sv.apply_diagonal_layer1(terms, diags, gates, subset)
        for (a, b), d in zip(terms, diags):
            ex.apply_q2(np.diag(d).reshape(2, 2, 2, 2), a, b)
        for q in (range(n) if subset is None else subset):
            ex.apply_q1(gates[q], q)
        assert _diff(sv, ex) < TOL, subset


def test_set_basis_and_measure_vs_oracle():
    n = 13
    rng = np.random.default_rng(7)
    sv = dx.StateVector(n)
    idx = int(rng.integers(0, 2**n))
    sv.set_basis(idx)
    host = sv.psi.cpu().numpy()
    assert host[idx] == 1.0 and np.count_nonzero(host) == 1
    psi = _state(rng, n)
    _load(sv, psi)
    ex = _oracle(n, psi)
    for q in [3, 0, 12, 7]:
        u = float(rng.uniform())
        r = sv.measure(q, u)
        # oracle measurement with the same uniform sample
        dens = ex.dens_q1(q)
        p = dens[0, 0].real
        ro, norm = (0, p) if p > u else (1, 1 - p)
        proj = np.zeros((2, 2))
        proj[ro, ro] = 1.0
        ex.apply_q1(proj, q)
        ex.state /= np.sqrt(norm)
        assert r == ro
        assert _diff(sv, ex) < TOL


def _same_graph():
    src = og.heavy_cube()
    gp = dq.TensorGraph()
    for nid in src.nodes:
        gp.add_node(2, nid)
    for eid, e in src.edges.items():
        gp.add_edge(eid, e.dimension)
    return gp, src


def test_emulator_api_vs_oracle():
    """ExactQuantumEmulator (graph API, traversal qubit order, keyed
    measurements) against the oracle's emulator on the same graph."""
    gp, go = _same_graph()
    key = dq.PRNGKey(5)
    key, sub = dq.split(key)
    gx = dx.ExactQuantumEmulator(gp, sub)
    ox = oe.ExactQuantumEmulator(go, sub)
    print(f"emulator API: {gp.nodes_number} qubits", flush=True)
    rng = dq.random.generator(key)
    state = {nid: int(rng.integers(0, 2)) for nid in go.nodes}
    gx.set_to_product_state(state)
    ox.set_to_product_state(state)
    assert np.abs(gx.state.cpu().numpy() - ox.state).max() == 0.0
    grng = np.random.default_rng(3)
    trav = list(gp.get_traversal_iterator())
    for _ in range(2):
        for el in trav:
            if isinstance(el, dq.Node):
                g = _unitary(grng, 2)
                gx.apply_q1(g, el.id)
                ox.apply_q1(g, el.id)
            else:
                g = _unitary(grng, 4).reshape(2, 2, 2, 2)
                gx.apply_q2(g, *el.id)
                ox.apply_q2(g, *el.id)
    assert np.abs(gx.state.cpu().numpy() - ox.state).max() < TOL
    for nid in go.nodes:
        assert np.abs(gx.dens_q1(nid) - ox.dens_q1(nid)).max() < TOL
    for nid in list(go.nodes)[:6]:
        assert gx.measure(nid) == ox.measure(nid)
    assert gx.prng_key == ox.prng_key
    assert np.abs(gx.state.cpu().numpy() - ox.state).max() < TOL


def test_exact_annealer_vs_oracle_gates():
    """run_exact_quantum_annealer (fused layers) against the gate-by-gate
    sequence of exact_quantum_annealer.py:92-122 on the oracle emulator."""
    ef = da.random_regular(12, 3, seed=4, couplings="normal", field_std=0.3)
    sched = da.get_scheduler(0.2, da.linear_schedule, 8)
    key = dq.PRNGKey(9)
    res = dx.run_exact_quantum_annealer(ef, sched, key, True, True)
    n = ef.nodes_number
    psi = np.zeros(2**n, dtype=complex)
    psi[0] = 1
    ex = _oracle(n, psi)
    for q in range(n):
        ex.apply_q1(da.HADAMARD, q)
    deg = ef.node_degrees
    hist = []
    for mix, cpl in sched:
        for k, (a, b) in enumerate(ef.coupled_spin_pairs):
            ham = (ef.coupling_amplitudes[k] * np.array([1, -1, -1, 1.0])
                   + ef.fields[a] / deg[a] * np.array([1, 1, -1, -1.0])
                   + ef.fields[b] / deg[b] * np.array([1, -1, 1, -1.0]))
            ex.apply_q2(np.diag(np.exp(-1j * cpl * ham)).reshape(2, 2, 2, 2), int(a), int(b))
        c, s = np.cos(mix), np.sin(mix)
        for q in range(n):
            ex.apply_q1(np.array([[c, -1j * s], [-1j * s, c]]), q)
        hist.append(np.array([ex.dens_q1(q) for q in range(n)]))
    assert len(res.density_matrices_history) == len(hist)
    for h_gpu, h_ref in zip(res.density_matrices_history, hist):
        assert np.abs(h_gpu - h_ref).max() < 1e-11
    assert np.abs(res.density_matrices - hist[-1]).max() < 1e-11
    # sampling: same keys, same rule
    k = key
    out = []
    for q in range(n):
        k, sub = dq.split(k)
        u = dq.random.generator(sub).uniform()
        dens = ex.dens_q1(q)
        p = dens[0, 0].real
        r, norm = (0, p) if p > u else (1, 1 - p)
        proj = np.zeros((2, 2))
        proj[r, r] = 1.0
        ex.apply_q1(proj, q)
        ex.state /= np.sqrt(norm)
        out.append(r)
    assert np.array_equal(res.configuration, 1 - 2 * np.array(out))


def test_exact_annealer_vs_bp_on_a_tree():
    """On a tree BP is exact: the exact and the BP annealers give the same
    one-site density matrices (two independent GPU paths)."""
    rng = np.random.default_rng(2)
    n = 10  # a chain: Schmidt ranks <= 2^5 = max_chi, so no truncation
    pairs = np.array([[i, i + 1] for i in range(n - 1)])
    ef = da.EnergyFunction(rng.standard_normal(n - 1), pairs, 0.3 * rng.standard_normal(n))
    sched = da.get_scheduler(0.3, da.linear_schedule, 10)
    ex = dx.run_exact_quantum_annealer(ef, sched, dq.PRNGKey(1), False, False)
    bp = da.run_quantum_annealer(ef, sched, dq.PRNGKey(1), 32, ef.edges_number, 1e-10, 100, False, "dfs",
                                 dtype=torch.complex128)
    assert np.abs(ex.density_matrices - bp.density_matrices).max() < 1e-6


def test_large_state_inverse_layers():
    """28 qubits (4 GiB): random single-qubit layer and diagonal layer, then
    their inverses in reverse order, restore the initial state; all density
    matrices have unit trace."""
    n = 28
    print(f"exact kernels n={n}", flush=True)
    sv = dx.StateVector(n)
    gen = torch.Generator(device="cuda").manual_seed(0)
    re = torch.randn(2**n, dtype=torch.float64, device="cuda", generator=gen)
    im = torch.randn(2**n, dtype=torch.float64, device="cuda", generator=gen)
    psi0 = torch.complex(re, im)
    del re, im
    psi0 /= torch.linalg.vector_norm(psi0)
    sv.psi.copy_(psi0)
    rng = np.random.default_rng(28)
    gates = np.array([_unitary(rng, 2) for _ in range(n)])
    terms = [(int(a), int(b)) for a, b in (rng.choice(n, size=2, replace=False) for _ in range(60))]
    diags = np.exp(1j * rng.uniform(-np.pi, np.pi, (len(terms), 4)))
    sv.apply1_layer(gates)
    sv.apply_diagonal(terms, diags)
    rho = sv.rdm1().cpu().numpy()
    assert np.abs(np.trace(rho, axis1=1, axis2=2) - 1.0).max() < 1e-10
    sv.apply_diagonal(terms, diags.conj())
    sv.apply1_layer(np.conj(np.transpose(gates, (0, 2, 1))))
    err = float((sv.psi - psi0).abs().max())
    assert err < 1e-12


def test_argument_errors():
    sv = dx.StateVector(4)
    with pytest.raises(ValueError):
        sv.apply1(4, np.eye(2))
    with pytest.raises(ValueError):
        sv.apply_diagonal([(1, 1)], [[1, 1, 1, 1]])
    with pytest.raises(dq._lib.DQError):
        dq._lib.check(dq._lib.lib().dq_sv_apply2(dq._lib.ptr(sv.psi), 4, 2, 2, dq._lib.ptr(sv.psi), None))
