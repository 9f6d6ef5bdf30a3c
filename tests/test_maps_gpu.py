"""GPU parity of the whole-graph maps against the complex128 oracle.

Every map of dequantizer_amd.mappings runs through libdq (C ABI) and is
compared with oracle.mappings on identical inputs:
  * BP sweeps (synchronous, asynchronous in bfs and dfs order): messages
    elementwise;
  * Vidal gauge: lambdas elementwise (they are unique), Gammas through
    gauge-invariant quantities (Vidal distance, one-site density matrices);
  * Vidal distance, symmetric gauge, density matrices, residual: elementwise on
    identical inputs.
"""
import numpy as np
import pytest

from tests.conftest import gpu_available
from tests.helpers import host, random_messages, random_tensors, to_oracle_messages

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not gpu_available(), reason="needs a ROCm GPU")]

import dequantizer_amd as dq  # noqa: E402
from oracle import graph as og  # noqa: E402
from oracle import mappings as om  # noqa: E402


def lattice_pair(sizes, phys, bond, open_boundary=True):
    return dq.get_nd_lattice(sizes, phys, bond, open_boundary), og.nd_lattice(sizes, phys, bond, open_boundary)


def regular_pair(n, bond, seed):
    import networkx as nx
    pairs = list(nx.random_regular_graph(3, n, seed=seed).edges)
    g1, g2 = dq.TensorGraph(), og.TensorGraph()
    for g in (g1, g2):
        for _ in range(n):
            g.add_node(2)
        for a, b in pairs:
            g.add_edge((a, b), bond)
    return g1, g2


CASES = {
    "lattice3x3_chi3": lambda: lattice_pair([3, 3], 2, 3),
    "cube2_chi2": lambda: lattice_pair([2, 2, 2], 2, 2, open_boundary=False),
    "rr3_12_chi4": lambda: regular_pair(12, 4, 3),
    "rr3_10_chi16": lambda: regular_pair(10, 16, 3),
    "rr3_8_chi32": lambda: regular_pair(8, 32, 5),
}


def _rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


@pytest.mark.parametrize("case", list(CASES))
@pytest.mark.parametrize("sync", [True, False])
@pytest.mark.parametrize("order", ["bfs", "dfs"])
@pytest.mark.parametrize("dtype,tol", [(np.complex128, 1e-10), (np.complex64, 2e-4)])
def test_bp_sweep_matches_oracle(case, sync, order, dtype, tol):
    gp, go = CASES[case]()
    rng = np.random.default_rng(7)
    tensors = random_tensors(gp, rng)
    messages = random_messages(gp, rng)
    trav_p = list(gp.get_traversal_iterator(ordering=order))
    trav_o = go.traversal(ordering=order)
    bp_p = dq.get_belief_propagation_map(trav_p, sync)
    bp_o = om.belief_propagation_map(trav_o, sync)
    tp = {k: v.astype(dtype) for k, v in tensors.items()}
    mp_ = {k: v.astype(dtype) for k, v in messages.items()}
    out_o = bp_o(tensors, to_oracle_messages(messages))
    out_p = bp_p(tp, mp_)
    # two sweeps, the second from the packed device output
    out_o2 = bp_o(tensors, out_o)
    out_p2 = bp_p(tp, out_p)
    assert len(out_p) == len(out_o)
    for k, v in out_o2.items():
        got = host(out_p2[dq.MessageID(k.src, k.dst)])
        assert _rel(got, v) < tol, (k, _rel(got, v))


@pytest.mark.parametrize("case", list(CASES))
@pytest.mark.parametrize("dtype,tol", [(np.complex128, 1e-9), (np.complex64, 5e-4)])
def test_vidal_gauge_and_distance(case, dtype, tol):
    gp, go = CASES[case]()
    rng = np.random.default_rng(11)
    tensors = random_tensors(gp, rng)
    messages = random_messages(gp, rng)
    trav_p = list(gp.get_traversal_iterator())
    trav_o = go.traversal()
    eps = 1e-10
    # converge BP on the oracle first so the gauge is well defined
    bp_o = om.belief_propagation_map(trav_o, False)
    mo = to_oracle_messages(messages)
    for _ in range(60):
        mo = bp_o(tensors, mo)
    vg_o = om.vidal_gauge_fixing_map(trav_o, eps)
    gam_o, lam_o = vg_o(tensors, mo)
    vg_p = dq.get_vidal_gauge_fixing_map(trav_p, eps)
    mp_ = {dq.MessageID(k.src, k.dst): v.astype(dtype) for k, v in mo.items()}
    gam_p, lam_p = vg_p({k: v.astype(dtype) for k, v in tensors.items()}, mp_)
    for eid, lm in lam_o.items():
        assert np.abs(host(lam_p[eid]) - lm).max() < max(tol, 1e-12) * 10, eid
    # gauge-invariant checks of Gamma: oracle Vidal distance and symmetric-gauge rho
    gam_h = {nid: host(gam_p[nid]).astype(np.complex128) for nid in gam_o}
    lam_h = {eid: host(lam_p[eid]) for eid in lam_o}
    vd_o = om.vidal_gauge_distance_map(trav_o, eps)
    assert abs(vd_o(gam_h, lam_h) - vd_o(gam_o, lam_o)) < tol * 10
    sg_o = om.symmetric_gauge_fixing_map(trav_o, eps)
    t_o, t_p = sg_o(gam_o, lam_o), sg_o(gam_h, lam_h)
    msg_o, msg_p = om.lambdas2messages(lam_o), om.lambdas2messages(lam_h)
    for node in go.nodes.values():
        r_o = om.one_side_density_matrix(node, t_o, msg_o)
        r_p = om.one_side_density_matrix(node, t_p, msg_p)
        assert np.abs(r_o - r_p).max() < tol * 10, node.id
    # Vidal distance map on identical inputs
    vd_p = dq.get_vidal_gauge_distance_map(trav_p, eps)
    assert abs(float(vd_p({k: v.astype(dtype) for k, v in gam_o.items()}, lam_o)) - vd_o(gam_o, lam_o)) < tol * 10


@pytest.mark.parametrize("case", list(CASES))
def test_symmetric_gauge_rdm_and_residual(case):
    gp, go = CASES[case]()
    rng = np.random.default_rng(5)
    tensors = random_tensors(gp, rng)
    messages = random_messages(gp, rng)
    messages2 = random_messages(gp, rng)
    trav_p = list(gp.get_traversal_iterator())
    trav_o = go.traversal()
    lambdas = {e.id: np.sort(rng.uniform(0, 1, e.dimension))[::-1] for e in go.edges.values()}
    lambdas[next(iter(lambdas))][-1] = 0.0  # a zero lambda
    sg_o = om.symmetric_gauge_fixing_map(trav_o, 1e-10)(tensors, lambdas)
    sg_p = dq.get_symmetric_gauge_fixing_map(trav_p, 1e-10)(tensors, lambdas)
    for nid, v in sg_o.items():
        assert _rel(host(sg_p[nid]), v) < 1e-12
    bp_p = dq.get_belief_propagation_map(trav_p, True)
    packed = bp_p(tensors, messages)  # device-resident messages of this layout
    mo = to_oracle_messages({k: host(v) for k, v in packed.items()})
    for node in gp.nodes.values():
        r_p = host(dq.get_one_side_density_matrix(node, tensors, packed))
        r_o = om.one_side_density_matrix(go.nodes[node.id], tensors, mo)
        assert _rel(r_p, r_o) < 1e-11, node.id
    for edge in gp.edges.values():
        r_p = host(dq.get_two_side_density_matrix(edge, tensors, packed))
        r_o = om.two_side_density_matrix(go.edges[edge.id], tensors, mo)
        assert _rel(r_p, r_o) < 1e-11, edge.id
    d_p = float(dq.messages_frob_distance(packed, messages2))
    d_o = om.messages_frob_distance(to_oracle_messages({k: host(v) for k, v in packed.items()}),
                                    to_oracle_messages(messages2))
    assert abs(d_p - d_o) < 1e-12 * max(1.0, d_o)


def test_ghz_tree_pipeline():
    """utils_testing.py:20-100 on the GPU maps: BP to 1e-8, Vidal distance,
    truncation back to phys, BP again gives I/p messages and rho."""
    key = dq.PRNGKey(42)
    tree = dq.get_random_tree_tensor_graph(20, 3, [3, 4, 5], key)
    trav = list(tree.get_traversal_iterator())
    acc = 1e-8
    bp = dq.get_belief_propagation_map(trav)
    vg = dq.get_vidal_gauge_fixing_map(trav, acc)
    vd = dq.get_vidal_gauge_distance_map(trav, acc)
    tensors = tree.init_tensors(dq.get_tensor_bloated_ghz_initializer(key))
    _, k2 = dq.split(key)
    messages = tree.init_messages(dq.get_message_random_nonnegative_initializer(k2))
    dist = 1e300
    it = 0
    while dist > acc and it < 200:
        new = bp(tensors, messages)
        dist = float(dq.messages_frob_distance(new, messages))
        messages = new
        it += 1
    assert dist <= acc
    gam, lam = vg(tensors, messages)
    assert float(vd(gam, lam)) < acc
    # all bonds have exactly phys (=3) non-zero lambdas
    for eid in lam:
        nz = int((host(lam[eid]) > acc).sum())
        assert nz == 3, (eid, host(lam[eid]))


@pytest.mark.parametrize("sync,order", [(True, "bfs"), (False, "bfs"), (False, "dfs")])
def test_bp_matrix_core_path_chi16(sync, order):
    """Degree-3 graphs at chi=16, complex64 run the matrix-core node update;
    compared with the complex128 oracle (FP32 accumulation, rel. 2e-5)."""
    gp, go = regular_pair(24, 16, 11)
    rng = np.random.default_rng(21)
    tensors = random_tensors(gp, rng)
    messages = random_messages(gp, rng)
    trav_p = list(gp.get_traversal_iterator(ordering=order))
    trav_o = go.traversal(ordering=order)
    out_o = om.belief_propagation_map(trav_o, sync)(tensors, to_oracle_messages(messages))
    bp_p = dq.get_belief_propagation_map(trav_p, sync)
    out_p = bp_p({k: v.astype(np.complex64) for k, v in tensors.items()},
                 {k: v.astype(np.complex64) for k, v in messages.items()})
    worst = max(_rel(host(out_p[dq.MessageID(k.src, k.dst)]), v) for k, v in out_o.items())
    assert worst < 2e-5, worst


@pytest.mark.parametrize("zero_tail", [False, True])
def test_vidal_distance_matrix_core_path_chi16(zero_tail):
    """At chi=16, degree 3, complex64 the Vidal distance runs through the
    matrix-core node kernel (diag(lambda^2) messages) and a wave-level nuclear
    norm; compared with the complex128 oracle.  zero_tail: the last four
    lambdas of every other edge are zero, which changes the ideal's count."""
    gp, go = regular_pair(24, 16, 5)
    rng = np.random.default_rng(8)
    tensors = random_tensors(gp, rng)
    lambdas = {}
    for k, eid in enumerate(go.edges):
        lm = np.sort(rng.uniform(0.05, 1.0, 16))[::-1]
        if zero_tail and k % 2:
            lm[12:] = 0.0
        lambdas[eid] = lm
    trav_p = list(gp.get_traversal_iterator())
    trav_o = go.traversal()
    ref = om.vidal_gauge_distance_map(trav_o, 1e-10)(tensors, lambdas)
    got = float(dq.get_vidal_gauge_distance_map(trav_p, 1e-10)(
        {k: v.astype(np.complex64) for k, v in tensors.items()}, lambdas))
    assert abs(got - ref) < 2e-5 * max(1.0, abs(ref)), (got, ref)
    # a canonical state has distance ~0: Vidal gauge of converged messages
    bp = dq.get_belief_propagation_map(trav_p, True)
    t64 = {k: v.astype(np.complex64) for k, v in tensors.items()}
    msgs = {k: v.astype(np.complex64) for k, v in random_messages(gp, rng).items()}
    for _ in range(60):
        msgs = bp(t64, msgs)
    gam, lam = dq.get_vidal_gauge_fixing_map(trav_p, 1e-10)(t64, msgs)
    assert float(dq.get_vidal_gauge_distance_map(trav_p, 1e-10)(gam, lam)) < 1e-4


@pytest.mark.parametrize("chi", [16, 32])
@pytest.mark.parametrize("dtype,tol", [(np.complex128, 1e-10), (np.complex64, 2e-4)])
def test_vidal_gauge_indefinite_messages(chi, dtype, tol):
    """chi=16 edges run the wave-level decomposition (two-sided Jacobi on each
    message, one-sided Jacobi on the core); chi=32 edges the block-level one
    (one-sided Jacobi on each message with the eigenvalue signs from Rayleigh
    quotients, the two-sided Jacobi when a pair +w, -w shares a singular
    subspace).  Hermitian messages with negative, clustered and opposite
    eigenvalues: lambdas and gauge-invariant products against the oracle's
    eigh/SVD path."""
    gp, go = regular_pair(10, chi, 4)
    rng = np.random.default_rng(17)
    tensors = random_tensors(gp, rng)
    msgs = {}
    for el in gp.get_traversal_iterator():
        for nb in el.neighbors:
            q, _ = np.linalg.qr(rng.standard_normal((chi, chi)) + 1j * rng.standard_normal((chi, chi)))
            w = rng.uniform(-0.2, 1.0, chi)
            w[:3] = 0.5  # a threefold eigenvalue
            w[3] = 1e-13  # below eps
            w[4], w[5] = 0.3, -0.3  # opposite eigenvalues
            m = (q * w) @ q.conj().T
            msgs[dq.MessageID(el.id, nb.id)] = m
    trav_p = list(gp.get_traversal_iterator())
    trav_o = go.traversal()
    gam_o, lam_o = om.vidal_gauge_fixing_map(trav_o, 1e-10)(tensors, to_oracle_messages(msgs))
    gam_p, lam_p = dq.get_vidal_gauge_fixing_map(trav_p, 1e-10)(
        {k: v.astype(dtype) for k, v in tensors.items()}, {k: v.astype(dtype) for k, v in msgs.items()})
    for eid, lm in lam_o.items():
        assert np.abs(host(lam_p[eid]) - lm).max() < tol * 10, eid
    # Gamma up to the gauge of degenerate singular values: the symmetric-gauge
    # one-site density matrices are invariant
    gam_h = {nid: host(gam_p[nid]).astype(np.complex128) for nid in gam_o}
    lam_h = {eid: host(lam_p[eid]) for eid in lam_o}
    sg_o = om.symmetric_gauge_fixing_map(trav_o, 1e-10)
    t_o, t_p = sg_o(gam_o, lam_o), sg_o(gam_h, lam_h)
    msg_o, msg_p = om.lambdas2messages(lam_o), om.lambdas2messages(lam_h)
    for node in go.nodes.values():
        r_o = om.one_side_density_matrix(node, t_o, msg_o)
        r_p = om.one_side_density_matrix(node, t_p, msg_p)
        assert np.abs(r_o - r_p).max() < tol * 100, node.id
