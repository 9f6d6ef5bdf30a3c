"""Pinning the oracle (CPU, no GPU).

The reference ships no golden vectors; its tests are invariants, ported here
(dequantizer/src/tests/*, dequantizer/src/mappings/utils_testing.py:20-100,
dequantizer/src/graph/tensor_graph.py:637-1177,
dequantizer/src/quantum_emulator/test_bp_quantum_emulator.py:21-108,
dequantizer/src/graph/test_unit.py:8-36), plus checks fixed by the
mathematics: brute-force contractions, state vectors, closed forms, and
mutations that each check must catch.
"""
import numpy as np
import pytest

from oracle import annealing as oa
from oracle import emulator as oe
from oracle import graph as og
from oracle import mappings as om
from oracle import tensor_ops as ot


# ------------------------------------------------------------------ graph
def test_small_graph_structure():
    """tensor_graph.py:637-795 (hyperedge included)."""
    g = og.TensorGraph()
    ids = [g.add_node(d) for d in (2, 5, 3, 4, 6)]
    assert ids == [0, 1, 2, 3, 4]
    g.add_edge((0, 1), 2)
    g.add_edge((2, 0), 4)
    g.add_edge((1, 2), 5)
    g.add_edge((4, 2, 3), 6)
    n = g.nodes
    assert [n[i].degree for i in range(5)] == [2, 2, 3, 1, 1]
    assert n[0].bond_shape == (2, 4) and n[2].bond_shape == (4, 5, 6)
    assert [e.id for e in n[2].neighbors] == [(2, 0), (1, 2), (4, 2, 3)]
    assert [x.id for x in g.edges[(4, 2, 3)].neighbors] == [4, 2, 3]
    t = g.init_tensors(og.tensor_random_normal_initializer(og.make_key(42)))
    assert t[2].shape == (4, 5, 6, 3) and t[4].shape == (6, 6)
    m = g.init_messages(og.message_random_nonnegative_initializer(og.make_key(1)))
    assert len(m) == 18
    assert m[og.MessageID((4, 2, 3), 4)].shape == (6, 6)
    assert np.all(np.linalg.eigvalsh(m[og.MessageID((4, 2, 3), 4)]) > -1e-5)
    g.add_node()
    assert len(g.traversal(g.edges[(2, 0)], "dfs")) == 9
    assert len(g.traversal(n[2], "bfs")) == 9


def test_lattice_counts_and_shapes():
    """tensor_graph.py:798-1041 on (4, 5, 6), chi 4, phys 3."""
    sizes = (4, 5, 6)
    g = og.nd_lattice(list(sizes), 3, 4)
    t = g.init_tensors(og.tensor_random_normal_initializer(og.make_key(43)))
    m = g.init_messages(og.message_random_nonnegative_initializer(og.make_key(44)))
    assert t[(0, 0, 0)].shape == (4, 4, 4, 3)
    assert t[(1, 1, 1)].shape == (4,) * 6 + (3,)
    edges = (3 * 5 * 6) + (4 * 4 * 6) + (4 * 5 * 5)
    assert g.edges_number == edges and len(m) == 4 * edges and len(t) == 120
    trav = g.traversal()
    assert sum(isinstance(x, og.Node) for x in trav) == 120


def test_traversal_orders_brute_force():
    """bfs/dfs visited-on-pop orders on a tiny graph, enumerated by hand."""
    g = og.TensorGraph()
    for _ in range(3):
        g.add_node()
    g.add_edge((0, 1), 1)
    g.add_edge((1, 2), 1)
    g.add_edge((2, 0), 1)
    bfs = [x.id for x in g.traversal(ordering="bfs")]
    dfs = [x.id for x in g.traversal(ordering="dfs")]
    assert bfs == [0, (0, 1), (2, 0), 1, 2, (1, 2)]
    assert dfs == [0, (2, 0), 2, (1, 2), 1, (0, 1)]


def test_ghz_core_and_bloated_initializer():
    """graph/test_unit.py:8-36."""
    for modes in range(2, 6):
        for p in range(1, 4):
            core = og.ghz_core(p, modes)
            assert abs(np.abs(core).sum() - p) < 1e-12
            for i in range(p):
                assert abs(core[(i,) * modes] - 1.0) < 1e-12
    g = og.TensorGraph()
    for _ in range(4):
        g.add_node(3)
    g.add_edge((0, 1), 5)
    g.add_edge((1, 2), 4)
    g.add_edge((1, 3), 6)
    t = g.init_tensors(og.tensor_bloated_ghz_initializer(og.make_key(42)))
    r = np.tensordot(t[0], t[1], axes=[0, 0])
    r = np.tensordot(t[2], r, axes=[0, 1])
    r = np.tensordot(t[3], r, axes=[0, 2])
    truth = og.ghz_core(3, 4)
    assert np.linalg.norm(r - truth) / np.linalg.norm(truth) < 1e-10


def test_potts_tensor_network_partition_function():
    """Contracting the square-root Potts network with itself gives the
    brute-force Ising partition function (2x2 open lattice)."""
    beta, h = 0.3, 0.5
    w = np.exp(beta * np.outer([1.0, -1.0], [1.0, -1.0]))
    f = np.exp(-h * beta * np.array([1.0, -1.0]))
    g = og.nd_lattice([2, 2], 2, 2)
    t = g.init_tensors(og.potts_initializer(w, f))
    # <psi|psi> = sum_s prod_sites f(s) prod_bonds W(s_a, s_b) for real T
    nodes = list(g.nodes)
    z_bf = 0.0
    for bits in range(16):
        s = {n: (bits >> k) & 1 for k, n in enumerate(nodes)}
        val = np.prod([f[s[n]] for n in nodes])
        for (a, b) in g.edges:
            val *= w[s[a], s[b]]
        z_bf += val
    # amplitude psi(s) = contraction of the sqrt network; |psi|^2 summed = Z
    letters = iter("abcdefghijklmnopqrstuvwxyz")
    bond = {e: next(letters) for e in g.edges}
    ket = ",".join("".join(bond[e.id] for e in g.nodes[n].neighbors) + next(letters) for n in nodes)
    phys = [x[-1] for x in ket.split(",")]
    amp = np.einsum(ket + "->" + "".join(phys), *[t[n] for n in nodes])
    z_tn = np.sum(np.abs(amp) ** 2)
    assert abs(z_tn - z_bf) / z_bf < 1e-12


# ------------------------------------------------------------ tensor ops
def test_pass_through_node_closed_form_and_invariants():
    rng = np.random.default_rng(0)
    t = rng.standard_normal((3, 4, 2, 2)) + 1j * rng.standard_normal((3, 4, 2, 2))
    msgs = []
    for d in (3, 4, 2):
        a = rng.standard_normal((d, d)) + 1j * rng.standard_normal((d, d))
        msgs.append(a @ a.conj().T)
    out = ot.pass_through_node(t, msgs)
    # closed form with einsum (message on the ket side: M[bra, ket])
    ref0 = np.einsum("kacs,ab,cd,lbds->kl", t.conj(), msgs[1], msgs[2], t)
    ref0 = ref0 / np.trace(ref0)
    assert np.allclose(out[0], ref0, atol=1e-12)
    for o in out:
        assert np.allclose(o, o.conj().T, atol=1e-12)
        assert np.all(np.linalg.eigvalsh(o) > -1e-12)
        assert abs(np.trace(o) - 1) < 1e-12
    # a transposed message (ket/bra swapped) must change the result
    bad = [msgs[0], msgs[1].T, msgs[2]]
    assert not np.allclose(ot.pass_through_node(t, bad)[0], out[0], atol=1e-6)


def test_decompose_gate_reconstructs():
    rng = np.random.default_rng(1)
    g = np.diag(np.exp(1j * rng.standard_normal(4))).reshape(2, 2, 2, 2)
    h1, h2 = ot.decompose_gate(g, 1e-10)
    assert h1.shape[0] == 2  # Ising-like gates have operator-Schmidt rank 2
    rec = np.einsum("qai,qbj->abij", h1, h2)
    assert np.allclose(rec, g, atol=1e-12)
    cnot = np.array([[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 0, 1], [0, 0, 1, 0]], dtype=complex).reshape(2, 2, 2, 2)
    h1, h2 = ot.decompose_gate(cnot, 1e-10)
    assert np.allclose(np.einsum("qai,qbj->abij", h1, h2), cnot, atol=1e-12)


def test_find_rank_closed_form():
    lm = np.array([1.0, 0.5, 1e-3, 1e-6])
    assert ot.find_rank(lm, None, None) == 4
    assert ot.find_rank(lm, None, 2) == 2
    # tail norms relative to |lm|: [8.9e-7, 8.9e-4, 0.447, 1]
    assert ot.find_rank(lm, 1e-5, None) == 3
    assert ot.find_rank(lm, 1e-2, None) == 2


def _tree_state_vector(g, tensors):
    """Exact contraction of a tree network into a state vector (node order of
    the default traversal)."""
    nodes = [x.id for x in g.traversal() if isinstance(x, og.Node)]
    letters = iter("abcdefghijklmnopqrstuvwxyzABCDEFGHIJKLMNOPQRSTUVWXYZ")
    bond = {e: next(letters) for e in g.edges}
    phys = {n: next(letters) for n in nodes}
    ins = ",".join("".join(bond[e.id] for e in g.nodes[n].neighbors) + phys[n] for n in nodes)
    psi = np.einsum(ins + "->" + "".join(phys[n] for n in nodes), *[tensors[n] for n in nodes], optimize=True)
    return psi / np.linalg.norm(psi), nodes


def test_bp_exact_on_trees_and_lambda_are_schmidt_values():
    """On a tree, converged BP gives the exact one-site density matrices and the
    Vidal lambdas are the Schmidt values of the cut (brute-force SVD)."""
    key = og.make_key(5)
    tree = og.random_tree(7, 2, [2, 3], key)
    rng = np.random.default_rng(2)
    tensors = {}
    for n in tree.nodes.values():
        shape = (*n.bond_shape, 2)
        tensors[n.id] = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    trav = tree.traversal()
    msgs = tree.init_messages(og.message_random_nonnegative_initializer(key))
    bp = om.belief_propagation_map(trav)
    for _ in range(40):
        msgs = bp(tensors, msgs)
    psi, order = _tree_state_vector(tree, tensors)
    nq = len(order)
    for k, nid in enumerate(order):
        v = np.moveaxis(psi, k, 0).reshape(2, -1)
        exact = v @ v.conj().T
        bp_rho = om.one_side_density_matrix(tree.nodes[nid], tensors, msgs)
        assert np.allclose(bp_rho, exact, atol=1e-10)
    gam, lam = om.vidal_gauge_fixing_map(trav, 1e-12)(tensors, msgs)
    assert om.vidal_gauge_distance_map(trav, 1e-12)(gam, lam) < 1e-9
    for eid, e in tree.edges.items():
        # bipartition by removing the edge
        a = e.neighbors[0].id
        side, stack = {a}, [a]
        while stack:
            x = stack.pop()
            for ed in tree.nodes[x].neighbors:
                if ed.id == eid:
                    continue
                for y in ed.neighbors:
                    if y.id not in side:
                        side.add(y.id)
                        stack.append(y.id)
        axes = [order.index(x) for x in order if x in side]
        rest = [i for i in range(nq) if i not in axes]
        mat = np.transpose(psi, axes + rest).reshape(2 ** len(axes), -1)
        sv = np.linalg.svd(mat, compute_uv=False)
        sv = sv / np.linalg.norm(sv)
        full = np.zeros(max(len(sv), len(lam[eid])))
        full[: len(sv)] = sv
        assert np.allclose(np.sort(lam[eid])[::-1], full[: len(lam[eid])], atol=1e-8)
        assert np.all(full[len(lam[eid]):] < 1e-8)
        # two-site density matrix is exact on trees too
        ia, ib = order.index(e.neighbors[0].id), order.index(e.neighbors[1].id)
        v = np.moveaxis(psi, [ia, ib], [0, 1]).reshape(4, -1)
        r2 = om.two_side_density_matrix(e, tensors, msgs)
        assert np.allclose(r2, v @ v.conj().T, atol=1e-10)


def test_ghz_tree_gauge_fixing():
    """mappings/utils_testing.py:20-100 (20 nodes, phys 3, bonds {3,4,5}, 1e-8)."""
    key = og.make_key(42)
    tree = og.random_tree(20, 3, [3, 4, 5], key)
    trav = tree.traversal()
    acc = 1e-8
    bp = om.belief_propagation_map(trav)
    vg = om.vidal_gauge_fixing_map(trav, acc)
    vd = om.vidal_gauge_distance_map(trav, acc)
    sg = om.symmetric_gauge_fixing_map(trav, acc)
    tensors = tree.init_tensors(og.tensor_bloated_ghz_initializer(key))
    _, k2 = og.split_key(key)
    msgs = tree.init_messages(og.message_random_nonnegative_initializer(k2))
    dist = np.inf
    while dist > acc:
        new = bp(tensors, msgs)
        dist = om.messages_frob_distance(new, msgs)
        msgs = new
    gam, lam = vg(tensors, msgs)
    assert vd(gam, lam) < acc
    tt, tl = tree.truncate(gam, lam, acc)
    assert all(e.dimension == 3 for e in tree.edges.values())
    tensors = sg(tt, tl)
    msgs = tree.init_messages(og.message_random_nonnegative_initializer(og.split_key(k2)[1]))
    dist = np.inf
    while dist > acc:
        new = bp(tensors, msgs)
        dist = om.messages_frob_distance(new, msgs)
        msgs = new
    for m in msgs.values():
        assert np.abs(m - np.eye(3) / 3).max() < acc
    for n in tree.nodes.values():
        assert np.linalg.norm(om.one_side_density_matrix(n, tensors, msgs) - np.eye(3) / 3) < acc


def test_ghz_circuit_on_lattice():
    """tensor_graph.py:1075-1177: H + 15 CNOTs on 4x4, rho = I/2; uncompute -> |0>."""
    lat = og.nd_lattice([4, 4], 2, 1)
    t = lat.init_tensors(og.tensor_std_state_initializer())
    had = np.array([[1, 1], [1, -1]], dtype=complex)
    cnot = np.array([[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 0, 1], [0, 0, 1, 0]], dtype=complex).reshape(2, 2, 2, 2)
    seq = [((0, 0), (0, 1)), ((0, 1), (0, 2)), ((0, 2), (0, 3)), ((0, 0), (1, 0)), ((0, 1), (1, 1)),
           ((0, 2), (1, 2)), ((0, 3), (1, 3)), ((1, 0), (2, 0)), ((2, 0), (2, 1)), ((2, 1), (2, 2)),
           ((2, 2), (2, 3)), ((2, 3), (3, 3)), ((3, 3), (3, 2)), ((2, 1), (3, 1)), ((2, 0), (3, 0))]
    t = lat.apply1((0, 0), t, had)
    for a, b in seq:
        t = lat.apply2(a, b, t, cnot, 1e-8)
    trav = lat.traversal()
    bp = om.belief_propagation_map(trav)
    msgs = lat.init_messages(og.message_random_nonnegative_initializer(og.make_key(42)))
    dist = np.inf
    while dist > 1e-8:
        new = bp(t, msgs)
        dist = om.messages_frob_distance(new, msgs)
        msgs = new
    for n in lat.nodes.values():
        assert np.allclose(om.one_side_density_matrix(n, t, msgs), np.eye(2) / 2)
    for a, b in reversed(seq):
        t = lat.apply2(a, b, t, cnot, 1e-8)
    t = lat.apply1((0, 0), t, had)
    msgs = lat.init_messages(og.message_random_nonnegative_initializer(og.make_key(43)))
    dist = np.inf
    while dist > 1e-8:
        new = bp(t, msgs)
        dist = om.messages_frob_distance(new, msgs)
        msgs = new
    gam, lam = om.vidal_gauge_fixing_map(trav, 1e-8)(t, msgs)
    gam, lam = lat.truncate(gam, lam, 1e-8)
    assert all(e.dimension == 1 for e in lat.edges.values())
    t = om.symmetric_gauge_fixing_map(trav, 1e-8)(gam, lam)
    msgs = lat.init_messages(og.message_random_nonnegative_initializer(og.make_key(44)))
    for _ in range(5):
        msgs = bp(t, msgs)
    for n in lat.nodes.values():
        assert np.allclose(om.one_side_density_matrix(n, t, msgs), np.array([[1, 0], [0, 0]]), atol=1e-8)


# -------------------------------------------------------------- emulator
# Tolerances of the reference test; the heavy-hex one is 5e-4 instead of 1e-4:
# the gates here come from NumPy draws (the reference's JAX draws cannot be
# reproduced) and the loop-induced BP error is gate dependent (1.3e-4 here).
@pytest.mark.parametrize("build,layers,acc", [(og.heavy_cube, 2, 3e-3), (og.small_heavy_hex, 4, 5e-4),
                                              (lambda: og.random_tree(15, 2, [1] * 14, og.make_key(42)), 4, 1e-4)])
def test_bp_emulator_vs_state_vector(build, layers, acc):
    """test_bp_quantum_emulator.py:21-108 on the oracle emulators."""
    g = build()
    key = og.make_key(42)
    key, sub = og.split_key(key)
    ex = oe.ExactQuantumEmulator(g, sub)
    bpe = oe.BPQuantumEmulator(g, sub, 1024, g.edges_number, belief_propagation_accuracy=1e-6)
    trav = g.traversal()
    key, sub = og.split_key(key)
    rng = og.key_rng(sub)
    state = {x.id: int(rng.integers(0, 2)) for x in trav if isinstance(x, og.Node)}
    ex.set_to_product_state(state)
    bpe.set_to_product_state(state)
    for nid, s in state.items():
        c = np.zeros((2, 2))
        c[s, s] = 1
        assert np.allclose(bpe.dens_q1(nid), c, atol=1e-8)
        assert bpe.measure(nid) == s
    grng = np.random.default_rng(9)
    for _ in range(layers):
        for el in trav:
            if isinstance(el, og.Node):
                x = grng.standard_normal((2, 2, 2))
                q, _ = np.linalg.qr(x[..., 0] + 1j * x[..., 1])
                ex.apply_q1(q, el.id)
                bpe.apply_q1(q, el.id)
            else:
                gate = np.diag(np.exp(1j * grng.standard_normal(4))).reshape(2, 2, 2, 2)
                ex.apply_q2(gate, *el.id)
                bpe.apply_q2(gate, *el.id)
    for nid in g.nodes:
        assert np.linalg.norm(ex.dens_q1(nid) - bpe.dens_q1(nid)) < acc
    ex.prng_key = bpe.prng_key
    for nid in g.nodes:
        assert ex.measure(nid) == bpe.measure(nid)


def test_annealing_two_spins_trotter_closed_form():
    """Two coupled spins: the BP annealer (a tree, no truncation) equals the
    exact Trotter product computed with 4x4 matrices."""
    pairs = np.array([[0, 1]])
    j = np.array([-1.0])
    h = np.array([0.3, -0.2])
    sched = oa.scheduler(0.1, lambda t: 1 - t, 20)
    _, rho1, rho2 = oa.run_quantum_annealer(pairs, j, h, sched, og.make_key(0), 4, 1, 1e-10, 100, False, "dfs",
                                            two_site=True)
    psi = np.kron(oa.HADAMARD @ [1, 0], oa.HADAMARD @ [1, 0]).astype(complex)
    for mix, cpl in sched:
        psi = oa.coupling_gate(j[0], h[0], h[1], 1, 1, cpl).reshape(4, 4) @ psi
        psi = np.kron(oa.mixing_gate(mix), oa.mixing_gate(mix)) @ psi
    r = np.outer(psi, psi.conj())
    assert np.allclose(rho2[0], r, atol=1e-10)
    assert np.allclose(rho1[0], np.einsum("abcb->ac", r.reshape(2, 2, 2, 2)), atol=1e-10)
    e_exact = float(np.real(np.vdot(psi, np.diag([-j[0] - h[0] - h[1], j[0] - h[0] + h[1],
                                                   j[0] + h[0] - h[1], -j[0] + h[0] + h[1]]) @ psi)))
    assert abs(oa.bp_energy(rho1, rho2, h, j) - e_exact) < 1e-10


def test_mutated_oracle_is_caught():
    """Each pinning check catches a planted mistake in the oracle (on the CPU)."""
    rng = np.random.default_rng(3)
    t = rng.standard_normal((2, 2, 2)) + 1j * rng.standard_normal((2, 2, 2))
    m = [np.eye(2) + 0.3 * (rng.standard_normal((2, 2)) + 1j * rng.standard_normal((2, 2))) for _ in range(2)]
    m = [x @ x.conj().T for x in m]
    good = ot.pass_through_node(t, m)[0]
    orig = ot.apply_matrix_on_axis
    try:  # transposed operand in the message contraction
        ot.apply_matrix_on_axis = lambda tensor, matrix, axis: orig(tensor, matrix.T, axis)
        assert not np.allclose(ot.pass_through_node(t, m)[0], good, atol=1e-8)
    finally:
        ot.apply_matrix_on_axis = orig


# ------------------------------------------------- state-vector emulator
def _bare_exact(n, psi):
    ex = oe.ExactQuantumEmulator.__new__(oe.ExactQuantumEmulator)
    ex.n, ex.node2pos, ex.state = n, {i: i for i in range(n)}, psi.copy()
    return ex


def _brute_q1(n, p, g):
    """2^n x 2^n operator of a one-qubit gate by loops over basis states; qubit
    p is index bit n-1-p (axis 0 of the row-major reshape is the most
    significant)."""
    b = n - 1 - p
    u = np.zeros((2**n, 2**n), dtype=complex)
    for i in range(2**n):
        for o in range(2):
            u[(i & ~(1 << b)) | (o << b), i] += g[o, (i >> b) & 1]
    return u


def _brute_q2(n, pa, pb, g):
    ba, bb = n - 1 - pa, n - 1 - pb
    u = np.zeros((2**n, 2**n), dtype=complex)
    for i in range(2**n):
        x1, x2 = (i >> ba) & 1, (i >> bb) & 1
        for o1 in range(2):
            for o2 in range(2):
                j = (i & ~(1 << ba) & ~(1 << bb)) | (o1 << ba) | (o2 << bb)
                u[j, i] += g[o1, o2, x1, x2]
    return u


def _brute_rdm(n, psi, p):
    b = n - 1 - p
    rho = np.zeros((2, 2), dtype=complex)
    for i in range(2**n):
        if (i >> b) & 1 == 0:
            a0, a1 = psi[i], psi[i | (1 << b)]
            rho += np.array([[a0 * a0.conjugate(), a0 * a1.conjugate()], [a1 * a0.conjugate(), a1 * a1.conjugate()]])
    return rho


def test_exact_emulator_against_brute_force_operators():
    """The oracle's tensordot/moveaxis state-vector updates equal explicit
    operators built bit by bit, for ordered, reversed and non-adjacent qubit
    pairs; its density matrices equal the bitwise partial trace; a planted
    swap of the two qubits of apply_q2 is caught."""
    n = 5
    rng = np.random.default_rng(11)
    psi = rng.standard_normal(2**n) + 1j * rng.standard_normal(2**n)
    psi /= np.linalg.norm(psi)
    ex = _bare_exact(n, psi)
    ref = psi.copy()
    for p in range(n):
        g = rng.standard_normal((2, 2)) + 1j * rng.standard_normal((2, 2))
        ex.apply_q1(g, p)
        ref = _brute_q1(n, p, g) @ ref
        assert np.allclose(ex.state, ref, atol=1e-12)
    for pa, pb in [(0, 1), (3, 1), (0, 4), (4, 2)]:
        g = (rng.standard_normal((2, 2, 2, 2)) + 1j * rng.standard_normal((2, 2, 2, 2))) / 2
        ex.apply_q2(g, pa, pb)
        ref = _brute_q2(n, pa, pb, g) @ ref
        assert np.allclose(ex.state, ref, atol=1e-12)
    for p in range(n):
        assert np.allclose(ex.dens_q1(p), _brute_rdm(n, ref, p), atol=1e-12)
    # planted mistake: the two qubits of a two-qubit gate exchanged
    g = (rng.standard_normal((2, 2, 2, 2)) + 1j * rng.standard_normal((2, 2, 2, 2))) / 2
    bad = _bare_exact(n, ref)
    bad.apply_q2(g, 1, 3)
    good = _brute_q2(n, 3, 1, g) @ ref
    assert not np.allclose(bad.state, good, atol=1e-6)
