"""GPU parity of the dictionary API of TensorGraph (apply1, apply2,
apply2_to_vidal_canonical, truncate) against the oracle, plus the reference's
GHZ-circuit known-answer test on a 4x4 lattice
(dequantizer/src/graph/tensor_graph.py:1075-1177).

Simple-update results are compared through quantities that do not depend on
the SVD's column phases: the lambdas, the truncation error and entropy, and
the two-site tensor theta = sum_k Gamma_1[.., k, ..] lambda_k Gamma_2[.., k, ..].
"""
import numpy as np
import pytest

from tests.conftest import gpu_available
from tests.helpers import host

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not gpu_available(), reason="needs a ROCm GPU")]

import dequantizer_amd as dq  # noqa: E402
from dequantizer_amd import local  # noqa: E402
from oracle import graph as og  # noqa: E402
from oracle import tensor_ops as ot  # noqa: E402

CNOT = np.array([[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 0, 1], [0, 0, 1, 0]], dtype=np.complex128).reshape(2, 2, 2, 2)
HADAMARD = np.array([[1, 1], [1, -1]], dtype=np.complex128)


def _cgauss(rng, shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def _random_gate(rng, p=2):
    q, _ = np.linalg.qr(_cgauss(rng, (p * p, p * p)))
    return q.reshape(p, p, p, p)


def _rel(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(np.asarray(b)), 1e-300)


def _gate_zoo(rng):
    swap = np.eye(4, dtype=np.complex128)[[0, 2, 1, 3]].reshape(2, 2, 2, 2)
    diag = np.diag(np.exp(1j * rng.uniform(0, 2 * np.pi, 4))).reshape(2, 2, 2, 2)
    return [("cnot", CNOT), ("swap", swap), ("identity", np.eye(4, dtype=np.complex128).reshape(2, 2, 2, 2)),
            ("diagonal", diag), ("random", _random_gate(rng)), ("random_p3", _random_gate(rng, 3)),
            ("nonunitary", _cgauss(rng, (2, 2, 2, 2)))]


@pytest.mark.parametrize("which", range(7))
def test_device_gate_split(which):
    """dq_decompose_gates: the halves multiply back to the gate, the ranks
    and the singular values (norms of the halves' slices) match the oracle's
    SVD-based split, and the rank bound holds."""
    import torch
    from dequantizer_amd.emulator import decompose_gates_device, gate_rank_bounds
    name, gate = _gate_zoo(np.random.default_rng(11))[which]
    p = gate.shape[0]
    d_g = torch.from_numpy(gate.reshape(1, -1).copy()).cuda()
    halves = torch.empty((1, 2 * p ** 4), dtype=torch.complex128, device="cuda")
    ranks = torch.zeros(1, dtype=torch.int32, device="cuda")
    decompose_gates_device(d_g, p, 1e-10, halves, ranks.data_ptr(), 1, torch.cuda.current_stream().cuda_stream)
    r = int(ranks.item())
    h = host(halves).reshape(2, p ** 4)
    h1 = h[0].reshape(p * p, p, p)
    h2 = h[1].reshape(p * p, p, p)
    o1, o2 = ot.decompose_gate(gate, 1e-10)
    assert r == o1.shape[0], name
    assert r <= gate_rank_bounds(gate)[0]
    assert np.abs(h1[r:]).max(initial=0) == 0 and np.abs(h2[r:]).max(initial=0) == 0
    rebuilt = np.einsum("qac,qbd->abcd", h1[:r], h2[:r])
    assert np.abs(rebuilt - gate).max() < 1e-13 * max(1.0, np.abs(gate).max()), name
    s_dev = np.linalg.norm(h1[:r].reshape(r, -1), axis=1) ** 2
    s_ora = np.linalg.norm(o1.reshape(r, -1), axis=1) ** 2
    assert np.abs(s_dev - s_ora).max() < 1e-12, name


@pytest.mark.parametrize("shape", [(2,), (3, 4, 2), (5, 2, 3, 2), (2, 2, 2, 3)])
@pytest.mark.parametrize("dtype,tol", [(np.complex128, 1e-13), (np.complex64, 2e-6)])
def test_apply1_matches_oracle(shape, dtype, tol):
    rng = np.random.default_rng(1)
    t = _cgauss(rng, shape).astype(dtype)
    g = _cgauss(rng, (shape[-1], shape[-1]))
    out = host(local.apply1(t, g))
    assert out.dtype == dtype
    assert _rel(out, ot.apply1(t.astype(np.complex128), g)) < tol


@pytest.mark.parametrize("shape", [(3, 2), (3, 4, 2), (2, 3, 4, 2)])
@pytest.mark.parametrize("dtype,tol", [(np.complex128, 1e-12), (np.complex64, 2e-6)])
def test_apply_half_gate_matches_oracle(shape, dtype, tol):
    rng = np.random.default_rng(2)
    t = _cgauss(rng, shape).astype(dtype)
    gate = _random_gate(rng)
    h1, h2 = ot.decompose_gate(gate, 1e-12)
    t128 = t.astype(np.complex128)
    for index in range(len(shape) - 1):
        a, b, r = local.apply_gate_halves(t, t, index, index, gate, 1e-12)
        assert r == h1.shape[0]
        a, b = host(a), host(b)
        oa, ob = ot.apply_half_gate(t128, h1, index), ot.apply_half_gate(t128, h2, index)
        assert a.shape == oa.shape and b.shape == ob.shape
        # the halves carry the SVD's column phases; the pair contracted over
        # the fused bond does not
        assert _rel(np.tensordot(a, b, axes=([index], [index])), np.tensordot(oa, ob, axes=([index], [index]))) < tol
        assert abs(np.linalg.norm(a) - 1) < tol and abs(np.linalg.norm(b) - 1) < tol


def _theta(g1, g2, lam, i1, i2):
    g1 = np.moveaxis(np.asarray(g1), i1, -1)
    g2 = np.moveaxis(np.asarray(g2), i2, -1)
    return np.tensordot(g1 * np.asarray(lam), g2, axes=([g1.ndim - 1], [g2.ndim - 1]))


def _su_case(rng, dims1, dims2, i1, i2):
    """A two-node pair with random tensors and random positive lambdas."""
    de = dims1[i1]
    assert dims2[i2] == de
    t1 = _cgauss(rng, (*dims1, 2))
    t2 = _cgauss(rng, (*dims2, 2))
    shared = np.sort(rng.uniform(0.1, 1.0, de))[::-1]
    lam1 = [shared if k == i1 else np.sort(rng.uniform(0.1, 1.0, d))[::-1] for k, d in enumerate(dims1)]
    lam2 = [shared if k == i2 else np.sort(rng.uniform(0.1, 1.0, d))[::-1] for k, d in enumerate(dims2)]
    return t1, t2, lam1, lam2


def _graph_for(dims1, dims2, i1, i2):
    """A TensorGraph whose nodes 0 and 1 have these bond orders (outer bonds
    to leaf nodes): the leaves before the shared bond are added first, then
    the shared edge, then the remaining leaves."""
    g = dq.TensorGraph()
    a, b = g.add_node(2), g.add_node(2)
    leaves = {}

    def leaf(node, k, d):
        lf = g.add_node(2)
        g.add_edge((node, lf), d)
        leaves[(node, k)] = (node, lf)

    for node, dims, shared in ((a, dims1, i1), (b, dims2, i2)):
        for k in range(shared):
            leaf(node, k, dims[k])
    g.add_edge((a, b), dims1[i1])
    for node, dims, shared in ((a, dims1, i1), (b, dims2, i2)):
        for k in range(shared + 1, len(dims)):
            leaf(node, k, dims[k])
    return g, a, b, leaves


@pytest.mark.parametrize("dims1,dims2,i1,i2", [
    ((3, 4, 2), (4, 3), 1, 0),
    ((4, 3), (2, 4, 3), 0, 1),
    ((2, 2), (2, 2), 0, 0),
    ((1, 3), (2, 3), 1, 1),
    ((16, 16, 16), (16, 16, 16), 2, 0),  # cp = 32: FP64 matrix-core environment / Gamma path
])
@pytest.mark.parametrize("accuracy,max_rank", [(None, None), (None, 3), (1e-1, None), (None, 16)])
@pytest.mark.parametrize("dtype,tol", [(np.complex128, 1e-9), (np.complex64, 1e-4)])
def test_simple_update_matches_oracle(dims1, dims2, i1, i2, accuracy, max_rank, dtype, tol):
    rng = np.random.default_rng(5)
    t1, t2, lam1, lam2 = _su_case(rng, dims1, dims2, i1, i2)
    g, a, b, leaves = _graph_for(dims1, dims2, i1, i2)
    na, nb = g.get_node(a), g.get_node(b)
    assert na.indices[(a, b)] == i1 and nb.indices[(a, b)] == i2
    lambdas = {}
    for node, lams, shared in ((a, lam1, i1), (b, lam2, i2)):
        for k, lm in enumerate(lams):
            lambdas[(a, b) if k == shared else leaves[(node, k)]] = lm
    tensors = {a: t1.astype(dtype), b: t2.astype(dtype)}
    gate = _random_gate(rng)
    if dims1 == (16, 16, 16) and (accuracy is not None or max_rank is None):
        # the full spectrum (16 * 4 values) needs chi = 64 padding: outside
        # the kernel's chi * phys <= 64, chi * phys^2 <= 128
        with pytest.raises(dq.DQError, match="supports chi"):
            g.apply2_to_vidal_canonical(a, b, tensors, lambdas, gate, 1e-10, accuracy, max_rank)
        return
    tensors, lambdas, err, ent = g.apply2_to_vidal_canonical(a, b, tensors, lambdas, gate, 1e-10, accuracy, max_rank)
    h1, h2 = ot.decompose_gate(gate, 1e-10)
    o1, o2, ol, oerr, oent = ot.simple_update(t1, t2, lam1, lam2, h1, h2, i1, i2, 1e-10, accuracy, max_rank)
    lam = host(lambdas[(a, b)])
    assert lam.shape == ol.shape
    assert tuple(host(tensors[a]).shape) == o1.shape and tuple(host(tensors[b]).shape) == o2.shape
    assert g.get_edge((a, b)).dimension == ol.shape[0]
    assert na.bond_shape == o1.shape[:-1] and nb.bond_shape == o2.shape[:-1]
    assert np.abs(lam - ol).max() < tol * max(1.0, np.abs(ol).max())
    assert abs(float(host(err)) - oerr) < tol
    assert abs(float(host(ent)) - oent) < tol
    assert _rel(_theta(host(tensors[a]), host(tensors[b]), lam, i1, i2), _theta(o1, o2, ol, i1, i2)) < tol


@pytest.mark.parametrize("dims1,dims2,i1,i2", [
    ((32, 32, 32), (32, 32, 32), 2, 0),  # the degree-3 nodes of the 127-qubit heavy-hex lattice
    ((32, 32), (32, 32, 32), 1, 0),  # a degree-2 bridge node against a degree-3 node
    ((32, 32), (32, 32), 0, 1),
])
@pytest.mark.parametrize("max_rank", [32, 20])
@pytest.mark.parametrize("dtype,tol", [(np.complex128, 1e-9), (np.complex64, 2e-4)])
def test_simple_update_chi32_matches_oracle(dims1, dims2, i1, i2, max_rank, dtype, tol):
    """chi = 32 with a rank-2 gate: a 64 x 64 core (chi * phys = 64 storage),
    truncated to max_rank, against the oracle's QR-based update
    (tensor_ops.py:265-295,318-389; a 2048 x 64 QR per degree-3 side)."""
    rng = np.random.default_rng(11)
    t1, t2, lam1, lam2 = _su_case(rng, dims1, dims2, i1, i2)
    g, a, b, leaves = _graph_for(dims1, dims2, i1, i2)
    lambdas = {}
    for node, lams, shared in ((a, lam1, i1), (b, lam2, i2)):
        for k, lm in enumerate(lams):
            lambdas[(a, b) if k == shared else leaves[(node, k)]] = lm
    tensors = {a: t1.astype(dtype), b: t2.astype(dtype)}
    # the kicked-Ising coupling RZZ(-pi/2) (operator-Schmidt rank 2)
    gate = np.diag(np.exp(1j * np.pi / 4 * np.array([1, -1, -1, 1]))).reshape(2, 2, 2, 2)
    tensors, lambdas, err, ent = g.apply2_to_vidal_canonical(a, b, tensors, lambdas, gate, 1e-10, None, max_rank)
    h1, h2 = ot.decompose_gate(gate, 1e-10)
    o1, o2, ol, oerr, oent = ot.simple_update(t1, t2, lam1, lam2, h1, h2, i1, i2, 1e-10, None, max_rank)
    lam = host(lambdas[(a, b)])
    assert lam.shape == ol.shape == (max_rank,)
    assert tuple(host(tensors[a]).shape) == o1.shape and tuple(host(tensors[b]).shape) == o2.shape
    assert np.abs(lam - ol).max() < tol * max(1.0, np.abs(ol).max())
    assert oerr > 1e-3  # the cap truncates: the error is a real check
    assert abs(float(host(err)) - oerr) < tol
    assert abs(float(host(ent)) - oent) < tol
    assert _rel(_theta(host(tensors[a]), host(tensors[b]), lam, i1, i2), _theta(o1, o2, ol, i1, i2)) < tol


def test_truncate_matches_oracle():
    rng = np.random.default_rng(3)
    gp, go = dq.get_nd_lattice([2, 3], 2, 4), og.nd_lattice([2, 3], 2, 4)
    tensors, lambdas = {}, {}
    for nid, node in go.nodes.items():
        tensors[nid] = _cgauss(rng, (*node.bond_shape, 2))
    for eid in go.edges:
        lm = np.sort(rng.uniform(0, 1, 4))[::-1]
        lm[2:] *= 1e-3
        lambdas[eid] = lm
    pt, pl = gp.truncate(dict(tensors), dict(lambdas), 1e-2)
    ot_, ol = go.truncate(dict(tensors), dict(lambdas), 1e-2)
    for eid in ol:
        assert np.array_equal(host(pl[eid]), ol[eid])
        assert gp.get_edge(eid).dimension == go.edges[eid].dimension
    for nid in ot_:
        assert np.array_equal(host(pt[nid]), ot_[nid])
        assert gp.get_node(nid).bond_shape == go.nodes[nid].bond_shape


def test_ghz_circuit_on_lattice():
    """tensor_graph.py:1075-1177: a 16-qubit GHZ circuit built with apply1 and
    the non-truncating apply2 gives one-site density matrices I/2 after BP;
    the inverse circuit, Vidal gauge and truncation return every bond to 1 and
    every qubit to |0><0|."""
    lattice = dq.get_nd_lattice([4, 4], 2, 1)
    tensors = lattice.init_tensors(dq.get_tensor_std_state_initializer())
    forward = [((0, 0), (0, 1)), ((0, 1), (0, 2)), ((0, 2), (0, 3)), ((0, 0), (1, 0)), ((0, 1), (1, 1)),
               ((0, 2), (1, 2)), ((0, 3), (1, 3)), ((1, 0), (2, 0)), ((2, 0), (2, 1)), ((2, 1), (2, 2)),
               ((2, 2), (2, 3)), ((2, 3), (3, 3)), ((3, 3), (3, 2)), ((2, 1), (3, 1)), ((2, 0), (3, 0))]
    tensors = lattice.apply1((0, 0), tensors, HADAMARD)
    for a, b in forward:
        tensors = lattice.apply2(a, b, tensors, CNOT, 1e-8)
    for nid, node in lattice.nodes.items():
        assert tuple(host(tensors[nid]).shape[:-1]) == node.bond_shape
    trav = list(lattice.get_traversal_iterator())
    bp = dq.get_belief_propagation_map(trav)

    def converge(tensors, key):
        messages = lattice.init_messages(dq.get_message_random_nonnegative_initializer(dq.PRNGKey(key)))
        dist, it = 1e300, 0
        while dist > 1e-8 and it < 500:
            new = bp(tensors, messages)
            dist = float(dq.messages_frob_distance(new, messages))
            messages = new
            it += 1
        assert dist <= 1e-8
        return messages

    messages = converge(tensors, 42)
    for i in range(4):
        for j in range(4):
            rho = host(dq.get_one_side_density_matrix(lattice.get_node((i, j)), tensors, messages))
            assert np.abs(rho - np.eye(2) / 2).max() < 1e-7
    for a, b in reversed(forward):
        tensors = lattice.apply2(a, b, tensors, CNOT, 1e-8)
    tensors = lattice.apply1((0, 0), tensors, HADAMARD)
    messages = converge(tensors, 43)
    gam, lam = dq.get_vidal_gauge_fixing_map(trav, 1e-8)(tensors, messages)
    gam, lam = lattice.truncate(gam, lam, 1e-8)
    tensors = dq.get_symmetric_gauge_fixing_map(trav, 1e-8)(gam, lam)
    messages = converge(tensors, 44)
    zero = np.array([[1, 0], [0, 0]], dtype=np.complex128)
    for i in range(4):
        for j in range(4):
            node = lattice.get_node((i, j))
            for d, nb in zip(node.bond_shape, node.neighbors):
                assert nb.dimension == d == 1
            rho = host(dq.get_one_side_density_matrix(node, tensors, messages))
            assert np.abs(rho - zero).max() < 1e-7
