"""The node update at its largest configured shape: degree 6, chi = 16,
phys 2 (one node tensor of 33.5 M elements, 256 MiB in complex64), as on the
16^3 periodic lattice.  A star whose centre has six chi = 16 bonds is a tree,
so its BP messages are exact marginals; one synchronous sweep is compared
with the oracle (complex128 to rounding, complex64 to FP32 accumulation)."""
import numpy as np
import pytest

from tests.conftest import gpu_available

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not gpu_available(), reason="needs a ROCm GPU")]

import dequantizer_amd as dq  # noqa: E402
from oracle import graph as og  # noqa: E402
from oracle import mappings as om  # noqa: E402
from tests.helpers import host, to_oracle_messages  # noqa: E402


def _star():
    gp, go = dq.TensorGraph(), og.TensorGraph()
    for g in (gp, go):
        for _ in range(7):
            g.add_node(2)
        for i in range(1, 7):
            g.add_edge((0, i), 16)
    return gp, go


@pytest.fixture(scope="module")
def star_case():
    gp, go = _star()
    rng = np.random.default_rng(3)
    tensors = {}
    for n in go.nodes.values():
        shape = n.bond_shape + (2,)
        tensors[n.id] = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)) / np.sqrt(2.0 * 16 ** 6)
    msgs = {}
    for el in gp.get_traversal_iterator():
        for nb in el.neighbors:
            a = rng.standard_normal((16, 16)) + 1j * rng.standard_normal((16, 16))
            m = a @ a.conj().T
            msgs[dq.MessageID(el.id, nb.id)] = m / np.trace(m)
    trav_o = go.traversal()
    ref = om.belief_propagation_map(trav_o, True)(tensors, to_oracle_messages(msgs))
    return gp, tensors, msgs, ref


@pytest.mark.parametrize("dtype,tol", [(np.complex128, 1e-11), (np.complex64, 2e-5)])
def test_degree6_chi16_node_update_matches_oracle(star_case, dtype, tol):
    gp, tensors, msgs, ref = star_case
    bp = dq.get_belief_propagation_map(list(gp.get_traversal_iterator()), True)
    out = bp({k: v.astype(dtype) for k, v in tensors.items()}, {k: v.astype(dtype) for k, v in msgs.items()})
    worst = 0.0
    for k, v in ref.items():
        got = host(out[dq.MessageID(k.src, k.dst)])
        worst = max(worst, np.linalg.norm(got - v) / np.linalg.norm(v))
        if k.src == 0:  # the centre's outputs: Hermitian, unit trace
            assert abs(np.trace(got) - 1.0) < 10 * tol
            assert np.abs(got - got.conj().T).max() < 10 * tol
    assert worst < tol, worst


def test_degree6_chi16_vidal_gauge_matches_oracle(star_case):
    """BP on the star converges in two synchronous sweeps (a tree of depth
    one); the Vidal gauge of the converged messages (edge decompositions and
    the absorb into the 33.5 M-element centre) against the oracle: lambdas to
    rounding and the symmetric-gauge one-site density matrices."""
    gp, tensors, msgs, _ = star_case
    go = _star()[1]
    trav_p, trav_o = list(gp.get_traversal_iterator()), go.traversal()
    bp_o = om.belief_propagation_map(trav_o, True)
    mo = to_oracle_messages(msgs)
    for _ in range(2):
        mo = bp_o(tensors, mo)
    gam_o, lam_o = om.vidal_gauge_fixing_map(trav_o, 1e-10)(tensors, mo)
    mp = {dq.MessageID(k.src, k.dst): v for k, v in mo.items()}
    gam_p, lam_p = dq.get_vidal_gauge_fixing_map(trav_p, 1e-10)(tensors, mp)
    for eid, lm in lam_o.items():
        assert np.abs(host(lam_p[eid]) - lm).max() < 1e-10, eid
    gam_h = {nid: host(gam_p[nid]) for nid in gam_o}
    lam_h = {eid: host(lam_p[eid]) for eid in lam_o}
    sg = om.symmetric_gauge_fixing_map(trav_o, 1e-10)
    t_o, t_p = sg(gam_o, lam_o), sg(gam_h, lam_h)
    m_o, m_p = om.lambdas2messages(lam_o), om.lambdas2messages(lam_h)
    for node in go.nodes.values():
        r_o = om.one_side_density_matrix(node, t_o, m_o)
        r_p = om.one_side_density_matrix(node, t_p, m_p)
        assert np.abs(r_o - r_p).max() < 1e-9, node.id


@pytest.mark.parametrize("dtype,tol", [(np.complex128, 1e-9), (np.complex64, 2e-5)])
def test_degree6_chi16_distance_and_density_matrices(star_case, dtype, tol):
    """The other per-node operations at the 33.5 M-element centre against the
    oracle: the Vidal distance (streaming bond Grams with diag(lambda^2)
    messages) of arbitrary tensors and lambdas, where every bond term is far
    from zero, and the one- and two-site density matrices from the BP
    messages (tiled environments)."""
    gp, tensors, msgs, ref = star_case
    go = _star()[1]
    trav_p, trav_o = list(gp.get_traversal_iterator()), go.traversal()
    rng = np.random.default_rng(5)
    lam = {}
    for eid in go.edges:
        v = np.sort(rng.uniform(0.05, 1.0, 16))[::-1]
        lam[eid] = v / np.linalg.norm(v)
    ten = {k: v.astype(dtype) for k, v in tensors.items()}
    d_o = om.vidal_gauge_distance_map(trav_o, 1e-10)(tensors, lam)
    d_p = float(dq.get_vidal_gauge_distance_map(trav_p, 1e-10)(ten, lam))
    assert d_o > 0.1, d_o
    assert abs(d_p - d_o) < tol * d_o, (d_p, d_o)
    out = dq.get_belief_propagation_map(trav_p, True)(ten, {k: v.astype(dtype) for k, v in msgs.items()})
    for nid in (0, 3):
        r_p = host(dq.get_one_side_density_matrix(gp.nodes[nid], ten, out))
        r_o = om.one_side_density_matrix(go.nodes[nid], tensors, ref)
        assert np.abs(r_p - r_o).max() < tol, nid
    for eid in ((0, 1), (0, 6)):
        r_p = host(dq.get_two_side_density_matrix(gp.edges[eid], ten, out))
        r_o = om.two_side_density_matrix(go.edges[eid], tensors, ref)
        assert np.abs(r_p - r_o).max() < tol, eid


def _uniform_pair(name):
    """Graphs whose nodes all have one degree >= 4 at chi = 16: the streaming
    node update fuses each output slot's last product with its Gram."""
    gp, go = dq.TensorGraph(), og.TensorGraph()
    if name == "torus4x4":  # degree 4, periodic 4 x 4 lattice
        return dq.get_nd_lattice([4, 4], 2, 16, False), og.nd_lattice([4, 4], 2, 16, False)
    for g in (gp, go):  # K7: degree 6, 33.5 M elements per node
        for _ in range(7):
            g.add_node(2)
        for a in range(7):
            for b in range(a + 1, 7):
                g.add_edge((a, b), 16)
    return gp, go


@pytest.mark.parametrize("name", ["torus4x4", "k7"])
@pytest.mark.parametrize("dtype,tol", [(np.complex128, 1e-11), (np.complex64, 2e-5)])
def test_uniform_degree_fused_node_update_matches_oracle(name, dtype, tol):
    """One synchronous sweep on graphs of uniform degree 4 (every (product
    axis, output slot) pair of the degree-4 plan, the contiguous and the
    strided tile layouts) and 6 against the oracle."""
    gp, go = _uniform_pair(name)
    rng = np.random.default_rng(11)
    tensors, msgs = {}, {}
    for n in go.nodes.values():
        shape = n.bond_shape + (2,)
        scale = np.sqrt(2.0 * np.prod(shape))
        tensors[n.id] = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)) / scale
    for el in gp.get_traversal_iterator():
        for nb in el.neighbors:
            a = rng.standard_normal((16, 16)) + 1j * rng.standard_normal((16, 16))
            m = a @ a.conj().T
            msgs[dq.MessageID(el.id, nb.id)] = m / np.trace(m)
    ref = om.belief_propagation_map(go.traversal(), True)(tensors, to_oracle_messages(msgs))
    bp = dq.get_belief_propagation_map(list(gp.get_traversal_iterator()), True)
    out = bp({k: v.astype(dtype) for k, v in tensors.items()}, {k: v.astype(dtype) for k, v in msgs.items()})
    worst = 0.0
    for k, v in ref.items():
        got = host(out[dq.MessageID(k.src, k.dst)])
        worst = max(worst, np.linalg.norm(got - v) / np.linalg.norm(v))
    assert worst < tol, worst
