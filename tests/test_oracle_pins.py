"""Pins of the oracle's Vidal distance and of the simple update's entropy,
truncation error and kept spectrum (CPU, no GPU).

* ``vidal_dist`` (dequantizer/src/tensor_ops.py:143-175) and its map
  (dequantizer/src/mappings/vidal_distance.py:17-42) away from zero: node
  tensors whose lambda-weighted bond Gram R is known in closed form, so the
  distance is the trace norm of R - ideal computed by hand (a diagonal R gives
  sum |p_k - 1/count|; a 2x2 R = [[a, b], [b*, 1 - a]] gives
  2 sqrt((a - 1/2)^2 + |b|^2)).
* ``simple_update`` (dequantizer/src/tensor_ops.py:318-389): on a four-qubit
  chain in exact Vidal form (built by sequential SVDs of a state vector, not
  by the oracle), a gate on the middle bond followed by the update must give
  the Schmidt spectrum of the new state across that bond (brute-force SVD),
  its entanglement entropy -sum s^2 log s^2 and the discarded weight
  sqrt(sum_{k >= rank} s_k^2), with and without a rank cap.

Each check also catches a planted mistake in the oracle.
"""
import numpy as np
import pytest

from oracle import graph as og
from oracle import mappings as om
from oracle import tensor_ops as ot

EPS = 1e-10


# ------------------------------------------------------------ vidal_dist
def _degree1_tensor(r: np.ndarray) -> np.ndarray:
    """T[b, s] with sum_s conj(T[k, s]) T[l, s] = r[k, l] (r Hermitian PSD):
    T = conj(sqrt(r)), since conj(T) T^T = sqrt(r) sqrt(r)^H = r."""
    w, v = np.linalg.eigh(r)
    root = (v * np.sqrt(np.clip(w, 0, None))) @ v.conj().T
    return np.conj(root)


def test_vidal_dist_diagonal_closed_form():
    """R = diag(0.7, 0.3): ideal I/2, distance |0.7-0.5| + |0.3-0.5| = 0.4."""
    t = _degree1_tensor(np.diag([0.7, 0.3]).astype(complex))
    assert ot.vidal_dist(t, [np.array([0.9, 0.1])], EPS) == pytest.approx(0.4, abs=1e-13)
    # a degree-1 node has no other bonds: its own lambda does not enter
    assert ot.vidal_dist(t, [np.array([0.5, 0.5])], EPS) == pytest.approx(0.4, abs=1e-13)


def test_vidal_dist_off_diagonal_closed_form():
    """R = [[a, b], [b*, 1-a]]: R - I/2 has eigenvalues +-sqrt((a-1/2)^2 + |b|^2)."""
    a, b = 0.62, 0.21 - 0.13j
    r = np.array([[a, b], [np.conj(b), 1 - a]])
    t = 1.7 * _degree1_tensor(r)  # the trace normalisation removes the scale
    want = 2 * np.sqrt((a - 0.5) ** 2 + abs(b) ** 2)
    assert ot.vidal_dist(t, [np.array([1.0, 0.0])], EPS) == pytest.approx(want, abs=1e-12)


def test_vidal_dist_ideal_counts_only_nonzero_diagonal():
    """R = diag(0.6, 0.4, 0): ideal = diag(1, 1, 0)/2, distance 0.2."""
    t = _degree1_tensor(np.diag([0.6, 0.4, 0.0]).astype(complex))
    assert ot.vidal_dist(t, [np.ones(3)], EPS) == pytest.approx(0.2, abs=1e-12)


def _weighted_node():
    """Degree-2 node T[a, c, s] = delta_{a c} delta_{s 0} x_a: the Gram of bond
    0 weighted by lambda_1^2 is diag(|x_a|^2 l1_a^2) and that of bond 1
    weighted by lambda_0^2 is diag(|x_c|^2 l0_c^2)."""
    x = np.array([0.8, 0.5 + 0.2j, 0.3j])
    t = np.zeros((3, 3, 2), dtype=complex)
    for a in range(3):
        t[a, a, 0] = x[a]
    l0 = np.array([0.9, 0.35, 0.25])
    l1 = np.array([0.7, 0.6, 0.38])

    def dist_of(p):
        p = p / p.sum()
        nz = p > EPS
        return float(np.sum(np.abs(p - nz / nz.sum())))

    want = dist_of(np.abs(x) ** 2 * l1**2) + dist_of(np.abs(x) ** 2 * l0**2)
    return t, [l0, l1], want


def test_vidal_dist_lambda_squared_weights_of_the_other_bonds():
    t, lams, want = _weighted_node()
    assert ot.vidal_dist(t, lams, EPS) == pytest.approx(want, abs=1e-12)


def test_vidal_distance_map_divides_by_twice_the_edges():
    """Two nodes on one edge: (d_a + d_b) / (2 |E|)."""
    g = og.TensorGraph()
    g.add_node(2)
    g.add_node(2)
    g.add_edge((0, 1), 2)
    a, b = 0.62, 0.21 - 0.13j
    tensors = {0: _degree1_tensor(np.diag([0.7, 0.3]).astype(complex)),
               1: _degree1_tensor(np.array([[a, b], [np.conj(b), 1 - a]]))}
    lambdas = {(0, 1): np.array([0.8, 0.6])}
    want = (0.4 + 2 * np.sqrt((a - 0.5) ** 2 + abs(b) ** 2)) / 2
    assert om.vidal_gauge_distance_map(g.traversal(), EPS)(tensors, lambdas) == pytest.approx(want, abs=1e-12)


def test_vidal_dist_planted_mistakes_are_caught():
    """Frobenius instead of trace norm, an ideal divided by the full bond
    dimension, and lambda instead of lambda^2 weights: each changes a value
    pinned above."""
    t1 = _degree1_tensor(np.diag([0.7, 0.3]).astype(complex))
    t3 = _degree1_tensor(np.diag([0.6, 0.4, 0.0]).astype(complex))
    tw, lams, want_w = _weighted_node()
    orig_svd = np.linalg.svd

    class _FrobLinalg:
        def __getattr__(self, name):
            return getattr(np.linalg, name)

        @staticmethod
        def svd(a, compute_uv=True, **kw):
            s = orig_svd(a, compute_uv=compute_uv, **kw)
            if compute_uv:
                return s
            return np.array([np.sqrt(np.sum(s**2))])  # Frobenius norm as a one-term "spectrum"

    orig_np = ot.np

    class _Np:
        linalg = _FrobLinalg()

        def __getattr__(self, name):
            return getattr(orig_np, name)

    try:
        ot.np = _Np()
        assert abs(ot.vidal_dist(t1, [np.ones(2)], EPS) - 0.4) > 1e-3
    finally:
        ot.np = orig_np

    orig_gt = ot._complex_gt
    try:  # every diagonal entry counted in the ideal
        ot._complex_gt = lambda z, eps: np.ones(z.shape, dtype=bool)
        assert abs(ot.vidal_dist(t3, [np.ones(3)], EPS) - 0.2) > 1e-3
    finally:
        ot._complex_gt = orig_gt

    orig_fn = ot.vidal_dist

    def lam_not_squared(tensor, lambdas, eps):
        return orig_fn(tensor, [np.sqrt(np.abs(l)) for l in lambdas], eps)

    assert abs(lam_not_squared(tw, lams, EPS) - want_w) > 1e-3


# ----------------------------------------------------- simple update spectra
def _vidal_chain(psi: np.ndarray):
    """Exact Vidal form of a 4-qubit state by sequential SVDs.  Returns the
    node tensors in the graph layout (bonds in edge-insertion order, phys
    last) and the normalised lambdas of the bonds (0,1), (1,2), (2,3)."""
    psi = psi / np.linalg.norm(psi)
    u, s1, vh = np.linalg.svd(psi.reshape(2, 8), full_matrices=False)
    g0 = u.T  # (bond01, s0)
    rest = (s1[:, None] * vh).reshape(len(s1) * 2, 4)
    u2, s2, vh2 = np.linalg.svd(rest, full_matrices=False)
    g1 = u2.reshape(len(s1), 2, len(s2)) / s1[:, None, None]  # (b01, s1, b12)
    rest2 = (s2[:, None] * vh2).reshape(len(s2) * 2, 2)
    u3, s3, vh3 = np.linalg.svd(rest2, full_matrices=False)
    g2 = u3.reshape(len(s2), 2, len(s3)) / s2[:, None, None]  # (b12, s2, b23)
    g3 = vh3  # (b23, s3)
    tensors = {0: g0, 1: np.transpose(g1, (0, 2, 1)), 2: np.transpose(g2, (0, 2, 1)), 3: g3}
    return tensors, [s1, s2, s3]


def _chain_state(t, lams):
    """Contract the Vidal chain back to the state vector (axes s0..s3)."""
    l1, l2, l3 = lams
    return np.einsum("as,a,abt,b,bcu,c,cv->stuv", t[0], l1, t[1], l2, t[2], l3, t[3])


def _schmidt(psi):
    sv = np.linalg.svd(psi.reshape(4, 4), compute_uv=False)
    return sv / np.linalg.norm(sv)


def _su_middle(seed: int, max_rank):
    rng = np.random.default_rng(seed)
    psi = rng.standard_normal((2, 2, 2, 2)) + 1j * rng.standard_normal((2, 2, 2, 2))
    t, lams = _vidal_chain(psi)
    assert np.allclose(_chain_state(t, lams), psi / np.linalg.norm(psi), atol=1e-12)
    g = rng.standard_normal((4, 4)) + 1j * rng.standard_normal((4, 4))
    gate = np.linalg.qr(g)[0].reshape(2, 2, 2, 2)  # (o1, o2, i1, i2)
    new = np.einsum("abij,sijv->sabv", gate, _chain_state(t, lams))
    h1, h2 = ot.decompose_gate(gate, EPS)
    out = ot.simple_update(t[1], t[2], [lams[0], lams[1]], [lams[1], lams[2]], h1, h2, 1, 0, EPS, None,
                           max_rank)
    return new, out, t, lams


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_simple_update_spectrum_entropy_and_error_without_cap(seed):
    new, (g1, g2, s, err, ent), t, lams = _su_middle(seed, None)
    sigma = _schmidt(new)
    assert np.allclose(s / np.linalg.norm(s), sigma[: len(s)], atol=1e-12)
    assert err == pytest.approx(0.0, abs=1e-12)
    p = sigma**2
    assert ent == pytest.approx(float(-np.sum(p[p > 0] * np.log(p[p > 0]))), abs=1e-12)
    # the untruncated update is exact: the new Vidal chain is the new state
    lam_new = s / np.linalg.norm(s)
    rebuilt = np.einsum("as,a,abt,b,bcu,c,cv->stuv", t[0], lams[0], g1, lam_new, g2, lams[2], t[3])
    rebuilt = rebuilt / np.linalg.norm(rebuilt)
    phase = np.vdot(rebuilt, new)
    assert np.allclose(rebuilt * phase / abs(phase), new, atol=1e-10)


@pytest.mark.parametrize("seed,rank", [(0, 2), (1, 3), (2, 1)])
def test_simple_update_truncation_error_and_entropy_with_cap(seed, rank):
    new, (_, _, s, err, ent), _, _ = _su_middle(seed, rank)
    sigma = _schmidt(new)
    assert len(s) == rank
    assert err == pytest.approx(float(np.sqrt(np.sum(sigma[rank:] ** 2))), abs=1e-12)
    p = sigma**2
    assert ent == pytest.approx(float(-np.sum(p * np.log(p))), abs=1e-12)
    # kept values are the leading Schmidt values up to one common scale
    assert np.allclose(s / np.linalg.norm(s) * np.sqrt(1 - err**2), sigma[:rank], atol=1e-12)


def test_simple_update_planted_mistakes_are_caught():
    """Entropy of s instead of s^2, and the error from the unnormalised
    values: each moves a value pinned above."""
    new, (_, _, _, err, ent), _, _ = _su_middle(1, 2)
    sigma = _schmidt(new)
    want_err = float(np.sqrt(np.sum(sigma[2:] ** 2)))
    want_ent = float(-np.sum(sigma**2 * np.log(sigma**2)))
    orig_entr = ot.entr
    try:  # entr applied to s_hat (its square root) instead of s_hat^2
        ot.entr = lambda x: orig_entr(np.sqrt(x))
        _, (_, _, _, _, bad_ent), _, _ = _su_middle(1, 2)
        assert abs(bad_ent - want_ent) > 1e-3
    finally:
        ot.entr = orig_entr
    orig_norm = ot.np.linalg.norm
    calls = {"n": 0}

    def norm_skip_spectrum(x, *a, **k):
        # the first norm of a 1-D real spectrum inside simple_update is the
        # s_hat normalisation: return 1 once so the error comes out of the
        # unnormalised values
        if calls["n"] == 0 and isinstance(x, np.ndarray) and x.ndim == 1 and np.isrealobj(x) and not a and not k:
            calls["n"] += 1
            return 1.0
        return orig_norm(x, *a, **k)

    try:
        ot.np.linalg.norm = norm_skip_spectrum
        _, (_, _, _, bad_err, _), _, _ = _su_middle(1, 2)
    finally:
        ot.np.linalg.norm = orig_norm
    assert calls["n"] > 0
    assert abs(bad_err - want_err) > 1e-3
    assert err == pytest.approx(want_err, abs=1e-12) and ent == pytest.approx(want_ent, abs=1e-12)


# ------------------------------------------------------- golden fixtures
def _load_script(name):
    import importlib.util
    import os
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", name)
    spec = importlib.util.spec_from_file_location(name[:-3], path)
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_eagle_fixture_matches_the_oracle_on_its_first_steps():
    """tests/golden/eagle_chi32.json holds the oracle's <Z_62> trajectory; its
    first four steps (bonds up to 16, about a second of oracle work) are
    recomputed here."""
    import json
    import os
    mod = _load_script("make_eagle_chi32.py")
    with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "eagle_chi32.json")) as f:
        fx = json.load(f)
    for mode, colour in (("edge_order", False), ("colour_layers", True)):
        z, _ = mod.run(4, colour)
        assert np.abs(np.asarray(z) - np.asarray(fx[mode]["z62"][:5])).max() < 1e-12, mode
