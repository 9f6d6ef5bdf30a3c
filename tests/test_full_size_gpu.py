"""Accuracy at the full sizes of the annealing and lattice workloads.

* The 1000-spin random 3-regular annealing run at chi = 16 (the instance of
  examples/quantum_annealing/quantum_annealing/src/quantum_annealer.py:90-152
  at seed 7): the first 80 steps (truncating from about step 45) in complex128
  and complex64 against the oracle's values
  stored in tests/golden/rr3_n1000_chi16.npz (written by
  tests/golden/make_rr3_n1000.py), and the whole 200-step run in complex64
  against complex128 (the complex128 path is the one held to the oracle).
* The 64-instance set (N from 100 to 1000): per-instance energy per spin,
  complex64 against complex128, over the whole 200-step run.
* The periodic 4^3 lattice (every node of degree 6, chi = 16): BP converged
  on the GPU in complex128 is a fixed point of the oracle's node update
  (dequantizer/src/tensor_ops.py:36-60) on sampled nodes, and the GPU's
  lambdas equal the oracle's edge decomposition (tensor_ops.py:108-130) of
  the converged messages.

The accuracy target for magnetisation and energy is 1e-5.
"""
import os

import numpy as np
import pytest

from tests.conftest import gpu_available

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not gpu_available(), reason="needs a ROCm GPU")]

import torch  # noqa: E402

import dequantizer_amd as dq  # noqa: E402
from dequantizer_amd import PRNGKey  # noqa: E402
from dequantizer_amd.annealing import (Annealer, EnergyFunction, edge_colored, get_scheduler,  # noqa: E402
                                       linear_schedule, random_regular)

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(autouse=True)
def _release_cached_memory():
    """These tests hold tens of GB (the 4^3 lattice: 32 GiB of complex128
    node tensors); blocks cached by earlier tests are returned first."""
    import gc
    gc.collect()
    torch.cuda.empty_cache()
    yield
    gc.collect()
    torch.cuda.empty_cache()


def _z_zz(res):
    r1, r2 = res.density_matrices, res.two_site_density_matrices
    z = (r1[:, 0, 0] - r1[:, 1, 1]).real
    zz = (r2[:, 0, 0] - r2[:, 1, 1] - r2[:, 2, 2] + r2[:, 3, 3]).real
    return z, zz


def _run(efs, keys, steps, dtype, schedule_steps=200):
    ann = Annealer(efs, keys, 16, 1, 1e-5, 100, True, "bfs", dtype=dtype, chi=16, fixed_sweeps=10)
    for mix, cpl in get_scheduler(0.1, linear_schedule, schedule_steps)[:steps]:
        ann.step(mix, cpl)
    return ann.finish(two_site=True)


@pytest.mark.parametrize("dtype,tol", [(torch.complex128, 1e-9), (torch.complex64, 1e-5)])
def test_rr3_n1000_first_steps_match_oracle(dtype, tol):
    fx = np.load(os.path.join(GOLDEN, "rr3_n1000_chi16.npz"))
    ef = EnergyFunction(fx["couplings"], fx["pairs"], fx["fields"])
    steps = int(fx["steps"])
    print(f"1000 spins, chi=16, {steps} steps, {dtype}", flush=True)
    res = _run([ef], [PRNGKey(7)], steps, dtype)[0]
    z, zz = _z_zz(res)
    dz, dzz = np.abs(z - fx["z"]).max(), np.abs(zz - fx["zz"]).max()
    de = abs(res.bp_energy - float(fx["energy"])) / ef.nodes_number
    print(f"max |dZ| {dz:.2e}, max |dZZ| {dzz:.2e}, |dE|/N {de:.2e}", flush=True)
    assert dz < tol and dzz < tol and de < tol
    assert len(res.truncation_errors) == steps * ef.edges_number
    tail = fx["truncation_errors_tail"]
    assert np.abs(res.truncation_errors[-len(tail):] - tail).max() < max(tol, 1e-9) * 10
    assert tail.max() > 1e-6  # the run truncates
    vd = np.asarray(res.vidal_distances_after_regauging, dtype=np.float64)
    assert np.abs(vd - fx["vidal_distances"]).max() < max(tol, 1e-9) * 10


def test_rr3_n1000_whole_run_complex64_vs_complex128():
    """All 200 steps plus the final observables, complex64 against complex128:
    the energy per spin and the magnetisation (mean <Z>) within the 1e-5
    target.  Single spins and edges drift further over the 200 steps (the
    largest per-spin |d<Z>| measured on MI355X is 1.6e-5, per-edge |d<ZZ>|
    8.9e-6): they are held to 3e-5."""
    ef = edge_colored(random_regular(1000, 3, seed=7, couplings="pm1", field_std=0.1))
    out = {}
    for dtype in (torch.complex64, torch.complex128):
        print(f"1000 spins, chi=16, 200 steps, {dtype}", flush=True)
        out[dtype] = _run([ef], [PRNGKey(7)], 200, dtype)[0]
    (z64, zz64), (z128, zz128) = _z_zz(out[torch.complex64]), _z_zz(out[torch.complex128])
    dz, dzz = np.abs(z64 - z128).max(), np.abs(zz64 - zz128).max()
    de = abs(out[torch.complex64].bp_energy - out[torch.complex128].bp_energy) / ef.nodes_number
    dm = abs(float(z64.mean() - z128.mean()))
    print(f"max |dZ| {dz:.2e}, max |dZZ| {dzz:.2e}, |dE|/N {de:.2e}, |dm| {dm:.2e}, "
          f"E/N {out[torch.complex128].bp_energy / 1000:.6f}", flush=True)
    assert de < 1e-5 and dm < 1e-5
    assert dz < 3e-5 and dzz < 3e-5


def test_instance_set_whole_run_complex64_vs_complex128():
    """The 64 random 3-regular instances (37,356 spins) through the whole
    200-step run in one engine, complex64 against complex128: per-instance
    energy per spin and magnetisation within 1e-5; single spins drift further (2.9e-5
    measured on MI355X) and are held to 5e-5."""
    from dequantizer_amd.drivers import annealing_instances
    efs = [edge_colored(ef) for ef in annealing_instances(64)]
    keys = [PRNGKey(1000 + k) for k in range(len(efs))]
    out = {}
    for dtype in (torch.complex64, torch.complex128):
        print(f"64 instances, chi=16, 200 steps, {dtype}", flush=True)
        out[dtype] = _run(efs, keys, 200, dtype)
    worst_e = worst_z = worst_m = 0.0
    for r64, r128, ef in zip(out[torch.complex64], out[torch.complex128], efs):
        z64, z128 = _z_zz(r64)[0], _z_zz(r128)[0]
        worst_e = max(worst_e, abs(r64.bp_energy - r128.bp_energy) / ef.nodes_number)
        worst_m = max(worst_m, abs(float(z64.mean() - z128.mean())))
        worst_z = max(worst_z, float(np.abs(z64 - z128).max()))
    print(f"64 instances: max |dE|/N {worst_e:.2e}, max |dm| {worst_m:.2e}, max |dZ| {worst_z:.2e}", flush=True)
    assert worst_e < 1e-5 and worst_m < 1e-5
    assert worst_z < 5e-5


def test_periodic_lattice_l4_is_an_oracle_fixed_point():
    """Periodic 4^3 lattice, degree 6, chi = 16, phys 2, Gaussian node tensors
    (x + iy)/sqrt(2 chi^6): synchronous BP on the GPU in complex128 to a
    residual of 1e-8, then the Vidal gauge.  On the host the oracle's node
    update of two nodes, fed the GPU's converged incoming messages, gives back
    the GPU's outgoing messages, and the oracle's edge decomposition of two
    edges' converged messages gives the GPU's lambdas."""
    from oracle import tensor_ops as ot
    chi = 16
    lat = dq.get_nd_lattice([4, 4, 4], 2, chi, False)
    trav = list(lat.get_traversal_iterator())
    gen = torch.Generator(device="cuda").manual_seed(5)
    tensors = {}
    for el in trav:
        if isinstance(el, dq.Node):
            x = torch.randn((*el.bond_shape, 2), dtype=torch.float64, device="cuda", generator=gen)
            y = torch.randn((*el.bond_shape, 2), dtype=torch.float64, device="cuda", generator=gen)
            tensors[el.id] = torch.complex(x, y) / np.sqrt(2.0 * chi ** 6)
            del x, y
    rng = np.random.default_rng(5)
    msgs = {}
    for el in trav:
        for nb in el.neighbors:
            a = rng.standard_normal((chi, chi)) + 1j * rng.standard_normal((chi, chi))
            m = a @ a.conj().T
            msgs[dq.MessageID(el.id, nb.id)] = m / np.trace(m)
    bp = dq.get_belief_propagation_map(trav, True)
    m = msgs
    r = np.inf
    for sweep in range(60):
        new = bp(tensors, m)
        r = float(dq.messages_frob_distance(new, m))
        m = new
        if r < 1e-8:
            break
    print(f"L=4 lattice: residual {r:.2e} after {sweep + 1} sweeps", flush=True)
    assert r < 1e-8
    gam, lam = dq.get_vidal_gauge_fixing_map(trav, 1e-10)(tensors, m)
    host_m = {k: v.cpu().numpy() for k, v in m.items()}
    nodes = [el for el in trav if isinstance(el, dq.Node)]
    for node in (nodes[0], nodes[37]):
        t = tensors[node.id].cpu().numpy()
        inc = [host_m[dq.MessageID(e.id, node.id)] for e in node.neighbors]
        out = ot.pass_through_node(t, inc)
        del t
        for e, o in zip(node.neighbors, out):
            got = host_m[dq.MessageID(node.id, e.id)]
            rel = np.linalg.norm(got - o) / np.linalg.norm(o)
            assert rel < 1e-7, (node.id, e.id, rel)
    edges = [el for el in trav if isinstance(el, dq.Edge)]
    for e in (edges[0], edges[100]):
        inc = [host_m[dq.MessageID(nb.id, e.id)] for nb in e.neighbors]
        _, lam_o = ot.edge_svd(inc, 1e-10)
        assert np.abs(lam[e.id].cpu().numpy() - lam_o).max() < 1e-9, e.id
