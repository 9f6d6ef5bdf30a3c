"""Workload drivers on the GPU engine.

* 3D Ising square-root network (examples/3d_ising_canonicalization.py):
  on an open chain the network is a tree, so BP is exact and the mean
  magnetisation equals the brute-force Boltzmann average; on a 3x3x3 and the
  full 6x6x6 open cube the driver is compared with the oracle flow.
* Eagle kicked Ising (examples/ibm_eagle_processor.py): <Z_62> dynamics
  against the oracle emulator on the 127-qubit lattice, in the reference's
  edge order and in the three-colour-layer mode.
* The batched annealing instance set: sizes and one batched step.
"""
import itertools

import numpy as np
import pytest

from tests.conftest import gpu_available

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not gpu_available(), reason="needs a ROCm GPU")]

import dequantizer_amd as dq  # noqa: E402
from dequantizer_amd import drivers  # noqa: E402
from oracle import emulator as oe  # noqa: E402
from oracle import graph as og  # noqa: E402
from oracle import mappings as om  # noqa: E402


def _chain_magnetisation(n, beta, coupling, field):
    """<s> of the open Ising chain with weight exp(-beta (coupling sum s_i s_{i+1}
    + field sum s_i)), spin index 0 -> +1, by enumeration."""
    num = den = 0.0
    for cfg in itertools.product([1.0, -1.0], repeat=n):
        s = np.asarray(cfg)
        w = np.exp(-beta * (coupling * np.sum(s[:-1] * s[1:]) + field * np.sum(s)))
        num += w * s.mean()
        den += w
    return num / den


@pytest.mark.parametrize("beta", [0.1, 0.4, 0.6])
def test_ising_chain_magnetisation_is_exact(beta):
    res = drivers.process_ising_sqrt_partition_network(1, 8, [beta], -1.0, 0.5, 1e-10, True, criterion="residual",
                                                       max_iterations=500)
    assert res.number_of_iterations[0] < 500
    assert abs(res.average_magnetization[0] - _chain_magnetisation(8, beta, -1.0, 0.5)) < 1e-9


def _oracle_ising(size, dim, beta, coupling, field, acc, key):
    lattice = og.nd_lattice(dim * [size], 2, 2, True)
    trav = lattice.traversal(ordering="dfs")
    bp = om.belief_propagation_map(trav, False)
    spin = np.array([1.0, -1.0])
    tensors = lattice.init_tensors(og.potts_initializer(np.exp(-coupling * beta * np.outer(spin, spin)),
                                                        np.exp(-field * beta * spin)))
    messages = lattice.init_messages(og.message_random_nonnegative_initializer(key))
    dist, it = np.inf, 0
    while dist > acc and it < 500:
        new = bp(tensors, messages)
        dist = om.messages_frob_distance(new, messages)
        messages = new
        it += 1
    mag = np.mean([2 * om.one_side_density_matrix(n, tensors, messages)[0, 0].real - 1 for n in lattice.nodes.values()])
    return mag, it


@pytest.mark.parametrize("size,betas", [(3, [0.2, 0.5]), (6, [0.1, 0.2, 0.3, 0.4, 0.5, 0.6])])
def test_ising_cube_matches_oracle(size, betas):
    """The 6x6x6 case is the full C1 workload (residual 1e-10)."""
    res = drivers.process_ising_sqrt_partition_network(3, size, betas, -1.0, 0.5, 1e-10, True, criterion="residual",
                                                       max_iterations=500)
    for k, beta in enumerate(betas[:2]):
        mag, it = _oracle_ising(size, 3, beta, -1.0, 0.5, 1e-10, og.make_key(42))
        assert res.number_of_iterations[k] == it
        assert abs(res.average_magnetization[k] - mag) < 1e-9
    assert all(abs(m) <= 1.0 for m in res.average_magnetization)
    # the Vidal gauge of the converged messages is canonical
    vd = dq.get_vidal_gauge_distance_map(
        list(dq.get_nd_lattice([size] * 3, 2, 2, True).get_traversal_iterator(ordering="dfs")), 1e-10)
    assert float(vd(res.gammas[-1], res.lambdas[-1])) < 1e-7


def _oracle_eagle(chi, steps, theta, colour_layers):
    lattice = og.heavy_hex_eagle()
    trav = lattice.traversal()
    per = 10 ** 9 if colour_layers else lattice.edges_number
    emu = oe.BPQuantumEmulator(lattice, og.make_key(42), chi, per, 1e-5, None, False, "dfs")
    edges = [el.id for el in trav if isinstance(el, og.Edge)]
    classes = []
    if colour_layers:
        colours = drivers.edge_colouring(dq.get_heavy_hex_ibm_eagle_lattice())
        classes = [[e for e in edges if colours[e] == c] for c in range(3)]
    mixing = drivers.rx(theta)
    out = [float((lambda d: (d[0, 0] - d[1, 1]).real)(emu.dens_q1(62)))]
    for _ in range(steps):
        for n in range(lattice.nodes_number):
            emu.apply_q1(mixing, n)
        if colour_layers:
            for cls in classes:
                for a, b in cls:
                    emu.apply_q2(drivers.RZZ, a, b)
                emu._vidal_regauging()
                emu.gates_since = 0
        else:
            for a, b in edges:
                emu.apply_q2(drivers.RZZ, a, b)
        d = emu.dens_q1(62)
        out.append(float((d[0, 0] - d[1, 1]).real))
    return out


@pytest.mark.parametrize("colour_layers", [False, True])
def test_eagle_dynamics_matches_oracle(colour_layers):
    z = drivers.run_eagle_simulation(8, 1, 3, 1e-5, 0.8, is_infinite=False, colour_layers=colour_layers)
    z_o = _oracle_eagle(8, 3, 0.8, colour_layers)
    assert np.abs(np.asarray(z) - np.asarray(z_o)).max() < 1e-8, (z, z_o)


def _eagle_fixture():
    import json
    import os
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "eagle_chi32.json")
    with open(path) as f:
        return json.load(f)


@pytest.mark.parametrize("mode", ["edge_order", "colour_layers"])
def test_eagle_chi32_complex128_matches_oracle(mode):
    """The 127-qubit kicked-Ising run at max_chi = 32 in complex128 against
    the oracle emulator's <Z_62> trajectory (tests/golden/eagle_chi32.json,
    written by tests/golden/make_eagle_chi32.py): the bonds reach 32 at step 5
    and from step 6 on every simple update truncates a 64-value spectrum to
    32 (the chi * phys = 64 storage path), in the reference's edge order and
    in three colour layers with a regauge after each."""
    import torch
    fx = _eagle_fixture()
    want = fx[mode]["z62"]
    steps = len(want) - 1
    print(f"eagle chi=32 c128 {mode}, {steps} steps", flush=True)
    z = drivers.run_eagle_simulation(32, 1, steps, fx["accuracy"], fx["theta"], is_infinite=False,
                                     dtype=torch.complex128, colour_layers=(mode == "colour_layers"))
    err = np.abs(np.asarray(z) - np.asarray(want))
    print("max |dZ| per step", err, flush=True)
    assert err.max() < 1e-6, (z, want)


def test_eagle_chi32_complex64_vs_complex128_all_steps():
    """The C2 configuration (127 qubits, chi=32, complex64 bulk, three colour
    layers each followed by a regauge, 20 steps): <Z_62> within 1e-5 of the
    complex128 run at every step."""
    import torch
    runs = {}
    for dtype in (torch.complex64, torch.complex128):
        print(f"eagle chi=32 {dtype}, 20 steps", flush=True)
        runs[dtype] = np.asarray(drivers.run_eagle_simulation(32, 1, 20, 1e-5, 0.8, is_infinite=False, dtype=dtype,
                                                              colour_layers=True))
    err = np.abs(runs[torch.complex64] - runs[torch.complex128])
    print("max |dZ| c64 vs c128:", float(err.max()), flush=True)
    assert err.max() < 1e-5, err


def test_annealing_instance_set_batched_step():
    import torch
    from dequantizer_amd.annealing import Annealer, get_scheduler, linear_schedule
    efs = drivers.annealing_instances(64)
    assert len(efs) == 64 and all(100 <= ef.nodes_number <= 1000 for ef in efs)
    assert all(ef.nodes_number % 2 == 0 and ef.edges_number == 3 * ef.nodes_number // 2 for ef in efs)
    keys = [dq.PRNGKey(k) for k in range(64)]
    ann = Annealer(efs, keys, 16, 1, 1e-5, 100, True, "bfs", dtype=torch.complex64, chi=16, fixed_sweeps=4)
    for mix, cpl in get_scheduler(0.1, linear_schedule, 200)[:3]:
        ann.step(mix, cpl)
    res = ann.finish(two_site=True)
    assert len(res) == 64
    for r, ef in zip(res, efs):
        assert r.density_matrices.shape == (ef.nodes_number, 2, 2)
        assert np.isfinite(r.bp_energy)


def test_eagle_batched_angles_match_single_runs():
    """Several RX angles as instances of one engine give the single-angle
    dynamics (synchronous BP, chi=16, complex128)."""
    import torch
    thetas = [0.3, 0.8]
    zs = drivers.run_eagle_simulations(thetas, 16, 3, 1e-5, is_infinite=False, dtype=torch.complex128,
                                       synchronous_update=True)
    for th, z in zip(thetas, zs):
        z1 = drivers.run_eagle_simulation(16, 1, 3, 1e-5, th, is_infinite=False, synchronous_update=True,
                                          max_belief_propagation_iterations=100, dtype=torch.complex128,
                                          colour_layers=True)
        assert np.abs(np.asarray(z) - np.asarray(z1)).max() < 1e-9, (th, z, z1)
