"""GPU parity of the annealing driver (batched plan path) against the oracle.

Random 3-regular instances, the reference annealing flow (Hadamards, coupling
gates in pair order, mixing gates, regauge once per layer), compared with the
oracle annealer run instance by instance with the same keys.
"""
import numpy as np
import pytest
import torch

from tests.conftest import gpu_available

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not gpu_available(), reason="needs a ROCm GPU")]

from dequantizer_amd import PRNGKey  # noqa: E402
from dequantizer_amd.annealing import (Annealer, bp_energy, edge_colored, get_scheduler,  # noqa: E402
                                       linear_schedule, random_regular)
from oracle import annealing as oa  # noqa: E402
from oracle import graph as og  # noqa: E402


def _oracle_run(ef, key, steps, chi, sync, traversal, fixed):
    sched = oa.scheduler(0.1, lambda t: 1 - t, steps)
    emu_kw = {}
    emu, rho1, rho2 = oa.run_quantum_annealer(ef.coupled_spin_pairs, ef.coupling_amplitudes, ef.fields, sched,
                                              key, chi, ef.edges_number, 1e-5, 100, sync, traversal,
                                              two_site=True, **emu_kw) if fixed is None else _fixed(
        ef, key, sched, chi, sync, traversal, fixed)
    return emu, rho1, rho2


def _fixed(ef, key, sched, chi, sync, traversal, fixed):
    from oracle import emulator as oe
    g = og.graph_from_pairs(ef.nodes_number, ef.coupled_spin_pairs)
    emu = oe.BPQuantumEmulator(g, key, chi, ef.edges_number, 1e-5, 100, sync, traversal, fixed_sweeps=fixed)
    deg = [g.nodes[i].degree for i in range(ef.nodes_number)]
    for i in range(ef.nodes_number):
        emu.apply_q1(oa.HADAMARD, i)
    for mix, cpl in sched:
        for (a, b), j in zip(ef.coupled_spin_pairs, ef.coupling_amplitudes):
            a, b = int(a), int(b)
            emu.apply_q2(oa.coupling_gate(j, ef.fields[a], ef.fields[b], deg[a], deg[b], cpl), a, b)
        mg = oa.mixing_gate(mix)
        for i in range(ef.nodes_number):
            emu.apply_q1(mg, i)
    rho1 = np.array([emu.dens_q1(i) for i in range(ef.nodes_number)])
    rho2 = np.array([emu.dens_q2(int(a), int(b)) for a, b in ef.coupled_spin_pairs])
    return emu, rho1, rho2


@pytest.mark.parametrize("sync,traversal,fixed", [(False, "dfs", None), (True, "bfs", None), (False, "bfs", 3)])
@pytest.mark.parametrize("dtype,tol", [(torch.complex128, 1e-9), (torch.complex64, 1e-5)])
def test_batched_annealing_matches_oracle(sync, traversal, fixed, dtype, tol):
    efs = [random_regular(n, 3, seed=s) for n, s in ((12, 1), (16, 2), (10, 3))]
    keys = [PRNGKey(100 + k) for k in range(len(efs))]
    steps, chi = 6, 4
    ann = Annealer(efs, keys, chi, 1, 1e-5, 100, sync, traversal, dtype=dtype, chi=chi, fixed_sweeps=fixed)
    for mix, cpl in get_scheduler(0.1, linear_schedule, steps):
        ann.step(mix, cpl)
    res = ann.finish(two_site=True)
    for k, ef in enumerate(efs):
        emu_o, r1, r2 = _oracle_run(ef, keys[k], steps, chi, sync, traversal, fixed)
        assert np.abs(res[k].density_matrices - r1).max() < tol
        assert np.abs(res[k].two_site_density_matrices - r2).max() < tol
        e_o = oa.bp_energy(r1, r2, ef.fields, ef.coupling_amplitudes)
        assert abs(res[k].bp_energy - e_o) < tol * ef.nodes_number
        z_p = (res[k].density_matrices[:, 0, 0] - res[k].density_matrices[:, 1, 1]).real
        z_o = (r1[:, 0, 0] - r1[:, 1, 1]).real
        assert np.abs(z_p - z_o).max() < tol
        if dtype == torch.complex128:
            assert res[k].bp_iterations == emu_o.bp_iterations
            np.testing.assert_allclose(res[k].truncation_errors, emu_o.truncation_errors, atol=1e-9)


def test_edge_colored_order_matches_oracle():
    """The colour-sorted pair list is an ordinary pair order for the oracle."""
    efs = [edge_colored(random_regular(14, 3, seed=5))]
    keys = [PRNGKey(9)]
    ann = Annealer(efs, keys, 4, 1, 1e-5, 100, True, "bfs", dtype=torch.complex128, chi=4, fixed_sweeps=4)
    for mix, cpl in get_scheduler(0.1, linear_schedule, 5):
        ann.step(mix, cpl)
    res = ann.finish(two_site=True)[0]
    _, r1, r2 = _oracle_run(efs[0], keys[0], 5, 4, True, "bfs", 4)
    assert np.abs(res.density_matrices - r1).max() < 1e-9
    assert len(ann.plan.launches) <= 3 + 1  # three colour classes + the mixing level


@pytest.mark.parametrize("dtype,tol", [(torch.complex128, 1e-9), (torch.complex64, 2e-5)])
def test_chi16_matrix_core_paths_match_oracle(dtype, tol):
    """Padded chi=16 on a degree-3 graph: the simple update runs its FP64
    matrix-core environment / Gamma products (chi * phys = 32) and, in
    complex64, BP and the Vidal distance run on the MFMA node kernel."""
    ef = random_regular(10, 3, seed=4)
    key = PRNGKey(31)
    steps = 7
    ann = Annealer([ef], [key], 16, 1, 1e-5, 100, True, "bfs", dtype=dtype, chi=16, fixed_sweeps=6)
    for mix, cpl in get_scheduler(0.1, linear_schedule, steps):
        ann.step(mix, cpl)
    res = ann.finish(two_site=True)[0]
    emu_o, r1, r2 = _oracle_run(ef, key, steps, 16, True, "bfs", 6)
    assert np.abs(res.density_matrices - r1).max() < tol
    assert np.abs(res.two_site_density_matrices - r2).max() < tol
    np.testing.assert_allclose(res.truncation_errors, emu_o.truncation_errors, atol=max(tol, 1e-9))
    vd_o = np.asarray(emu_o.vidal_distances_after_regauging, dtype=np.float64)
    assert np.abs(np.asarray(res.vidal_distances_after_regauging, dtype=np.float64) - vd_o).max() < max(tol, 1e-9) * 10
