"""Configuration layer and instance generators of the annealing runs
(examples/quantum_annealing/qbp.py, exact.py, configs/, lattice_generator.py,
energy_function.py).  CPU only, except the two end-to-end runs at the bottom.

The instance structures are pinned by what the reference fixes: the Eagle
coupling map's degree histogram (127 spins, 144 couplings, degrees
{1: 2, 2: 89, 3: 36}, as the heavy-hex tensor graph of tensor_graph.py:547-613),
a single 12-cycle, a 10-spin tree with a degree-3 root, regular degrees and
the max-cut convention (zero fields, couplings -1).  Drawn values are NumPy
draws from this package's keys (JAX threefry streams are not reproduced):
parity unpinned, by construction.
"""
import collections
import math
import os

import numpy as np
import pytest

from dequantizer_amd import runner

from tests.conftest import gpu_available


def _write(path, text):
    os.makedirs(os.path.dirname(path), exist_ok=True)
    with open(path, "w") as f:
        f.write(text)


def _config_dir(root, nodes=12):
    _write(os.path.join(root, "config.yaml"),
           "defaults:\n  - quantum_annealing_schedule: short\n  - task_generator: tree\n"
           "  - emulator_parameters: std\nhydra:\n  run:\n    dir: ./outputs\n")
    _write(os.path.join(root, "quantum_annealing_schedule", "short.yaml"),
           "total_time_step_size: 0.2\nschedule: 'lambda t: 1 - t'\nsteps_number: 12\n"
           "sample_measurements: true\nrecord_history: true\n")
    _write(os.path.join(root, "task_generator", "tree.yaml"),
           "lattice_type: \"small_tree\"\nfields_std: 1.\ncouplings_std: 1.\nseed: 42\n")
    _write(os.path.join(root, "task_generator", "rr.yaml"),
           f"lattice_type: \"random_regular\"\nfields_std: 1.\ncouplings_std: 1.\nseed: 42\n"
           f"nodes_number: {nodes}\ndegree: 3\n")
    _write(os.path.join(root, "emulator_parameters", "std.yaml"),
           "max_chi: 4\nmax_belief_propagation_iterations: 100\nlayers_per_regauging: 1\naccuracy: 1e-5\n"
           "synchronous_update: False\ntraversal_type: \"dfs\"\n")
    return root


def test_load_config_defaults_and_overrides(tmp_path):
    root = _config_dir(str(tmp_path))
    cfg = runner.load_config(root)
    assert cfg["task_generator"]["lattice_type"] == "small_tree"
    assert cfg["emulator_parameters"]["accuracy"] == 1e-5 or cfg["emulator_parameters"]["accuracy"] == "1e-5"
    assert "hydra" not in cfg
    cfg = runner.load_config(root, overrides=["task_generator=rr", "emulator_parameters.max_chi=16",
                                              "quantum_annealing_schedule.steps_number=3"])
    assert cfg["task_generator"]["nodes_number"] == 12
    assert cfg["emulator_parameters"]["max_chi"] == 16
    assert cfg["quantum_annealing_schedule"]["steps_number"] == 3
    with pytest.raises(FileNotFoundError):
        runner.load_config(root, overrides=["task_generator=missing"])


def test_parse_schedule_without_eval():
    f = runner.parse_schedule("lambda t: 1 - t")
    assert f(0.25) == 0.75
    g = runner.parse_schedule("lambda t: math.cos(pi * t / 2) ** 2")
    assert abs(g(0.5) - 0.5) < 1e-15
    h = runner.parse_schedule("lambda s: -(s - 1) * jnp.exp(0)")
    assert h(0.0) == 1.0
    for bad in ["__import__('os').system('true')", "lambda t: t.__class__", "lambda t, u: t",
                "lambda t: open('x')", "lambda t: [t]", "lambda t: os.sep", "lambda *a: 1"]:
        with pytest.raises((ValueError, SyntaxError)):
            runner.parse_schedule(bad)


def _degrees(ef):
    deg = collections.Counter()
    for a, b in ef.coupled_spin_pairs:
        deg[int(a)] += 1
        deg[int(b)] += 1
    return deg


def test_instance_structures():
    key = runner.dqr.PRNGKey(3)
    hh = runner.generate_lattice("ibm_heavy_hex", 1.0, 1.0, None, None, key)
    assert hh.nodes_number == 127 and hh.edges_number == 144
    assert collections.Counter(_degrees(hh).values()) == {1: 2, 2: 89, 3: 36}
    assert len({tuple(sorted(p)) for p in hh.coupled_spin_pairs.tolist()}) == 144
    loop = runner.generate_lattice("one_heavy_hex_loop", 1.0, 0.5, None, None, key)
    assert loop.nodes_number == 12 and loop.edges_number == 12
    assert set(_degrees(loop).values()) == {2}
    # one cycle through all 12 spins
    adj = collections.defaultdict(list)
    for a, b in loop.coupled_spin_pairs:
        adj[int(a)].append(int(b))
        adj[int(b)].append(int(a))
    seen, prev, cur = [0], None, 0
    while True:
        nxt = [x for x in adj[cur] if x != prev][0]
        if nxt == 0:
            break
        seen.append(nxt)
        prev, cur = cur, nxt
    assert sorted(seen) == list(range(12))
    tree = runner.generate_lattice("small_tree", 1.0, 1.0, None, None, key)
    assert tree.nodes_number == 10 and tree.edges_number == 9 and _degrees(tree)[0] == 3
    rr = runner.generate_lattice("random_regular", 0.3, 2.0, 20, 3, key)
    assert rr.nodes_number == 20 and set(_degrees(rr).values()) == {3}
    mc = runner.generate_lattice("random_regular_maxcut", None, None, 20, 3, key)
    assert np.all(mc.fields == 0) and np.all(mc.coupling_amplitudes == -1)
    with pytest.raises(NotImplementedError):
        runner.generate_lattice("square", 1.0, 1.0, None, None, key)
    # the same key gives the same instance
    again = runner.generate_lattice("random_regular", 0.3, 2.0, 20, 3, key)
    assert np.array_equal(again.coupled_spin_pairs, rr.coupled_spin_pairs)
    assert np.array_equal(again.fields, rr.fields)


QBP_KEYS = {"quantum_annealer_configuration", "quantum_annealer_energy", "density_matrices_history",
            "coupling_amplitudes", "coupled_spin_pairs", "fields", "density_matrices",
            "quantum_annealer_energy_wo_sampling", "vidal_distances_after_regauging",
            "truncation_affected_vidal_distances", "truncation_errors", "entropies"}
EXACT_KEYS = {"quantum_annealer_configuration", "quantum_annealer_energy", "density_matrices_history",
              "coupling_amplitudes", "coupled_spin_pairs", "fields", "density_matrices",
              "quantum_annealer_energy_wo_sampling"}


@pytest.mark.gpu
@pytest.mark.skipif(not gpu_available(), reason="needs a ROCm GPU")
def test_qbp_then_exact_on_the_small_tree(tmp_path):
    """qbp.py then exact.py on the reference's 10-spin tree: the result files
    hold the reference's dataset names, and on a tree (max_chi 8 >= every
    Schmidt rank) BP annealing equals exact annealing."""
    root = _config_dir(str(tmp_path / "configs"))
    out = str(tmp_path / "run")
    cfg = runner.load_config(root, overrides=["emulator_parameters.max_chi=8"])
    bp = runner.run_qbp(cfg, out)
    ex = runner.run_exact(cfg, out)
    with np.load(os.path.join(out, "result.npz")) as f:
        assert set(f.files) == QBP_KEYS
    with np.load(os.path.join(out, "exact_result.npz")) as f:
        assert set(f.files) == EXACT_KEYS
    assert bp["density_matrices"].shape == (10, 2, 2)
    assert np.abs(bp["density_matrices"] - ex["density_matrices"]).max() < 1e-6
    assert np.abs(bp["density_matrices_history"] - ex["density_matrices_history"]).max() < 1e-6
    assert math.isclose(float(bp["quantum_annealer_energy_wo_sampling"]),
                        float(ex["quantum_annealer_energy_wo_sampling"]))


@pytest.mark.gpu
@pytest.mark.skipif(not gpu_available(), reason="needs a ROCm GPU")
def test_bp_against_exact_on_a_loopy_instance(tmp_path):
    """The reference's exact-validation workflow (qbp.py, then exact.py on the
    stored instance) on a 26-spin random 3-regular graph (the size of
    configs/task_generator/exact_validation.yaml), chi = 16.  Loopy BP is an
    approximation: the bound is four times the error measured on MI355X
    (max |<Z>_BP - <Z>_exact| = 4.7e-3); the errors are printed."""
    root = _config_dir(str(tmp_path / "configs"), nodes=26)
    out = str(tmp_path / "run")
    cfg = runner.load_config(root, overrides=["task_generator=rr", "emulator_parameters.max_chi=16",
                                              "quantum_annealing_schedule.steps_number=40",
                                              "quantum_annealing_schedule.record_history=false"])
    bp = runner.run_qbp(cfg, out)
    ex = runner.run_exact(cfg, out)
    ef = runner._energy(bp["coupling_amplitudes"], bp["coupled_spin_pairs"], bp["fields"])
    z_bp = (bp["density_matrices"][:, 0, 0] - bp["density_matrices"][:, 1, 1]).real
    z_ex = (ex["density_matrices"][:, 0, 0] - ex["density_matrices"][:, 1, 1]).real
    dz = float(np.abs(z_bp - z_ex).max())
    e_field_bp = float(-np.dot(ef.fields, z_bp)) / ef.nodes_number
    e_field_ex = float(-np.dot(ef.fields, z_ex)) / ef.nodes_number
    print(f"26 spins, 40 steps: max |<Z>_BP - <Z>_exact| = {dz:.3e}; field energy per spin "
          f"BP {e_field_bp:.5f} exact {e_field_ex:.5f}", flush=True)
    assert dz < 0.02


def test_command_line(tmp_path):
    """python -m dequantizer_amd.runner: help text, a malformed override
    rejected, and (without a GPU) the run raises instead of computing on the
    host."""
    import subprocess
    import sys

    import torch

    r = subprocess.run([sys.executable, "-m", "dequantizer_amd.runner", "--help"], capture_output=True, text=True,
                       cwd=os.path.dirname(os.path.dirname(os.path.abspath(__file__))), timeout=120)
    assert r.returncode == 0 and "qbp" in r.stdout and "--config-dir" in r.stdout
    root = _config_dir(str(tmp_path / "configs"))
    with pytest.raises(ValueError):
        runner.main(["qbp", "--config-dir", root, "--output-dir", str(tmp_path / "o"), "emulator_parameters"])
    if not torch.cuda.is_available():
        with pytest.raises(RuntimeError, match="needs a ROCm GPU"):
            runner.main(["qbp", "--config-dir", root, "--output-dir", str(tmp_path / "o")])
