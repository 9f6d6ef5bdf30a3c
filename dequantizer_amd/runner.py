"""Annealing runs from the reference's YAML configuration groups, with NPZ
result files.

Mirrors examples/quantum_annealing/qbp.py:26-160 (BP annealing of a generated
instance) and examples/quantum_annealing/exact.py:26-123 (the exact annealer on
the instance of an earlier BP run), with the instance generators of
quantum_annealing/src/lattice_generator.py:14-71 and energy_function.py:78-260.

Configuration: a directory laid out like examples/quantum_annealing/configs
(``config.yaml`` with a ``defaults`` list of ``group: option`` entries, one
``<group>/<option>.yaml`` file per choice).  Files are read with
``yaml.safe_load``; overrides use the Hydra spelling (``group=option`` picks
another file, ``group.key=value`` sets one value).  The schedule string
(``'lambda t: 1 - t'``) is parsed into a closure by a whitelist of arithmetic
syntax, never by ``eval``.

Results: the HDF5 datasets of qbp.py:108-160 / exact.py:80-120 under the same
names, in ``result.npz`` / ``exact_result.npz`` (h5py is not a dependency).

Random draws (fields, couplings, the random-regular graph's seed) come from
this package's keys, not JAX threefry streams: the structure of an instance
follows the reference, its values do not.
"""
import argparse
import ast
import logging
import math
import os
from typing import Callable, Dict, List, Optional, Sequence

import numpy as np
import yaml

from . import random as dqr
from .annealing import EnergyFunction, eval_energy, get_scheduler, run_quantum_annealer
from .graph import _EAGLE_BRIDGES, _EAGLE_ROWS

log = logging.getLogger(__name__)

# ------------------------------------------------------------------ configs


def _set_path(cfg: Dict, dotted: str, value):
    keys = dotted.split(".")
    node = cfg
    for k in keys[:-1]:
        node = node.setdefault(k, {})
    node[keys[-1]] = value


def load_config(config_dir: str, config_name: str = "config", overrides: Sequence[str] = ()) -> Dict:
    """The composed configuration: the ``defaults`` groups of
    ``<config_dir>/<config_name>.yaml``, then the overrides in order."""
    with open(os.path.join(config_dir, config_name + ".yaml")) as f:
        top = yaml.safe_load(f) or {}
    choices: Dict[str, str] = {}
    for entry in top.get("defaults", []):
        if isinstance(entry, dict):
            choices.update({str(k): str(v) for k, v in entry.items()})
    values = []
    for ov in overrides:
        if "=" not in ov:
            raise ValueError(f"override {ov!r} is not key=value")
        key, raw = ov.split("=", 1)
        if "." not in key:
            choices[key] = raw
        else:
            values.append((key, yaml.safe_load(raw)))
    cfg: Dict = {k: v for k, v in top.items() if k not in ("defaults", "hydra")}
    for group, option in choices.items():
        path = os.path.join(config_dir, group, option + ".yaml")
        if not os.path.exists(path):
            raise FileNotFoundError(f"config group {group!r}: no option {option!r} ({path})")
        with open(path) as f:
            cfg[group] = yaml.safe_load(f) or {}
    for key, value in values:
        _set_path(cfg, key, value)
    return cfg


_UNARY = {ast.UAdd: lambda a: +a, ast.USub: lambda a: -a}
_BINARY = {ast.Add: lambda a, b: a + b, ast.Sub: lambda a, b: a - b, ast.Mult: lambda a, b: a * b,
           ast.Div: lambda a, b: a / b, ast.Pow: lambda a, b: a ** b, ast.Mod: lambda a, b: a % b}
_FUNCS = {"sin": math.sin, "cos": math.cos, "tan": math.tan, "exp": math.exp, "log": math.log,
          "sqrt": math.sqrt, "tanh": math.tanh, "abs": abs, "min": min, "max": max}
_CONSTS = {"pi": math.pi, "e": math.e}


def parse_schedule(text: str) -> Callable[[float], float]:
    """``'lambda t: <expr>'`` as a closure, where <expr> uses numbers, ``t``,
    ``pi``, ``e``, + - * / ** %, and sin cos tan exp log sqrt tanh abs min max
    (optionally as ``math.``/``jnp.``/``np.`` attributes).  Anything else is
    rejected."""
    tree = ast.parse(text.strip(), mode="eval").body
    if not isinstance(tree, ast.Lambda) or len(tree.args.args) != 1 or tree.args.vararg or tree.args.kwarg:
        raise ValueError(f"schedule {text!r} is not a one-argument lambda")
    arg = tree.args.args[0].arg

    def fname(node) -> str:
        if isinstance(node, ast.Name):
            return node.id
        if isinstance(node, ast.Attribute) and isinstance(node.value, ast.Name) and node.value.id in (
                "math", "jnp", "np", "numpy"):
            return node.attr
        raise ValueError(f"schedule {text!r}: unsupported callee")

    def check(node):
        if isinstance(node, ast.Constant) and isinstance(node.value, (int, float)):
            return
        if isinstance(node, ast.Name) and (node.id == arg or node.id in _CONSTS):
            return
        if isinstance(node, ast.UnaryOp) and type(node.op) in _UNARY:
            return check(node.operand)
        if isinstance(node, ast.BinOp) and type(node.op) in _BINARY:
            check(node.left)
            return check(node.right)
        if isinstance(node, ast.Call) and not node.keywords and fname(node.func) in _FUNCS:
            for a in node.args:
                check(a)
            return
        raise ValueError(f"schedule {text!r}: unsupported syntax {ast.dump(node)[:60]}")

    check(tree.body)

    def ev(node, t):
        if isinstance(node, ast.Constant):
            return node.value
        if isinstance(node, ast.Name):
            return t if node.id == arg else _CONSTS[node.id]
        if isinstance(node, ast.UnaryOp):
            return _UNARY[type(node.op)](ev(node.operand, t))
        if isinstance(node, ast.BinOp):
            return _BINARY[type(node.op)](ev(node.left, t), ev(node.right, t))
        return _FUNCS[fname(node.func)](*[ev(a, t) for a in node.args])

    body = tree.body
    return lambda t: float(ev(body, float(t)))


# -------------------------------------------------------------- instances


def _normal(key, n: int) -> np.ndarray:
    return dqr.generator(key).standard_normal(n)


def _energy(couplings, pairs, fields) -> EnergyFunction:
    return EnergyFunction(np.asarray(couplings, dtype=np.float64), np.asarray(pairs, dtype=np.int64).reshape(-1, 2),
                          np.asarray(fields, dtype=np.float64))


def random_on_ibm_heavy_hex(key, field_std: float = 1.0, coupling_std: float = 1.0) -> EnergyFunction:
    """energy_function.py:78-160: the 127-spin Eagle coupling map (row chains,
    then bridges, in the reference's order)."""
    key, sub = dqr.split(key)
    fields = _normal(sub, 127) * field_std
    pairs = [(i, i + 1) for lo, hi in _EAGLE_ROWS for i in range(lo, hi)] + list(_EAGLE_BRIDGES)
    key, sub = dqr.split(key)
    return _energy(_normal(sub, len(pairs)) * coupling_std, pairs, fields)


def random_on_one_heavy_hex_loop(key, field_std: float = 1.0, coupling_std: float = 1.0) -> EnergyFunction:
    """energy_function.py:175-196: one 12-spin heavy-hex loop (fields and
    couplings from the same subkey, as there)."""
    key, sub = dqr.split(key)
    g = dqr.generator(sub)
    fields = g.standard_normal(12) * field_std
    pairs = [(i, i + 1) for i in range(4)] + [(i, i + 1) for i in range(5, 9)] + [(0, 10), (10, 5), (4, 11), (11, 9)]
    couplings = dqr.generator(sub).standard_normal(len(pairs)) * coupling_std
    return _energy(couplings, pairs, fields)


def random_on_small_tree(key, field_std: float = 1.0, coupling_std: float = 1.0) -> EnergyFunction:
    """energy_function.py:199-220: a 10-spin tree (root 0 with three children,
    two leaves under each)."""
    key, sub = dqr.split(key)
    fields = dqr.generator(sub).standard_normal(10) * field_std
    pairs = [(0, 1), (0, 2), (0, 3), (1, 4), (1, 5), (2, 6), (2, 7), (3, 8), (3, 9)]
    couplings = dqr.generator(sub).standard_normal(len(pairs)) * coupling_std
    return _energy(couplings, pairs, fields)


def _regular_pairs(key, nodes_number: int, degree: int) -> List:
    import networkx as nx

    seed = int(dqr.as_key(key)[0])
    graph = nx.random_regular_graph(degree, nodes_number, seed)
    return [tuple(e) for e in graph.edges]


def random_regular(key, nodes_number: int = 150, degree: int = 3, field_std: float = 1.0,
                   coupling_std: float = 1.0) -> EnergyFunction:
    """energy_function.py:223-243."""
    key, sub = dqr.split(key)
    pairs = _regular_pairs(sub, nodes_number, degree)
    key, sub = dqr.split(key)
    fields = _normal(sub, nodes_number) * field_std
    key, sub = dqr.split(key)
    return _energy(_normal(sub, len(pairs)) * coupling_std, pairs, fields)


def random_regular_maxcut(key, nodes_number: int = 150, degree: int = 3) -> EnergyFunction:
    """energy_function.py:246-260: zero fields, every coupling -1."""
    key, sub = dqr.split(key)
    pairs = _regular_pairs(sub, nodes_number, degree)
    return _energy(-np.ones(len(pairs)), pairs, np.zeros(nodes_number))


def generate_lattice(lattice_type: str, field_std: Optional[float], coupling_std: Optional[float],
                     nodes_number: Optional[int], degree: Optional[int], key) -> EnergyFunction:
    """lattice_generator.py:14-71."""
    key, sub = dqr.split(key)
    if lattice_type in ("ibm_heavy_hex", "one_heavy_hex_loop", "small_tree", "random_regular"):
        if field_std is None or coupling_std is None:
            raise ValueError(f"{lattice_type} needs fields_std and couplings_std")
    if lattice_type == "ibm_heavy_hex":
        return random_on_ibm_heavy_hex(sub, field_std, coupling_std)
    if lattice_type == "one_heavy_hex_loop":
        return random_on_one_heavy_hex_loop(sub, field_std, coupling_std)
    if lattice_type == "small_tree":
        return random_on_small_tree(sub, field_std, coupling_std)
    if lattice_type in ("random_regular", "random_regular_maxcut"):
        if nodes_number is None or degree is None:
            raise ValueError(f"{lattice_type} needs nodes_number and degree")
        if lattice_type == "random_regular":
            return random_regular(sub, nodes_number, degree, field_std, coupling_std)
        return random_regular_maxcut(sub, nodes_number, degree)
    raise NotImplementedError(f"Lattice of type {lattice_type} is not implemented.")


# ----------------------------------------------------------------- drivers


def _opt(cast, x):
    return None if x is None else cast(x)


def _energy_wo_sampling(ef: EnergyFunction, dens: np.ndarray) -> float:
    """qbp.py:128-131: the spin configuration 2 [rho_00 > 0.5] - 1."""
    return eval_energy(ef, 2 * (np.asarray(dens)[:, 0, 0].real > 0.5) - 1)


def run_qbp(cfg: Dict, output_dir: str) -> Dict[str, np.ndarray]:
    """qbp.py:26-160: generate the instance, run BP annealing, write
    ``<output_dir>/result.npz``; returns the datasets written."""
    tg, sch, emu = cfg["task_generator"], cfg["quantum_annealing_schedule"], cfg["emulator_parameters"]
    key = dqr.PRNGKey(int(tg["seed"]))
    key, _ = dqr.split(key)
    ef = generate_lattice(tg["lattice_type"], _opt(float, tg.get("fields_std")), _opt(float, tg.get("couplings_std")),
                          _opt(int, tg.get("nodes_number")), _opt(int, tg.get("degree")), key)
    schedule = get_scheduler(float(sch["total_time_step_size"]), parse_schedule(sch["schedule"]),
                             int(sch["steps_number"]))
    key, sub = dqr.split(key)
    record_history = bool(sch["record_history"])
    log.info("BP annealing of %d spins, %d couplings, %d steps", ef.nodes_number, ef.edges_number, len(schedule))
    res = run_quantum_annealer(ef, schedule, sub, int(emu["max_chi"]),
                               ef.edges_number * int(emu["layers_per_regauging"]), float(emu["accuracy"]),
                               int(emu["max_belief_propagation_iterations"]), bool(emu["synchronous_update"]),
                               str(emu["traversal_type"]), bool(sch["sample_measurements"]), record_history)
    out: Dict[str, np.ndarray] = {}
    if res.configuration is not None:
        out["quantum_annealer_configuration"] = np.asarray(res.configuration)
        out["quantum_annealer_energy"] = np.asarray(eval_energy(ef, res.configuration))
    if record_history:
        out["density_matrices_history"] = np.asarray(res.density_matrices_history)
    out["coupling_amplitudes"] = ef.coupling_amplitudes
    out["coupled_spin_pairs"] = ef.coupled_spin_pairs
    out["fields"] = ef.fields
    out["density_matrices"] = np.asarray(res.density_matrices)
    out["quantum_annealer_energy_wo_sampling"] = np.asarray(_energy_wo_sampling(ef, res.density_matrices))
    out["vidal_distances_after_regauging"] = np.asarray(res.vidal_distances_after_regauging)
    out["truncation_affected_vidal_distances"] = np.asarray(res.truncation_affected_vidal_distances)
    out["truncation_errors"] = np.asarray(res.truncation_errors)
    out["entropies"] = np.asarray(res.entropies)
    os.makedirs(output_dir, exist_ok=True)
    np.savez(os.path.join(output_dir, "result.npz"), **out)
    return out


def run_exact(cfg: Dict, qbp_result_path: str) -> Dict[str, np.ndarray]:
    """exact.py:26-123: the exact annealer on the instance stored by an
    earlier BP run in ``<qbp_result_path>/result.npz``; writes
    ``exact_result.npz`` beside it."""
    from .exact import run_exact_quantum_annealer

    sch = cfg["quantum_annealing_schedule"]
    key = dqr.PRNGKey(int(cfg["task_generator"]["seed"]))
    key, _ = dqr.split(key)
    with np.load(os.path.join(qbp_result_path, "result.npz")) as f:
        ef = _energy(f["coupling_amplitudes"], f["coupled_spin_pairs"], f["fields"])
    schedule = get_scheduler(float(sch["total_time_step_size"]), parse_schedule(sch["schedule"]),
                             int(sch["steps_number"]))
    key, sub = dqr.split(key)
    record_history = bool(sch["record_history"])
    log.info("exact annealing of %d spins, %d steps", ef.nodes_number, len(schedule))
    res = run_exact_quantum_annealer(ef, schedule, sub, bool(sch["sample_measurements"]), record_history)
    out: Dict[str, np.ndarray] = {}
    if res.configuration is not None:
        out["quantum_annealer_configuration"] = np.asarray(res.configuration)
        out["quantum_annealer_energy"] = np.asarray(eval_energy(ef, res.configuration))
    if record_history:
        out["density_matrices_history"] = np.asarray(res.density_matrices_history)
    out["coupling_amplitudes"] = ef.coupling_amplitudes
    out["coupled_spin_pairs"] = ef.coupled_spin_pairs
    out["fields"] = ef.fields
    out["density_matrices"] = np.asarray(res.density_matrices)
    out["quantum_annealer_energy_wo_sampling"] = np.asarray(_energy_wo_sampling(ef, res.density_matrices))
    np.savez(os.path.join(qbp_result_path, "exact_result.npz"), **out)
    return out


def main(argv: Optional[Sequence[str]] = None) -> int:
    ap = argparse.ArgumentParser(description="BP (qbp) or exact annealing from YAML configuration groups")
    ap.add_argument("mode", choices=["qbp", "exact"])
    ap.add_argument("--config-dir", required=True)
    ap.add_argument("--config-name", default="config")
    ap.add_argument("--output-dir", default="outputs/run",
                    help="qbp: where result.npz goes; exact: where the BP run's result.npz is")
    ap.add_argument("overrides", nargs="*", help="group=option or group.key=value")
    # overrides may come before or after the options (group=option ...)
    args = ap.parse_intermixed_args(argv)
    logging.basicConfig(level=logging.INFO)
    cfg = load_config(args.config_dir, args.config_name, args.overrides)
    if args.mode == "qbp":
        run_qbp(cfg, args.output_dir)
    else:
        run_exact(cfg, args.output_dir)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
