"""Benchmark: forward BP annealing run on a 1000-spin random 3-regular graph.

Workload (one GPU): N = 1000 spins, random 3-regular graph (1500 edges),
J in {-1, +1}, h ~ N(0, 0.1^2), transverse-field annealing with dt = 0.1 over a
200-step linear schedule, chi = 16, complex64 contractions with FP64
decompositions, regauge (BP to Vidal distance 1e-5, at most 100 sweeps) once
per layer.  A timed step is one Trotter step: regauge + 1500 simple updates
+ 1000 one-site gates.  The metric is edge-updates/s (one EU = one directed
node->edge message recomputation) over the whole timed run.

Prints one JSON line.  ``--gpus`` above 1 is refused: this build measures one
GPU.
"""
import argparse
import gc
import json
import os
import sys
import time

import numpy as np

METRIC = "edge-updates/s & SU gates/s at χ=16 vs plain GPU and CPU; fwd+adjoint time"
PEAK_FP32_TFLOPS = 157.3
PEAK_HBM_GBS = 8000.0
# HBM bytes of one BP sweep (one node-kernel launch with the synchronous edge
# update fused in) of the default workload (N=1000, chi=16, complex64), from
# FETCH_SIZE / WRITE_SIZE counter passes scaled by a 256 MiB copy (FETCH_SIZE
# reads half its bytes, WRITE_SIZE all): 2 x 38304.5 KiB read + 12000 KiB
# written.
TRAFFIC_BP_SWEEP_DEFAULT = 2 * 38304.5 * 1024 + 12000.0 * 1024


def eu_flops(d: int, chi: int, p: int) -> int:
    """Algorithmic FLOPs of one edge update: d-1 message contractions and one
    Gram, each chi^(d+1) p complex multiply-adds (8 real flops)."""
    return 8 * d * chi ** (d + 1) * p


def eu_bytes(d: int, chi: int, p: int, elem: int = 8) -> float:
    """Algorithmic HBM bytes of one edge update: the node tensor and its d
    incoming messages read once and amortised over its d outputs, plus one
    message written."""
    return elem * (chi ** d * p + d * chi * chi) / d + elem * chi * chi


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=6)
    ap.add_argument("--nodes", type=int, default=1000)
    ap.add_argument("--chi", type=int, default=16)
    ap.add_argument("--schedule-steps", type=int, default=200)
    ap.add_argument("--dt", type=float, default=0.1)
    ap.add_argument("--seed", type=int, default=7)
    ap.add_argument("--traversal", default="bfs")
    ap.add_argument("--bp-schedule", default="sync", choices=["sync", "async"],
                    help="synchronous sweeps, or asynchronous sweeps in --traversal order")
    ap.add_argument("--fixed-sweeps", type=int, default=10,
                    help="BP sweeps per regauge with the convergence test off (fixed work); 0 = converge to 1e-5")
    ap.add_argument("--cpu-baseline", type=int, default=1, help="run the oracle CPU leg (1) or not (0)")
    ap.add_argument("--plain-gpu", type=int, default=1, help="run the torch-eager GPU comparator (1) or not (0)")
    ap.add_argument("--adjoint-steps", type=int, default=200,
                    help="Trotter steps of the forward+adjoint timing (0 skips it)")
    ap.add_argument("--knots", type=int, default=32, help="schedule knots of the forward+adjoint timing")
    ap.add_argument("--batch-instances", type=int, default=64,
                    help="instances of the batched annealing leg (the whole 200-step run of the random 3-regular "
                         "instance set, N in 100..1000, in one engine); 0 skips it")
    ap.add_argument("--eagle-steps", type=int, default=20,
                    help="Trotter steps of the kicked-Ising leg on the 127-qubit heavy-hex lattice (chi 32, "
                         "complex64, nine RX angles in one batched engine); 0 skips it")
    ap.add_argument("--lattice-size", type=int, default=4,
                    help="L of the periodic L^3 lattice leg (degree 6, chi 16, phys 2, Gaussian tensors): BP to "
                         "1e-8 and the Vidal gauge in complex64 and complex128; 0 skips it")
    ap.add_argument("--copy-probe-mb", type=int, default=0,
                    help="before the timed steps, one device-to-device copy of this many MiB (a kernel of known "
                         "traffic against which hardware byte counters are scaled); 0 skips it")
    ap.add_argument("--exact-qubits", type=int, default=30,
                    help="exact state-vector leg: qubits of a random 3-regular instance (16 * 2^n bytes of HBM); "
                         "0 skips it")
    ap.add_argument("--exact-layers", type=int, default=10, help="timed annealing layers of the exact leg")
    ap.add_argument("--gate-order", default="colored", choices=["colored", "pairs"],
                    help="coupling-gate order: pair list sorted by edge colour, or the generator's pair order")
    args = ap.parse_args()
    args.sync = args.bp_schedule == "sync"
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only (--gpus 1); runs over several GPUs are not part of this "
              "benchmark", file=sys.stderr)
        sys.exit(2)

    import torch

    from dequantizer_amd.annealing import Annealer, edge_colored, get_scheduler, linear_schedule, random_regular
    from dequantizer_amd import PRNGKey

    torch.cuda.init()
    ef = random_regular(args.nodes, 3, seed=args.seed, couplings="pm1", field_std=0.1)
    if args.gate_order == "colored":
        ef = edge_colored(ef)
    schedule = get_scheduler(args.dt, linear_schedule, args.schedule_steps)
    if args.warmup + args.steps > len(schedule):
        raise SystemExit("warmup + steps exceeds the schedule length")
    t_build = time.perf_counter()
    ann = Annealer([ef], [PRNGKey(args.seed)], max_chi=args.chi, layers_per_regauging=1, accuracy=1e-5,
                   max_belief_propagation_iterations=100, synchronous_update=args.sync,
                   traversal_type=args.traversal, dtype=torch.complex64, chi=args.chi,
                   fixed_sweeps=args.fixed_sweeps or None)
    emu = ann.emu
    for k in range(args.warmup):
        ann.step(*schedule[k])
        emu._flush()
        torch.cuda.synchronize()
        print(f"bench: warmup step {k} done, chi={emu.chi}, bp sweeps so far {sum(emu.bp_iterations[0])}",
              file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    t_build = time.perf_counter() - t_build

    if args.copy_probe_mb:
        src = torch.ones(args.copy_probe_mb << 18, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    it_before = sum(emu.bp_iterations[0])
    emu.sweep_events = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.warmup, args.warmup + args.steps):
        ann.step(*schedule[k])
        if k % 2 == 0:
            print(f"bench: timed step {k} queued", file=sys.stderr, flush=True)
    emu._flush()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    elapsed = t1 - t0
    sweeps = sum(emu.bp_iterations[0]) - it_before
    n_eu = sweeps * 2 * ef.edges_number
    n_gates = args.steps * ef.edges_number
    # live per-launch duration of the BP sweep (HIP events on torch's stream)
    # each entry times one launch of n sweeps (the regauge's sweeps replay as
    # one captured graph): the per-sweep duration is the total over the sweeps
    total_s, n_sweeps_timed = 0.0, 0
    for ev0, ev1, act, n_sw in emu.sweep_events:
        if (act if isinstance(act, bool) else int(act.sum().item()) > 0):
            total_s += ev0.elapsed_time(ev1) * 1e-3
            n_sweeps_timed += n_sw
    durs = [total_s / n_sweeps_timed] if n_sweeps_timed else []
    emu.sweep_events = None
    d, p, chi = 3, 2, args.chi
    flops_per_sweep = 2 * ef.edges_number * eu_flops(d, chi, p)
    bytes_per_sweep = 2 * ef.edges_number * eu_bytes(d, chi, p)
    sweep_s = float(np.mean(durs)) if durs else float("nan")
    achieved_tflops = flops_per_sweep / sweep_s / 1e12 if durs else None

    result = {
        "metric": METRIC,
        "value": n_eu / elapsed,
        "unit": "edge-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64",
        "data": "synthetic",
        "config": {
            "workload": "BP annealing, random 3-regular N=%d (%d edges), chi=%d, dt=%g, %d-step linear schedule, "
                        "steps %d..%d timed" % (args.nodes, ef.edges_number, chi, args.dt, args.schedule_steps,
                                               args.warmup, args.warmup + args.steps - 1),
            "nodes": args.nodes, "edges": ef.edges_number, "chi": chi, "phys": p,
            "bp_schedule": ("sync" if args.sync else "async-" + args.traversal),
            "bp_accuracy": 1e-5, "regauge": "every layer", "gate_order": args.gate_order,
            "sweeps_per_regauge": (args.fixed_sweeps or "to convergence"),
        },
        "su_gates_per_s": n_gates / elapsed,
        "autotune": [ln for ln in emu.tuning_report.strip().split("\n") if ln],
        "bp_sweeps_timed": int(sweeps),
        "bp_eu_per_s_in_sweeps": (2 * ef.edges_number / sweep_s) if durs else None,
        "setup_and_warmup_s": t_build,
        "roofline": {
            "bound": "mfma",
            "kernel": "BP sweep (k_bp_nodes_d3c16, the synchronous edge update fused in)",
            "achieved": achieved_tflops,
            "peak": PEAK_FP32_TFLOPS,
            "unit": "TFLOP/s",
            "frac": (achieved_tflops / PEAK_FP32_TFLOPS) if achieved_tflops else None,
            "traffic": (TRAFFIC_BP_SWEEP_DEFAULT if (args.nodes, args.chi) == (1000, 16) else None),
            "traffic_unit": "B per launch",
            "algorithmic_flops_per_launch": flops_per_sweep,
            "algorithmic_bytes_per_launch": bytes_per_sweep,
            "avg_launch_s": sweep_s,
            "launches": n_sweeps_timed,
            "timing": "HIP events around each sweep launch (a regauge's fixed sweeps replay as one captured HIP "
                      "graph), total time divided by the number of sweeps",
        },
        "cpu_baseline": None,
    }
    if args.batch_instances:
        result["batched"] = batched_run(args)
        print(f"bench: batched run done: {result['batched']['seconds']:.1f} s", file=sys.stderr, flush=True)
    if args.eagle_steps:
        result["eagle"] = eagle_run(args)
        print(f"bench: eagle leg done: {result['eagle']['seconds']:.1f} s", file=sys.stderr, flush=True)
    if args.exact_qubits:
        result["exact_sv"] = exact_run(args)
        print(f"bench: exact state-vector leg done: {result['exact_sv']['ms_per_layer']:.1f} ms per layer",
              file=sys.stderr, flush=True)
    if args.lattice_size:
        result["lattice"] = lattice_run(args)
        print(f"bench: lattice leg done: {result['lattice']['sweep_ms_c64']:.0f} ms per complex64 sweep",
              file=sys.stderr, flush=True)
    if args.adjoint_steps:
        result["fwd_adjoint"] = adjoint_timing(ef, args)
        print(f"bench: forward+adjoint done: {result['fwd_adjoint']['seconds']:.1f} s", file=sys.stderr, flush=True)
    if args.plain_gpu:
        result["plain_gpu"] = plain_gpu_baseline(ann, ef, schedule[args.warmup + args.steps], args)
        result["vs_plain_gpu"] = result["value"] / result["plain_gpu"]["value"]
        print(f"bench: plain GPU comparator done: {result['plain_gpu']['ms_per_step']:.0f} ms/step", file=sys.stderr,
              flush=True)
    if args.cpu_baseline:
        result["cpu_baseline"] = cpu_baseline(ann, ef, schedule[args.warmup + args.steps], args)
        result["vs_cpu_baseline"] = result["value"] / result["cpu_baseline"]["value"]
        if args.exact_qubits:
            cpu_sv = exact_cpu_baseline(args)
            result["exact_sv"]["cpu_baseline"] = cpu_sv
            result["exact_sv"]["vs_cpu_baseline"] = result["exact_sv"]["amplitude_layers_per_s"] / cpu_sv["value"]
    print(json.dumps(result))


def batched_run(args):
    """The instance set (random 3-regular graphs, N drawn from 100..1000,
    +-1 couplings, N(0, 0.1^2) fields) annealed together in one engine over
    the whole schedule: every launch covers all instances.  EU/s over the
    whole run (regauges, simple updates, one-site gates and the final one-
    and two-site density matrices), and the BP-estimated and sign-
    configuration energy per spin of each instance."""
    import torch
    from dequantizer_amd import PRNGKey
    from dequantizer_amd.annealing import Annealer, edge_colored, get_scheduler, linear_schedule
    from dequantizer_amd.drivers import annealing_instances
    t0 = time.perf_counter()
    efs = [edge_colored(ef) for ef in annealing_instances(args.batch_instances)]
    keys = [PRNGKey(1000 + k) for k in range(len(efs))]
    schedule = get_scheduler(args.dt, linear_schedule, args.schedule_steps)
    ann = Annealer(efs, keys, max_chi=args.chi, layers_per_regauging=1, accuracy=1e-5,
                   max_belief_propagation_iterations=100, synchronous_update=args.sync,
                   traversal_type=args.traversal, dtype=torch.complex64, chi=args.chi,
                   fixed_sweeps=args.fixed_sweeps or None)
    torch.cuda.synchronize()
    setup = time.perf_counter() - t0
    t0 = time.perf_counter()
    for k, (mix, cpl) in enumerate(schedule):
        ann.step(mix, cpl)
        if k % 50 == 0:
            print(f"bench: batched step {k} queued", file=sys.stderr, flush=True)
    res = ann.finish(two_site=True)
    torch.cuda.synchronize()
    seconds = time.perf_counter() - t0
    n_eu = sum(sum(r.bp_iterations) * 2 * ef.edges_number for r, ef in zip(res, efs))
    n_gates = len(schedule) * sum(ef.edges_number for ef in efs)
    e_bp = np.asarray([r.bp_energy / ef.nodes_number for r, ef in zip(res, efs)])
    e_sign = np.asarray([r.sign_energy / ef.nodes_number for r, ef in zip(res, efs)])
    return {"instances": len(efs), "nodes": int(sum(ef.nodes_number for ef in efs)),
            "edges": int(sum(ef.edges_number for ef in efs)), "steps": len(schedule), "seconds": seconds,
            "setup_seconds": setup, "value": n_eu / seconds, "unit": "edge-updates/s",
            "su_gates_per_s": n_gates / seconds, "ms_per_step": 1e3 * seconds / len(schedule),
            "energy_per_spin_bp": [round(float(x), 6) for x in e_bp],
            "energy_per_spin_sign": [round(float(x), 6) for x in e_sign],
            "energy_per_spin_bp_mean": float(e_bp.mean())}


def exact_run(args):
    """Exact state-vector annealing (exact_quantum_annealer.py:84-122) of a
    random 3-regular instance: per layer the coupling phases ride on the first
    pass of the single-qubit mixing layer (12 qubits per pass); HIP events on
    the launch stream time each layer, then the diagonal pass, the mixing
    layer and the density matrices on their own.  Algorithmic bytes of one
    pass: the 2^n complex128 amplitudes read and written once."""
    import torch

    from dequantizer_amd.annealing import HADAMARD, coupling_gates, mixing_gate, random_regular
    from dequantizer_amd.exact import StateVector

    n = args.exact_qubits
    ef = random_regular(n, 3, seed=args.seed)
    pairs = np.asarray(ef.coupled_spin_pairs, dtype=np.int64)
    sv = StateVector(n)
    sv.apply1_layer(HADAMARD)
    sched = [(args.dt * (1 - i / max(1, args.exact_layers)), args.dt * i / max(1, args.exact_layers))
             for i in range(args.exact_layers + 1)]
    stream = torch.cuda.current_stream()
    ev = [[torch.cuda.Event(enable_timing=True) for _ in range(2)] for _ in sched]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i, (mix, cpl) in enumerate(sched):
        diag = np.diagonal(coupling_gates(ef, cpl).reshape(-1, 4, 4), axis1=1, axis2=2)
        mg = mixing_gate(mix)
        ev[i][0].record(stream)
        sv.apply_diagonal_layer1(pairs, diag, mg)
        ev[i][1].record(stream)
        if i == 0:  # the first layer is the warm-up
            torch.cuda.synchronize()
            t0 = time.perf_counter()
    torch.cuda.synchronize()
    seconds = time.perf_counter() - t0
    layer_ms = float(np.mean([e[0].elapsed_time(e[1]) for e in ev[1:]]))
    # the two kernel groups on their own, for their rates
    q = [torch.cuda.Event(enable_timing=True) for _ in range(4)]
    q[0].record(stream)
    sv.apply_diagonal(pairs, diag)
    q[1].record(stream)
    sv.apply1_layer(mg)
    q[2].record(stream)
    rho = sv.rdm1()
    q[3].record(stream)
    torch.cuda.synchronize()
    trace_err = float((rho[:, 0, 0] + rho[:, 1, 1] - 1).abs().max())
    diag_ms, mix_ms, rdm_ms = q[0].elapsed_time(q[1]), q[1].elapsed_time(q[2]), q[2].elapsed_time(q[3])
    del sv, rho
    torch.cuda.empty_cache()
    passes = 1 + max(0, -(-(n - 12) // 8))
    pass_bytes = 2 * 16 * float(2 ** n)
    fused = n >= 12 and pairs.shape[0] <= 64
    layer_passes = passes + (0 if fused else 1)
    return {"qubits": n, "couplings": int(pairs.shape[0]), "layers": args.exact_layers,
            "state_gib": 16 * 2 ** n / 2 ** 30, "seconds": seconds, "ms_per_layer": 1e3 * seconds / args.exact_layers,
            "layer_event_ms": layer_ms, "layer_passes": layer_passes,
            "layer_pass_gbs": layer_passes * pass_bytes / layer_ms / 1e6,
            "layer_pass_frac": layer_passes * pass_bytes / layer_ms / 1e6 / PEAK_HBM_GBS,
            "diagonal_pass_ms": diag_ms, "diagonal_pass_gbs": pass_bytes / diag_ms / 1e6,
            "diagonal_pass_frac": pass_bytes / diag_ms / 1e6 / PEAK_HBM_GBS,
            "mixing_layer_ms": mix_ms, "mixing_passes": passes,
            "mixing_pass_gbs": passes * pass_bytes / mix_ms / 1e6,
            "mixing_pass_frac": passes * pass_bytes / mix_ms / 1e6 / PEAK_HBM_GBS,
            "rdm1_all_ms": rdm_ms, "rdm1_passes": passes,
            "rdm1_pass_gbs": passes * pass_bytes / 2 / rdm_ms / 1e6, "trace_err": trace_err,
            "amplitude_layers_per_s": 2 ** n * args.exact_layers / seconds}


def exact_cpu_baseline(args, n=22):
    """The oracle's state-vector emulator (the reference's tensordot updates,
    NumPy complex128) on one annealing layer of a random 3-regular instance
    with n qubits: the coupling gates one by one, then the mixing gates.
    Reported per amplitude (amplitude-layers/s), the unit of the exact leg."""
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    from dequantizer_amd.annealing import coupling_gates, mixing_gate, random_regular
    from oracle import emulator as oe

    ef = random_regular(n, 3, seed=args.seed)
    ex = oe.ExactQuantumEmulator.__new__(oe.ExactQuantumEmulator)
    rng = np.random.default_rng(args.seed)
    psi = rng.standard_normal(2 ** n) + 1j * rng.standard_normal(2 ** n)
    ex.n, ex.node2pos, ex.state = n, {i: i for i in range(n)}, psi / np.linalg.norm(psi)
    gates = coupling_gates(ef, 0.5 * args.dt)
    mg = mixing_gate(0.5 * args.dt)
    t0 = time.perf_counter()
    for (a, b), g in zip(ef.coupled_spin_pairs, gates):
        ex.apply_q2(g, int(a), int(b))
    for q in range(n):
        ex.apply_q1(mg, q)
    dt = time.perf_counter() - t0
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {"value": 2 ** n / dt, "unit": "amplitude-layers/s", "cores": threads, "kind": "port",
            "sample": "one annealing layer (%d coupling gates, %d mixing gates) of a %d-qubit random 3-regular "
                      "instance on the oracle's state vector (NumPy complex128 tensordot), %.1f s"
                      % (ef.edges_number, n, n, dt),
            "seconds": dt}


if __name__ == "__main__":
    main()
