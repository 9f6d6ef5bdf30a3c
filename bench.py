"""Benchmark: forward BP annealing run on a 1000-spin random 3-regular graph.

Workload (one GPU): N = 1000 spins, random 3-regular graph (1500 edges),
J in {-1, +1}, h ~ N(0, 0.1^2), transverse-field annealing with dt = 0.1 over a
200-step linear schedule, chi = 16, complex64 contractions with FP64
decompositions, a regauge of exactly 10 synchronous BP sweeps (then the Vidal
gauge and distance) before every layer.  A timed step is one whole forward
run: the engine set up on |+...+>, 200 Trotter steps (regauge + 1500 simple
updates + 1000 one-site gates each) and the final one- and two-site density
matrices.  The metric is edge-updates/s (one EU = one directed node->edge
message recomputation) over the timed runs.

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
PEAK_FP64_TFLOPS = 78.6
PEAK_HBM_GBS = 8000.0
# HBM bytes of one BP sweep (one node-kernel launch with the synchronous edge
# update fused in) of the default workload (N=1000, chi=16, complex64), from
# FETCH_SIZE / WRITE_SIZE counter passes scaled by a 256 MiB copy (FETCH_SIZE
# reads half its bytes, WRITE_SIZE all): 2 x 38304.5 KiB read + 12000 KiB
# written.
TRAFFIC_BP_SWEEP_DEFAULT = 2 * 38304.5 * 1024 + 12000.0 * 1024
# HBM bytes of one simple-update launch (500 gates, one edge-colour class) of
# the default workload, from the same counter passes over a whole run (median
# dispatch; the 256 MiB copy reads 131089 KiB by FETCH_SIZE and writes
# 262144 KiB by WRITE_SIZE): 84093.6 KiB fetched (x 2) + 112942 KiB written.
TRAFFIC_SU_LAUNCH_DEFAULT = int((2 * 84093.6 + 112942.0) * 1024)
# HBM bytes of one complex64 BP sweep of the periodic 4^3 lattice leg (64
# degree-6 nodes of 256 MiB): the streaming kernels' counter passes over the
# leg (FETCH_SIZE x 2 + WRITE_SIZE, calibrated as above), summed over a
# sweep's 5 two-product and 6 product-Gram launches: 160.0 GiB for the
# two-product passes (their algorithmic 32 GiB each) and 218.7 GiB for the
# product-Gram passes (192 GiB algorithmic; the 64-byte tile rows of the
# non-contiguous axis pairs are partly fetched twice).
TRAFFIC_LATTICE_SWEEP_L4 = 406597664796
ALGO_BYTES_LATTICE_SWEEP_L4 = 22 * 64 * 256 * 2 ** 20


def eu_flops(d: int, chi: int, p: int) -> int:
    """Algorithmic FLOPs of one edge update: d-1 message contractions and one
    Gram, each chi^(d+1) p complex multiply-adds (8 real flops)."""
    return 8 * d * chi ** (d + 1) * p


def eu_bytes(d: int, chi: int, p: int, elem: int = 8) -> float:
    """Algorithmic HBM bytes of one edge update: the node tensor and its d
    incoming messages read once and amortised over its d outputs, plus one
    message written."""
    return elem * (chi ** d * p + d * chi * chi) / d + elem * chi * chi


def su_flops(m_a: int, m_b: int, n: int, k: int) -> float:
    """Algorithmic FLOPs of one simple update: a reduced QR of each side's
    (m x n) matrix (2 m n^2 - 2/3 n^3 complex multiply-adds), the SVD of the
    n x n core (21 n^3 real flops, four times for complex arithmetic) and the
    two new Gammas (m n k complex multiply-adds per side, k kept values).  At
    chi = 16, degree 3, gate rank 2 (m = 512, n = 32, k = 16): 23.4 MFLOP."""
    qr = sum(2.0 * m * n * n - 2.0 / 3.0 * n ** 3 for m in (m_a, m_b))
    gam = (m_a + m_b) * n * k
    return 8.0 * (qr + gam) + 84.0 * n ** 3


def annealing_run(ef, schedule, args, dtype, sweep_events=None, su_events=None):
    """One whole forward annealing run (quantum_annealer.py:90-152): the
    engine set up on |+...+> (Vidal canonicalisation from random messages),
    every Trotter step of the schedule (regauge, coupling gates, mixing
    gates), then the one- and two-site density matrices of every spin and
    edge and the BP energy."""
    from dequantizer_amd import PRNGKey
    from dequantizer_amd.annealing import Annealer
    ann = Annealer([ef], [PRNGKey(args.seed)], max_chi=args.chi, layers_per_regauging=1, accuracy=1e-5,
                   max_belief_propagation_iterations=100, synchronous_update=args.sync,
                   traversal_type=args.traversal, dtype=dtype, chi=args.chi,
                   fixed_sweeps=args.fixed_sweeps or None)
    ann.emu.sweep_events = sweep_events
    ann.emu.su_events = su_events
    for mix, cpl in schedule:
        ann.step(mix, cpl)
    res = ann.finish(two_site=True)[0]
    ann.emu.sweep_events = ann.emu.su_events = None
    return ann, res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3, help="timed whole annealing runs")
    ap.add_argument("--warmup", type=int, default=1, help="untimed whole annealing runs before them")
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
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="BLAS threads of the CPU leg (0 = the host's usable cores); a 1-thread run follows")
    ap.add_argument("--plain-gpu", type=int, default=1, help="run the torch-eager GPU comparator (1) or not (0)")
    ap.add_argument("--adjoint-steps", type=int, default=200,
                    help="Trotter steps of the forward+adjoint timing (0 skips it)")
    ap.add_argument("--knots", type=int, default=32, help="schedule knots of the forward+adjoint timing")
    ap.add_argument("--adjoint-memory-gb", type=float, default=24.0,
                    help="device memory for the adjoint's records and checkpoints (GiB)")
    ap.add_argument("--adjoint-bp-tol", type=float, default=0.0,
                    help="forward+adjoint timing with every regauge iterated to the BP fixed point (message "
                         "residual below this) and the implicit adjoint; 0 = the fixed sweeps of --fixed-sweeps")
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
                    help="before the timed runs, one device-to-device copy of this many MiB (a kernel of known "
                         "traffic against which hardware byte counters are scaled); 0 skips it")
    ap.add_argument("--exact-qubits", type=int, default=30,
                    help="exact state-vector leg: qubits of a random 3-regular instance (16 * 2^n bytes of HBM); "
                         "0 skips it")
    ap.add_argument("--exact-layers", type=int, default=10, help="timed annealing layers of the exact leg")
    ap.add_argument("--async-steps", type=int, default=5,
                    help="Trotter steps timed with the emulator's default asynchronous depth-first BP schedule "
                         "(0 skips that leg)")
    ap.add_argument("--ising-cube", type=int, default=1,
                    help="time the 6^3 open-cube 3D Ising canonicalisation (launch-bound, chi = 2); 0 skips it")
    ap.add_argument("--gate-order", default="colored", choices=["colored", "pairs"],
                    help="coupling-gate order: pair list sorted by edge colour, or the generator's pair order")
    args = ap.parse_args()
    args.sync = args.bp_schedule == "sync"
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only (--gpus 1); runs over several GPUs are not part of this "
              "benchmark", file=sys.stderr)
        sys.exit(2)
    if args.steps < 1:
        raise SystemExit("--steps must be at least 1")

    import torch

    from dequantizer_amd.annealing import edge_colored, get_scheduler, linear_schedule, random_regular

    torch.cuda.init()
    ef = random_regular(args.nodes, 3, seed=args.seed, couplings="pm1", field_std=0.1)
    if args.gate_order == "colored":
        ef = edge_colored(ef)
    schedule = get_scheduler(args.dt, linear_schedule, args.schedule_steps)
    t_build = time.perf_counter()
    tuning = ""  # the autotuner runs once per process, in the first engine
    for k in range(args.warmup):
        ann_w, _ = annealing_run(ef, schedule, args, torch.complex64)
        tuning = tuning or ann_w.emu.tuning_report
        torch.cuda.synchronize()
        print(f"bench: warmup run {k} done", file=sys.stderr, flush=True)
    t_build = time.perf_counter() - t_build

    if args.copy_probe_mb:
        src = torch.ones(args.copy_probe_mb << 18, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    sweep_events, su_events = [], []
    runs = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        runs.append(annealing_run(ef, schedule, args, torch.complex64, sweep_events, su_events))
        print(f"bench: timed run {k} done", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    ann, res = runs[-1]
    emu = ann.emu
    sweeps = sum(sum(r.bp_iterations) for _, r in runs)
    n_eu = sweeps * 2 * ef.edges_number
    n_trotter = args.steps * len(schedule)
    n_gates = n_trotter * ef.edges_number
    d, p, chi = 3, 2, args.chi

    # BP sweeps: HIP events on torch's stream around each launch (a regauge's
    # fixed sweeps replay as one captured HIP graph); total time / sweeps
    bp_s, bp_sweeps_timed = 0.0, 0
    for ev0, ev1, act, n_sw in sweep_events:
        if (act if isinstance(act, bool) else int(act.sum().item()) > 0):
            bp_s += ev0.elapsed_time(ev1) * 1e-3
            bp_sweeps_timed += n_sw
    sweep_s = bp_s / bp_sweeps_timed if bp_sweeps_timed else float("nan")
    flops_per_sweep = 2 * ef.edges_number * eu_flops(d, chi, p)
    bytes_per_sweep = 2 * ef.edges_number * eu_bytes(d, chi, p)
    bp_tflops = flops_per_sweep / sweep_s / 1e12 if bp_sweeps_timed else None
    # simple updates: HIP events around each launch (one edge-colour class of
    # gates per launch); per-launch durations in launch order
    su_durs = np.asarray([a.elapsed_time(b) * 1e-3 for a, b, _ in su_events])
    su_counts = np.asarray([c for _, _, c in su_events])
    su_s = float(su_durs.sum())
    gate_flops = su_flops(p * chi ** (d - 1), p * chi ** (d - 1), 2 * chi, chi)
    su_launch_s = float(su_durs.mean()) if len(su_durs) else float("nan")
    su_gates_per_launch = float(su_counts.mean()) if len(su_counts) else float("nan")
    su_tflops = su_gates_per_launch * gate_flops / su_launch_s / 1e12 if len(su_durs) else None
    # the full-rank half of each run (the bond spectra fill around step 100)
    per_run = len(su_events) // args.steps if args.steps else 0
    tail = np.concatenate([su_durs[r * per_run + per_run // 2:(r + 1) * per_run] for r in range(args.steps)])
    tail_n = np.concatenate([su_counts[r * per_run + per_run // 2:(r + 1) * per_run] for r in range(args.steps)])
    su_tail_tflops = float(tail_n.mean() * gate_flops / tail.mean() / 1e12) if len(tail) else None
    vd = np.asarray(res.vidal_distances_after_regauging, dtype=np.float64)
    z = (res.density_matrices[:, 0, 0] - res.density_matrices[:, 1, 1]).real

    bp_roof = {
        "bound": "mfma",
        "kernel": "BP sweep (k_bp_nodes_d3c16, the synchronous edge update fused in)",
        "achieved": bp_tflops, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
        "frac": (bp_tflops / PEAK_FP32_TFLOPS) if bp_tflops else None,
        "traffic": (TRAFFIC_BP_SWEEP_DEFAULT if (args.nodes, args.chi) == (1000, 16) else None),
        "traffic_unit": "B per launch",
        "algorithmic_flops_per_launch": flops_per_sweep, "algorithmic_bytes_per_launch": bytes_per_sweep,
        "avg_launch_s": sweep_s, "launches": bp_sweeps_timed, "share_of_run": bp_s / elapsed,
        "timing": "HIP events around each sweep launch on torch's stream (a regauge's fixed sweeps replay as one "
                  "captured HIP graph), total time divided by the number of sweeps",
    }
    su_roof = {
        "bound": "mfma",
        "kernel": "simple update (k_su, one edge-colour class of gates per launch)",
        "achieved": su_tflops, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
        "frac": (su_tflops / PEAK_FP64_TFLOPS) if su_tflops else None,
        "traffic": (TRAFFIC_SU_LAUNCH_DEFAULT if (args.nodes, args.chi) == (1000, 16) else None),
        "traffic_unit": "B per launch",
        "algorithmic_bytes_per_launch": su_gates_per_launch * 2 * 2 * 8 * chi ** d * p,
        "algorithmic_flops_per_launch": su_gates_per_launch * gate_flops,
        "algorithmic_flops_per_gate": gate_flops,
        "avg_launch_s": su_launch_s, "launches": int(len(su_durs)), "share_of_run": su_s / elapsed,
        "achieved_second_half_of_run": su_tail_tflops,
        "frac_second_half_of_run": (su_tail_tflops / PEAK_FP64_TFLOPS) if su_tail_tflops else None,
        "timing": "HIP events around each simple-update launch on torch's stream, mean over all launches of the "
                  "timed runs (FP64 peak: the decompositions run in FP64)",
    }
    dominant, secondary = (su_roof, bp_roof) if su_s > bp_s else (bp_roof, su_roof)
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
            "workload": "BP annealing, random 3-regular N=%d (%d edges), chi=%d, dt=%g, %d-step linear schedule; "
                        "a step is one whole forward run (set-up, %d Trotter steps, final one- and two-site "
                        "density matrices)" % (args.nodes, ef.edges_number, chi, args.dt, args.schedule_steps,
                                               args.schedule_steps),
            "nodes": args.nodes, "edges": ef.edges_number, "chi": chi, "phys": p,
            "bp_schedule": ("sync" if args.sync else "async-" + args.traversal),
            "regauge": "every layer", "gate_order": args.gate_order,
            "sweeps_per_regauge": (args.fixed_sweeps or "to Vidal distance 1e-5"),
            "bp_accuracy": ("not enforced (fixed sweeps)" if args.fixed_sweeps else 1e-5),
        },
        "trotter_steps_timed": n_trotter,
        "ms_per_trotter_step": 1e3 * elapsed / n_trotter,
        "su_gates_per_s": n_gates / elapsed,
        "su_gates_per_s_in_launches": n_gates / su_s if su_s else None,
        "bp_eu_per_s_in_sweeps": (2 * ef.edges_number / sweep_s) if bp_sweeps_timed else None,
        "bp_sweeps_timed": int(sweeps),
        "vidal_distance_after_fixed_sweeps": {"max": float(vd.max()), "median": float(np.median(vd)),
                                              "final": float(vd[-1]), "regauges": int(len(vd))},
        "energy_per_spin_bp": res.bp_energy / ef.nodes_number,
        "magnetisation": float(z.mean()),
        "max_truncation_error": float(np.max(res.truncation_errors)),
        "autotune": [ln for ln in (tuning or runs[0][0].emu.tuning_report).strip().split("\n") if ln],
        "setup_and_warmup_s": t_build,
        "roofline": dominant,
        "roofline_other": secondary,
        "cpu_baseline": None,
    }
    del runs, ann, emu
    gc.collect()
    torch.cuda.empty_cache()
    if args.async_steps:
        result["async_dfs"] = async_dfs_run(ef, schedule, args)
        print(f"bench: asynchronous dfs leg done", file=sys.stderr, flush=True)
    if args.ising_cube:
        result["ising_cube"] = ising_cube_run(args)
        print(f"bench: Ising cube leg done", file=sys.stderr, flush=True)
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
    if args.plain_gpu or args.cpu_baseline:
        # the comparators start from the state of a run stopped halfway (full
        # bond spectra), and time one more Trotter step
        half = len(schedule) // 2
        ann, _ = annealing_run(ef, schedule[:half], args, torch.complex64)
        nxt = schedule[half]
        if args.plain_gpu:
            result["plain_gpu"] = plain_gpu_baseline(ann, ef, nxt, args)
            result["plain_gpu"]["gpu_ms_per_trotter_step"] = result["ms_per_trotter_step"]
            result["vs_plain_gpu"] = result["value"] / result["plain_gpu"]["value"]
            print(f"bench: plain GPU comparator done: {result['plain_gpu']['ms_per_step']:.0f} ms/step",
                  file=sys.stderr, flush=True)
        if args.cpu_baseline:
            result["cpu_baseline"] = cpu_baseline(ann, ef, nxt, args, args.cpu_threads or usable_cores())
            result["vs_cpu_baseline"] = result["value"] / result["cpu_baseline"]["value"]
            print(f"bench: CPU leg done", file=sys.stderr, flush=True)
            result["cpu_baseline_1thread"] = cpu_baseline(ann, ef, nxt, args, 1)
            print(f"bench: one-thread CPU leg done", file=sys.stderr, flush=True)
            if args.exact_qubits:
                cpu_sv = exact_cpu_baseline(args)
                result["exact_sv"]["cpu_baseline"] = cpu_sv
                result["exact_sv"]["vs_cpu_baseline"] = result["exact_sv"]["amplitude_layers_per_s"] / cpu_sv["value"]
    print(json.dumps(result))


def usable_cores() -> int:
    """CPUs this process may run on (its affinity mask)."""
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def cpu_model() -> str:
    """The host CPU's model name (lscpu), for the CPU legs."""
    try:
        import subprocess
        out = subprocess.run(["lscpu"], stdout=subprocess.PIPE, stderr=subprocess.DEVNULL, text=True).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def async_dfs_run(ef, schedule, args, warm=5):
    """The emulator's default BP schedule (asynchronous sweeps in depth-first
    traversal order, bp_quantum_emulator.py:56-59): every sweep reads the
    messages already updated in it, which levels the 1000-node graph into
    about 730 dependent launches per sweep.  Trotter steps ``warm`` onwards
    are timed with the same fixed sweeps per regauge."""
    import torch
    from dequantizer_amd import PRNGKey
    from dequantizer_amd.annealing import Annealer
    ann = Annealer([ef], [PRNGKey(args.seed)], max_chi=args.chi, layers_per_regauging=1, accuracy=1e-5,
                   max_belief_propagation_iterations=100, synchronous_update=False, traversal_type="dfs",
                   dtype=torch.complex64, chi=args.chi, fixed_sweeps=args.fixed_sweeps or 10)
    for mix, cpl in schedule[:warm]:
        ann.step(mix, cpl)
    ann.emu._flush()
    torch.cuda.synchronize()
    it0 = sum(ann.emu.bp_iterations[0])
    t0 = time.perf_counter()
    for mix, cpl in schedule[warm: warm + args.async_steps]:
        ann.step(mix, cpl)
    ann.emu._flush()
    torch.cuda.synchronize()
    seconds = time.perf_counter() - t0
    sweeps = sum(ann.emu.bp_iterations[0]) - it0
    n_levels = ann.emu.topo.schedule(False)[0]
    return {"steps": args.async_steps, "seconds": seconds, "ms_per_trotter_step": 1e3 * seconds / args.async_steps,
            "sweeps": int(sweeps), "levels_per_sweep": int(n_levels),
            "value": sweeps * 2 * ef.edges_number / seconds, "unit": "edge-updates/s"}


def ising_cube_run(args):
    """The 3D classical Ising square-root network on the open 6^3 cube (chi =
    2, 216 nodes, 540 edges; examples/3d_ising_canonicalization.py:89-160):
    asynchronous depth-first BP to a message residual of 1e-10, then the
    Vidal gauge and the mean magnetisation, for beta = 0.1 .. 0.6.  At chi = 2
    every launch is latency-bound: the figure is time per sweep, not a
    roofline fraction."""
    import torch
    from dequantizer_amd.drivers import process_ising_sqrt_partition_network
    betas = [0.1, 0.2, 0.3, 0.4, 0.5, 0.6]
    process_ising_sqrt_partition_network(3, 6, betas[:1], -1.0, 0.5, 1e-10, True, criterion="residual",
                                         max_iterations=500)  # set-up of the handles and schedules
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    res = process_ising_sqrt_partition_network(3, 6, betas, -1.0, 0.5, 1e-10, True, criterion="residual",
                                               max_iterations=500)
    torch.cuda.synchronize()
    seconds = time.perf_counter() - t0
    sweeps = int(sum(res.number_of_iterations))
    return {"L": 6, "nodes": 216, "edges": 540, "chi": 2, "betas": betas, "seconds": seconds, "sweeps": sweeps,
            "ms_per_sweep": 1e3 * seconds / sweeps, "iterations": [int(x) for x in res.number_of_iterations],
            "magnetisation": [round(float(m), 10) for m in res.average_magnetization],
            "note": "each sweep includes the host-side residual check (one device read-back per sweep)"}


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


def eagle_run(args, chi=32):
    """Kicked-Ising dynamics on the 127-qubit heavy-hex lattice: RX(theta) on
    every qubit, then RZZ(-pi/2) in three edge-colour layers, each followed by
    a regauge (synchronous BP to Vidal distance 1e-5, at most 100 sweeps);
    chi = 32, complex64; theta in linspace(0, pi/2, 9) as nine instances of
    one batched engine.  <Z> of qubit 62 after the last step, per angle."""
    import torch
    from dequantizer_amd.drivers import run_eagle_simulations
    from dequantizer_amd.graph import get_heavy_hex_ibm_eagle_lattice
    thetas = np.linspace(0.0, np.pi / 2, 9)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    zs = run_eagle_simulations(thetas, chi, args.eagle_steps, 1e-5, is_infinite=False, synchronous_update=True,
                               dtype=torch.complex64, chi=chi)
    torch.cuda.synchronize()
    seconds = time.perf_counter() - t0
    # the BP sweep launches of a second run, timed with HIP events (the
    # streaming node update at chi = 32: the plan of products and Grams)
    events: list = []
    run_eagle_simulations(thetas, chi, args.eagle_steps, 1e-5, is_infinite=False, synchronous_update=True,
                          dtype=torch.complex64, chi=chi, sweep_events=events)
    torch.cuda.synchronize()
    lat = get_heavy_hex_ibm_eagle_lattice()
    flops_inst = sum(len(n.neighbors) * eu_flops(len(n.neighbors), chi, 2) for n in lat.nodes.values())
    sweep_s, flops = 0.0, 0.0
    for ev0, ev1, act, _ in events:
        n_act = int(act.sum().item())
        if n_act:
            sweep_s += ev0.elapsed_time(ev1) * 1e-3
            flops += n_act * flops_inst
    achieved = flops / sweep_s / 1e12 if sweep_s > 0 else None
    return {"qubits": 127, "chi": chi, "steps": args.eagle_steps, "angles": 9, "seconds": seconds,
            "ms_per_step_all_angles": 1e3 * seconds / args.eagle_steps,
            "z62_final": [round(float(z[-1]), 6) for z in zs], "thetas": [round(float(t), 6) for t in thetas],
            "bp_sweeps": len(events), "bp_sweep_ms": 1e3 * sweep_s / max(1, len(events)),
            "roofline": {"bound": "mfma", "kernel": "BP sweep (streaming node update, chi=32, complex64)",
                         "achieved": achieved, "peak": 157.3, "unit": "TFLOP/s",
                         "frac": achieved / 157.3 if achieved else None, "traffic": None,
                         "algorithmic_flops_per_instance_sweep": flops_inst,
                         "timing": "HIP events around each BP sweep launch of a second identical run, "
                                   "algorithmic FLOPs (8 d^2 chi^(d+1) p per degree-d node, chi = 32 padded) of "
                                   "the instances active in that sweep"}}


def lattice_run(args, chi=16, tol=1e-8, max_sweeps=60):
    """Periodic L^3 cubic lattice, every node of degree 6, chi = 16, phys 2
    (256 MiB per node tensor in complex64): node tensors with entries
    (x + iy) / sqrt(2 chi^6), x, y ~ N(0, 1) (torch generator, seed 5),
    messages A A^H / tr.  Synchronous BP until the mean message change is
    below 1e-8, then the Vidal gauge, in complex64 and in complex128; the
    complex64 messages and lambdas are compared with the complex128 ones
    (that path matches the oracle at this node shape in the tests)."""
    import torch
    import dequantizer_amd as dq
    lat = dq.get_nd_lattice([args.lattice_size] * 3, 2, chi, False)
    trav = list(lat.get_traversal_iterator())
    out = {"L": args.lattice_size, "nodes": lat.nodes_number, "edges": lat.edges_number, "degree": 6, "chi": chi}
    res = {}
    for name, dt, np_dt in (("c64", torch.complex64, np.complex64), ("c128", torch.complex128, np.complex128)):
        gen = torch.Generator(device="cuda").manual_seed(5)
        tensors = {}
        for el in trav:
            if isinstance(el, dq.Node):
                x = torch.randn((*el.bond_shape, 2), dtype=torch.float64, device="cuda", generator=gen)
                y = torch.randn((*el.bond_shape, 2), dtype=torch.float64, device="cuda", generator=gen)
                tensors[el.id] = (torch.complex(x, y) / np.sqrt(2.0 * chi ** 6)).to(dt)
                del x, y
        rng = np.random.default_rng(5)
        msgs = {}
        for el in trav:
            for nb in el.neighbors:
                a = rng.standard_normal((chi, chi)) + 1j * rng.standard_normal((chi, chi))
                m = a @ a.conj().T
                msgs[dq.MessageID(el.id, nb.id)] = (m / np.trace(m)).astype(np_dt)
        bp = dq.get_belief_propagation_map(trav, True)
        vg = dq.get_vidal_gauge_fixing_map(trav, 1e-10)
        m = bp(tensors, msgs)  # first sweep: upload and workspace set-up
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        sweeps, r = 1, float("inf")
        while sweeps < max_sweeps:
            new = bp(tensors, m)
            r = float(dq.messages_frob_distance(new, m))
            m = new
            sweeps += 1
            if r < tol:
                break
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        gam, lam = vg(tensors, m)
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        out[f"sweeps_{name}"] = sweeps
        out[f"residual_{name}"] = r
        out[f"sweep_ms_{name}"] = 1e3 * (t1 - t0) / (sweeps - 1)
        out[f"vidal_gauge_ms_{name}"] = 1e3 * (t2 - t1)
        res[name] = ({k: v.cpu().numpy() for k, v in m.items()}, {k: v.cpu().numpy() for k, v in lam.items()})
        del tensors, gam, lam, m, new, bp, vg
        torch.cuda.empty_cache()
    eu = 2 * lat.edges_number
    out["eu_per_s_c64"] = eu / (out["sweep_ms_c64"] * 1e-3)
    out["algorithmic_tflops_c64"] = eu * eu_flops(6, chi, 2) / (out["sweep_ms_c64"] * 1e-3) / 1e12
    out["roofline"] = {"bound": "mfma", "kernel": "BP sweep (streaming node update, degree 6, chi=16, complex64)",
                       "achieved": out["algorithmic_tflops_c64"], "peak": 157.3, "unit": "TFLOP/s",
                       "frac": out["algorithmic_tflops_c64"] / 157.3,
                       "traffic": TRAFFIC_LATTICE_SWEEP_L4 if args.lattice_size == 4 else None,
                       "traffic_unit": "B per sweep (the streaming kernels, counter passes)",
                       "algorithmic_bytes_per_sweep": ALGO_BYTES_LATTICE_SWEEP_L4 if args.lattice_size == 4 else None,
                       "algorithmic_flops_per_sweep": eu * eu_flops(6, chi, 2),
                       "timing": "wall time of the sweeps after the first, synchronised (one sweep per map call)"}
    (m64, l64), (m128, l128) = res["c64"], res["c128"]
    out["message_rel_err_c64_vs_c128"] = float(max(np.linalg.norm(m64[k] - m128[k]) / np.linalg.norm(m128[k])
                                                   for k in m128))
    out["lambda_abs_err_c64_vs_c128"] = float(max(np.abs(l64[k] - l128[k]).max() for k in l128))
    return out


def adjoint_timing(ef, args):
    """Energy of the annealing run and its gradient with respect to the
    schedule knots and every coupling (hand-written adjoint, complex128,
    synchronous BP, fixed sweeps per regauge): wall time of one forward and
    of forward + adjoint on the same instance, with the records and
    checkpoints held to ``--adjoint-memory-gb`` (the backward recomputes all
    segments but the last from their checkpoints).  The forward + adjoint
    runs twice; ``seconds`` is the second call (a gradient-based schedule
    search calls it repeatedly, and the first takes its device memory from
    the driver), ``first_call_seconds`` the first."""
    import torch
    from dequantizer_amd import PRNGKey
    from dequantizer_amd.adjoint import AnnealingGradient
    gc.collect()
    torch.cuda.empty_cache()  # blocks cached by the earlier legs
    t0 = time.perf_counter()
    run = AnnealingGradient(ef, PRNGKey(args.seed), chi=args.chi, n_steps=args.adjoint_steps, dt=args.dt,
                            n_knots=args.knots,
                            fixed_sweeps=None if args.adjoint_bp_tol > 0 else (args.fixed_sweeps or 10),
                            bp_tol=args.adjoint_bp_tol or 1e-12)
    torch.cuda.synchronize()
    setup = time.perf_counter() - t0
    knots = np.linspace(0.0, 1.0, args.knots)
    torch.cuda.synchronize()
    torch.cuda.reset_peak_memory_stats()
    held_gb = torch.cuda.memory_allocated() / 1e9
    retries0 = torch.cuda.memory_stats().get("num_alloc_retries", 0)
    t0 = time.perf_counter()
    e_fwd = run.value(knots)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    run.value_and_grad(knots, memory_gb=args.adjoint_memory_gb)
    torch.cuda.synchronize()
    run.forward_sweeps.clear()
    run.adjoint_iterations.clear()
    t2 = time.perf_counter()
    e, dk, dj = run.value_and_grad(knots, memory_gb=args.adjoint_memory_gb)
    torch.cuda.synchronize()
    t3 = time.perf_counter()
    out = {"seconds": t3 - t2, "first_call_seconds": t2 - t1, "forward_seconds": t1 - t0, "setup_seconds": setup,
           "steps": args.adjoint_steps,
           "knots": args.knots, "couplings": int(ef.edges_number), "dtype": "c128", "backward_cut": run.cut,
           "sweeps_per_regauge": (f"to residual {args.adjoint_bp_tol:g} (implicit adjoint)" if args.adjoint_bp_tol > 0
                                  else args.fixed_sweeps or 10),
           "bp_sweeps_per_call": int(np.sum(run.forward_sweeps)) if args.adjoint_bp_tol > 0 else None,
           "adjoint_iterations_per_call": int(np.sum(run.adjoint_iterations)) if args.adjoint_bp_tol > 0 else None,
           "memory_budget_gb": args.adjoint_memory_gb,
           "segment_steps": run.segment_steps, "record_gb_per_step": run.record_bytes() / 2 ** 30, "energy": e,
           "energy_forward_only": e_fwd,
           "grad_norm_knots": float(np.linalg.norm(dk)), "grad_norm_couplings": float(np.linalg.norm(dj)),
           "peak_device_memory_gb": torch.cuda.max_memory_allocated() / 1e9,
           "held_before_gb": held_gb,
           "allocator_retries": torch.cuda.memory_stats().get("num_alloc_retries", 0) - retries0,
           "c64_forward": None}
    if args.adjoint_bp_tol <= 0:
        # the complex64 forward (the benchmark engine's precision) with the
        # complex128 adjoint: time, and energy and gradients against the above
        del run
        gc.collect()
        torch.cuda.empty_cache()
        r64 = AnnealingGradient(ef, PRNGKey(args.seed), chi=args.chi, n_steps=args.adjoint_steps, dt=args.dt,
                                n_knots=args.knots, fixed_sweeps=args.fixed_sweeps or 10,
                                forward_dtype=torch.complex64)
        r64.value_and_grad(knots, memory_gb=args.adjoint_memory_gb)
        torch.cuda.synchronize()
        t4 = time.perf_counter()
        e64, dk64, dj64 = r64.value_and_grad(knots, memory_gb=args.adjoint_memory_gb)
        torch.cuda.synchronize()
        t5 = time.perf_counter()
        out["c64_forward"] = {"seconds": t5 - t4, "forward_dtype": "c64", "adjoint_dtype": "c128",
                              "segment_steps": r64.segment_steps,
                              "record_gb_per_step": r64.record_bytes() / 2 ** 30, "energy": e64,
                              "energy_abs_diff_vs_c128_forward": abs(e64 - e),
                              "knot_grad_rel_diff": float(np.abs(dk64 - dk).max() / np.abs(dk).max()),
                              "coupling_grad_rel_diff": float(np.abs(dj64 - dj).max() / np.abs(dj).max())}
        del r64
        gc.collect()
        torch.cuda.empty_cache()
    return out


def plain_gpu_baseline(ann, ef, step_times, args, sample_sweeps=2, sample_gates=150):
    """The same Trotter step written as plain torch eager ops on the GPU: one
    node, edge or gate at a time (tensordot for contractions in complex64,
    torch.linalg eigh / svd / qr in complex128), no batching, no fused kernels.
    Started from the engine's state after the timed steps.  A bounded sample is
    timed (the first Vidal gauge and distance, ``sample_sweeps`` BP sweeps, the
    final gauge and distance, ``sample_gates`` simple updates, all one-site
    gates) and the step time is extrapolated to the full work (fixed sweeps and
    every gate).  Reports EU/s = EU per step / step time."""
    import torch
    from dequantizer_amd.annealing import coupling_gates, mixing_gate
    emu = ann.emu
    emu._flush()
    topo, chi, eps = emu.topo, emu.chi, 1e-10
    c64, c128 = torch.complex64, torch.complex128
    dims = emu.dims.copy()
    lam_all = emu.lam.detach().view(-1, chi)
    lams = [lam_all[j, : int(dims[j])].clone() for j in range(topo.n_edges)]
    adj = [[(int(topo.adj_edge[k]), int(topo.adj_side[k])) for k in range(topo.adj_ptr[i], topo.adj_ptr[i + 1])]
           for i in range(topo.n_nodes)]
    gam = []
    for i in range(topo.n_nodes):
        deg, off = int(topo.deg[i]), int(emu.h.offsets[i])
        full = emu.gam.detach()[off: off + chi ** deg * 2].view(*((chi,) * deg), 2)
        gam.append(full[tuple(slice(0, int(dims[e])) for e, _ in adj[i])].clone())

    def bcast(v, axis, nd):
        shape = [1] * nd
        shape[axis] = v.shape[0]
        return v.reshape(shape)

    def sym(t, ls):
        for a, lm in enumerate(ls):
            t = t * bcast(torch.sqrt(lm * (lm > eps)).to(c64), a, t.ndim)
        return t / torch.linalg.vector_norm(t)

    def on_axis(t, m, a):
        return torch.movedim(torch.tensordot(m, t, dims=([1], [a])), 0, a)

    def node_update(t, msgs):
        out = []
        for idx in range(t.ndim - 1):
            u = t
            for a, m in enumerate(msgs):
                if a != idx:
                    u = on_axis(u, m, a)
            others = [a for a in range(t.ndim) if a != idx]
            g = torch.tensordot(t.conj(), u, dims=(others, others))
            out.append(g / torch.trace(g))
        return out

    def edge_svd(mf, mb):
        lf, uf = torch.linalg.eigh(mf.to(c128))
        lb, ub = torch.linalg.eigh(mb.to(c128))
        inv = lambda x: 1.0 / ((x < eps) * 3.4028234663852886e38 + x)
        fn = lambda u, v: (u * v.to(c128)) @ u.conj().T
        isf, isb = fn(uf, torch.sqrt(inv(lf))), fn(ub, torch.sqrt(inv(lb)))
        sf, sb = fn(uf, torch.sqrt(lf * (lf > eps))), fn(ub, torch.sqrt(lb * (lb > eps)))
        u, lm, vh = torch.linalg.svd(sf @ sb.T, full_matrices=False)
        lm = lm / torch.linalg.vector_norm(lm)
        lm = lm * (lm > eps)
        mask = (lm > eps).to(c128)
        u = u * mask
        u = u / torch.linalg.vector_norm(u)
        vh = vh * mask[:, None]
        vh = vh / torch.linalg.vector_norm(vh)
        return (isf @ u).to(c64), (isb @ vh.T).to(c64), lm

    def absorb(t, xs):
        for a, x in enumerate(xs):
            t = on_axis(t, x.T, a)
        return t / torch.linalg.vector_norm(t)

    def vdist(t, ls):
        total = 0.0
        for idx in range(t.ndim - 1):
            u = t
            for a, lm in enumerate(ls):
                if a != idx:
                    u = u * bcast((lm ** 2).to(c64), a, t.ndim)
            others = [a for a in range(t.ndim) if a != idx]
            r = torch.tensordot(t.conj(), u, dims=(others, others)).to(c128)
            r = r / torch.trace(r)
            ideal = torch.diag((torch.diagonal(r).real > eps).to(c128))
            total = total + torch.linalg.svdvals(r - ideal / torch.trace(ideal)).sum()
        return total

    def regauge_vg(tensors, msgs):
        xs, new_l = {}, {}
        for j in range(topo.n_edges):
            xf, xb, lm = edge_svd(msgs[(j, 0)], msgs[(j, 1)])
            xs[(j, 0)], xs[(j, 1)], new_l[j] = xf, xb, lm
        new_t = [absorb(tensors[i], [xs[(e, s)] for e, s in adj[i]]) for i in range(topo.n_nodes)]
        dist = sum(vdist(new_t[i], [new_l[e] for e, _ in adj[i]]) for i in range(topo.n_nodes))
        return new_t, new_l, dist

    def sweep(tensors, e2n):
        n2e = {}
        for i in range(topo.n_nodes):
            outs = node_update(tensors[i], [e2n[(e, s)] for e, s in adj[i]])
            for (e, s), m in zip(adj[i], outs):
                n2e[(e, s)] = m
        return {(e, s): n2e[(e, 1 - s)] for (e, s) in n2e}

    def su(i_a, i_b, j, g):
        mat = g.permute(0, 2, 1, 3).reshape(4, 4)
        uu, ss, vvh = torch.linalg.svd(mat)
        r = int((ss > eps).sum())
        root = torch.sqrt(ss[:r])
        h1 = (uu[:, :r] * root).reshape(2, 2, r).permute(2, 0, 1).to(c64)
        h2 = (vvh[:r] * root[:, None]).reshape(r, 2, 2).to(c64)
        sides = []
        for node, h in ((i_a, h1), (i_b, h2)):
            t = gam[node]
            idx = [k for k, (e, _) in enumerate(adj[node]) if e == j][0]
            for a, (e, _) in enumerate(adj[node]):
                if a != idx:
                    t = t * bcast(lams[e].to(c64), a, t.ndim)
            th = torch.einsum("...i,qoi->...qo", t, h)
            th = torch.movedim(th, -2, idx + 1)
            shape = list(t.shape)
            shape[idx] *= r
            th = th.reshape(shape)
            th = torch.movedim(th / torch.linalg.vector_norm(th), idx, -1)
            q, rr = torch.linalg.qr(th.reshape(-1, th.shape[-1]).to(c128))
            q = q.reshape(*th.shape[:-1], -1)
            for a, (e, _) in enumerate([x for k, x in enumerate(adj[node]) if k != idx]):
                lm = lams[e]
                q = q * bcast((1.0 / ((lm < eps) * 3.4028234663852886e38 + lm)).to(c128), a, q.ndim)
            sides.append((q, rr, idx))
        (qa, ra, ia), (qb, rb, ib) = sides
        lam_rep = lams[j].repeat_interleave(r)
        u, sv, vh = torch.linalg.svd((ra * lam_rep.to(c128)) @ rb.T, full_matrices=False)
        sv = sv * (sv > eps)
        rank = min(args.chi, sv.shape[0])
        gam[i_a] = torch.movedim(torch.tensordot(qa, u[:, :rank], dims=([qa.ndim - 1], [0])), -1, ia).to(c64)
        gam[i_b] = torch.movedim(torch.tensordot(qb, vh[:rank], dims=([qb.ndim - 1], [1])), -1, ib).to(c64)
        lams[j] = sv[:rank].to(torch.float64)

    torch.cuda.synchronize()
    t0 = time.perf_counter()
    tensors = [sym(gam[i], [lams[e] for e, _ in adj[i]]) for i in range(topo.n_nodes)]
    msgs = {(j, s): torch.diag(lams[j].to(c64)) for j in range(topo.n_edges) for s in range(2)}
    regauge_vg(tensors, msgs)
    torch.cuda.synchronize()
    t_first = time.perf_counter() - t0
    t0 = time.perf_counter()
    for _ in range(sample_sweeps):
        msgs = sweep(tensors, msgs)
    torch.cuda.synchronize()
    t_sweep = (time.perf_counter() - t0) / sample_sweeps
    t0 = time.perf_counter()
    new_t, new_l, _ = regauge_vg(tensors, {k: v for k, v in msgs.items()})
    torch.cuda.synchronize()
    t_last = time.perf_counter() - t0
    gam[:] = new_t
    for j in range(topo.n_edges):
        lams[j] = new_l[j]
    mix, cpl = step_times
    gates = torch.from_numpy(coupling_gates(ef, cpl)).to("cuda")
    edge_of = {}
    for j in range(topo.n_edges):
        a, b = (int(x) for x in topo.edge_node[j])
        edge_of[(a, b)] = edge_of[(b, a)] = j
    nidx = topo.node_index[0]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k, (a, b) in enumerate(ef.coupled_spin_pairs[:sample_gates]):
        ia, ib = nidx[int(a)], nidx[int(b)]
        su(ia, ib, edge_of[(ia, ib)], gates[k])
    torch.cuda.synchronize()
    t_gate = (time.perf_counter() - t0) / min(sample_gates, ef.edges_number)
    mg = torch.from_numpy(mixing_gate(mix)).to("cuda", c64)
    t0 = time.perf_counter()
    for i in range(topo.n_nodes):
        t = torch.tensordot(gam[i], mg, dims=([gam[i].ndim - 1], [1]))
        gam[i] = t / torch.linalg.vector_norm(t)
    torch.cuda.synchronize()
    t_apply1 = time.perf_counter() - t0
    sweeps = args.fixed_sweeps or 10
    t_step = t_first + sweeps * t_sweep + t_last + ef.edges_number * t_gate + t_apply1
    eu = sweeps * 2 * ef.edges_number
    return {"value": eu / t_step, "unit": "edge-updates/s", "kind": "torch eager, one element at a time",
            "ms_per_step": 1e3 * t_step, "gates_per_s": 1.0 / t_gate,
            "sample": "timed: first Vidal gauge + distance %.2f s, %d BP sweeps at %.3f s each, final gauge + "
                      "distance %.2f s, %d simple updates at %.2f ms each, %d one-site gates %.3f s; step "
                      "extrapolated to %d sweeps and %d gates" % (
                          t_first, sample_sweeps, t_sweep, t_last, min(sample_gates, ef.edges_number),
                          1e3 * t_gate, topo.n_nodes, t_apply1, sweeps, ef.edges_number)}


def cpu_baseline(ann, ef, step_times, args, threads: int):
    """The oracle (NumPy complex128 restatement; the reference computes in
    complex128) on one Trotter step of the same instance, started from the
    GPU state: regauge + all simple updates + mixing gates, with the BLAS /
    OpenMP thread pools limited to ``threads``.  Reports EU/s = oracle BP EU in
    that step / its time."""
    from threadpoolctl import threadpool_limits
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    import torch
    from oracle import annealing as oa
    from oracle import emulator as oe
    from oracle import graph as og

    emu = ann.emu
    emu._flush()
    topo = emu.topo
    chi = emu.chi
    g = og.graph_from_pairs(ef.nodes_number, ef.coupled_spin_pairs)
    gam = emu.gam.detach().cpu().numpy().astype(np.complex128)
    lam = emu.lam.detach().cpu().numpy().reshape(-1, chi)
    tensors, lambdas = {}, {}
    for j, e in enumerate(topo.edges):
        d = int(emu.dims[j])
        lambdas[e.id] = lam[j, :d].copy()
        g.edges[e.id].dimension = d
    for i, node in enumerate(topo.nodes):
        deg = int(topo.deg[i])
        off = int(emu.h.offsets[i])
        full = gam[off: off + chi ** deg * 2].reshape((chi,) * deg + (2,))
        dims = [int(emu.dims[topo.adj_edge[k]]) for k in range(topo.adj_ptr[i], topo.adj_ptr[i + 1])]
        tensors[node.id] = full[tuple(slice(0, x) for x in dims) + (slice(None),)].copy()
        g.nodes[node.id].bond_shape = tuple(dims)
    # oracle emulator with this state (constructed on |0>, then replaced)
    o = oe.BPQuantumEmulator.__new__(oe.BPQuantumEmulator)
    o.tensor_graph = g
    o.prng_key = og.make_key(args.seed)
    o.max_chi, o.max_iters, o.per_regauge, o.acc = args.chi, 100, ef.edges_number, 1e-5
    o.gates_since = ef.edges_number
    o.traverser = g.traversal(ordering=args.traversal)
    from oracle import mappings as om
    o.bp_map = om.belief_propagation_map(o.traverser, args.sync)
    o.vg_map = om.vidal_gauge_fixing_map(o.traverser, 1e-10)
    o.vd_map = om.vidal_gauge_distance_map(o.traverser, 1e-10)
    o.sg_map = om.symmetric_gauge_fixing_map(o.traverser, 1e-10)
    o.tensors, o.lambdas = tensors, lambdas
    o.vidal_distances_after_regauging, o.truncation_affected_vidal_distances = [], []
    o.truncation_errors, o.entropies, o.bp_iterations = [], [], []
    o._initial_messages = None
    o.fixed_sweeps = args.fixed_sweeps or None
    mix, cpl = step_times
    deg = [g.nodes[i].degree for i in range(ef.nodes_number)]
    with threadpool_limits(limits=threads):
        t0 = time.perf_counter()
        for (a, b), jv in zip(ef.coupled_spin_pairs, ef.coupling_amplitudes):
            a, b = int(a), int(b)
            o.apply_q2(oa.coupling_gate(jv, ef.fields[a], ef.fields[b], deg[a], deg[b], cpl), a, b)
        mg = oa.mixing_gate(mix)
        for i in range(ef.nodes_number):
            o.apply_q1(mg, i)
        dt = time.perf_counter() - t0
    sweeps = sum(o.bp_iterations)
    eu = sweeps * 2 * ef.edges_number
    return {"value": eu / dt, "unit": "edge-updates/s", "cores": threads, "kind": "port",
            "cpu": cpu_model(), "host_cores_usable": usable_cores(), "dtype": "c128",
            "sample": "one Trotter step of the same instance on the oracle (NumPy complex128, %d BLAS threads), "
                      "started from the GPU state halfway through the run (full bond spectra): regauge (%d BP "
                      "sweeps) + %d simple updates + %d one-site gates, %.1f s" % (
                          threads, sweeps, ef.edges_number, ef.nodes_number, dt),
            "seconds": dt, "gates_per_s": ef.edges_number / dt}


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
    threads = usable_cores()
    return {"value": 2 ** n / dt, "unit": "amplitude-layers/s", "cores": threads, "kind": "port", "cpu": cpu_model(),
            "sample": "one annealing layer (%d coupling gates, %d mixing gates) of a %d-qubit random 3-regular "
                      "instance on the oracle's state vector (NumPy complex128 tensordot), %.1f s"
                      % (ef.edges_number, n, n, dt),
            "seconds": dt}


if __name__ == "__main__":
    main()
