"""Writes rr3_n1000_chi16.npz: the first steps of the 1000-spin random
3-regular annealing run at chi = 16 computed by the oracle BP emulator in
complex128, the workload of
examples/quantum_annealing/quantum_annealing/src/quantum_annealer.py:90-152:
instance seed 7 (+-1 couplings, N(0, 0.1^2) fields), pairs sorted by an edge
colouring, dt = 0.1 on a 200-step linear schedule, a regauge before every
layer of exactly 10 synchronous BP sweeps (bfs order) followed by one Vidal
gauge.  The first coupling layer is the identity (s = 0), so the bonds reach
16 at step 5 and are truncated (32 values to 16) from step 6 on.

Stored: the instance (pairs, couplings, fields) so that the GPU test feeds
identical inputs, and after the last step the per-spin <Z>, the per-edge
<ZZ>, the BP energy, the truncation error of every gate of the last
TAIL_STEPS steps and every regauge's Vidal distance.  The committed file was
made with 80 steps (truncation errors up to 5e-6 in its last steps).  Run from
the repository root:
    python tests/golden/make_rr3_n1000.py [steps]
"""
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from dequantizer_amd import PRNGKey  # noqa: E402  (host-side key and instance generators; no GPU)
from dequantizer_amd.annealing import edge_colored, random_regular  # noqa: E402
from oracle import annealing as oa  # noqa: E402
from oracle import emulator as oe  # noqa: E402
from oracle import graph as og  # noqa: E402

N, SEED, CHI, DT, SCHEDULE, SWEEPS = 1000, 7, 16, 0.1, 200, 10
TAIL_STEPS = 10  # steps whose per-gate truncation errors are stored


def run(steps):
    ef = edge_colored(random_regular(N, 3, seed=SEED, couplings="pm1", field_std=0.1))
    pairs, j, h = ef.coupled_spin_pairs, ef.coupling_amplitudes, ef.fields
    g = og.graph_from_pairs(N, pairs)
    emu = oe.BPQuantumEmulator(g, PRNGKey(SEED), CHI, ef.edges_number, 1e-5, 100, True, "bfs",
                               fixed_sweeps=SWEEPS)
    deg = [g.nodes[i].degree for i in range(N)]
    for i in range(N):
        emu.apply_q1(oa.HADAMARD, i)
    sched = oa.scheduler(DT, lambda t: 1 - t, SCHEDULE)
    for k, (mix, cpl) in enumerate(sched[:steps]):
        t0 = time.time()
        for (a, b), jj in zip(pairs, j):
            a, b = int(a), int(b)
            emu.apply_q2(oa.coupling_gate(jj, h[a], h[b], deg[a], deg[b], cpl), a, b)
        mg = oa.mixing_gate(mix)
        for i in range(N):
            emu.apply_q1(mg, i)
        print(f"step {k + 1}: max bond {max(e.dimension for e in g.edges.values())}, max truncation error "
              f"{max(emu.truncation_errors[-len(pairs):]):.2e}, {time.time() - t0:.1f} s", flush=True)
    rho1 = np.array([emu.dens_q1(i) for i in range(N)])
    rho2 = np.array([emu.dens_q2(int(a), int(b)) for a, b in pairs])
    z = (rho1[:, 0, 0] - rho1[:, 1, 1]).real
    zz = (rho2[:, 0, 0] - rho2[:, 1, 1] - rho2[:, 2, 2] + rho2[:, 3, 3]).real
    tail = TAIL_STEPS * len(pairs)
    return dict(pairs=np.asarray(pairs, dtype=np.int32), couplings=j, fields=h, steps=steps, z=z, zz=zz,
                energy=oa.bp_energy(rho1, rho2, h, j),
                truncation_errors_tail=np.asarray(emu.truncation_errors)[-tail:], tail_steps=TAIL_STEPS,
                vidal_distances=np.asarray(emu.vidal_distances_after_regauging, dtype=np.float64))


if __name__ == "__main__":
    steps = int(sys.argv[1]) if len(sys.argv) > 1 else 80
    out = run(steps)
    np.savez_compressed(os.path.join(os.path.dirname(os.path.abspath(__file__)), "rr3_n1000_chi16.npz"), **out)
    print("energy", out["energy"])
