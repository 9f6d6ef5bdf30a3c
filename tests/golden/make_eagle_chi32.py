"""Writes eagle_chi32.json: <Z_62> of the 127-qubit heavy-hex kicked-Ising
dynamics (RX(theta) on every qubit, then RZZ(-pi/2) on every edge) computed by
the oracle BP emulator in complex128 at max_chi = 32, the workload of
examples/ibm_eagle_processor.py:56-145 on the finite lattice.

Two gate orders: the reference's (traverser edge order, one regauge per full
layer) and three edge-colour classes with a regauge after each class.  The
bonds double every step, so from step 6 on the simple updates truncate 64
values to 32.  Run from the repository root:
    python tests/golden/make_eagle_chi32.py [steps]
"""
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from dequantizer_amd import drivers  # noqa: E402  (host-side colouring only; no GPU)
import dequantizer_amd as dq  # noqa: E402
from oracle import emulator as oe  # noqa: E402
from oracle import graph as og  # noqa: E402

THETA = 0.8
CHI = 32


def run(steps, colour_layers):
    lattice = og.heavy_hex_eagle()
    trav = lattice.traversal()
    per = 10 ** 9 if colour_layers else lattice.edges_number
    emu = oe.BPQuantumEmulator(lattice, og.make_key(42), CHI, per, 1e-5, None, False, "dfs")
    edges = [el.id for el in trav if isinstance(el, og.Edge)]
    classes = []
    if colour_layers:
        colours = drivers.edge_colouring(dq.get_heavy_hex_ibm_eagle_lattice())
        classes = [[e for e in edges if colours[e] == c] for c in range(3)]
    mixing = drivers.rx(THETA)

    def z():
        d = emu.dens_q1(62)
        return float((d[0, 0] - d[1, 1]).real)

    out = [z()]
    for k in range(steps):
        t0 = time.time()
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
        out.append(z())
        print(f"colour={colour_layers} step {k + 1}: <Z62> = {out[-1]:.12f}, max bond "
              f"{max(e.dimension for e in lattice.edges.values())}, {time.time() - t0:.1f} s", flush=True)
    return out, list(emu.bp_iterations)


if __name__ == "__main__":
    steps = int(sys.argv[1]) if len(sys.argv) > 1 else 7
    res = {"theta": THETA, "max_chi": CHI, "accuracy": 1e-5, "key": 42, "traversal": "dfs", "steps": steps}
    for mode, colour in (("edge_order", False), ("colour_layers", True)):
        z, iters = run(steps, colour)
        res[mode] = {"z62": z, "bp_iterations": iters}
    with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "eagle_chi32.json"), "w") as f:
        json.dump(res, f, indent=1)
