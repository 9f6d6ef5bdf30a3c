"""Autotuner: the report names the variants it timed, and every variant it
may choose is correct: each node-update variant against the oracle, each
simple-update storage mode against the other (same arithmetic: bitwise)."""
import numpy as np
import pytest

from tests.conftest import gpu_available

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not gpu_available(), reason="needs a ROCm GPU")]

import torch  # noqa: E402

import dequantizer_amd as dq  # noqa: E402
from dequantizer_amd import _lib  # noqa: E402
from oracle import mappings as om  # noqa: E402
from tests.helpers import host, random_messages, random_tensors, to_oracle_messages  # noqa: E402
from tests.test_maps_gpu import regular_pair  # noqa: E402


def test_autotune_report_and_choice():
    from dequantizer_amd.engine import Topology
    gp, _ = regular_pair(24, 16, 3)
    topo = Topology([list(gp.get_traversal_iterator())])
    h = topo.handle(16, 2, True)
    rep = _lib.autotune(h.ptr, _lib.DQ_C64, 2, force=True)
    assert "node update" in rep and "chose" in rep and "simple update" in rep, rep


@pytest.mark.parametrize("variant", [0, 1, 2, 3])
def test_every_node_update_variant_matches_oracle(variant):
    gp, go = regular_pair(24, 16, 11)
    rng = np.random.default_rng(21)
    tensors = random_tensors(gp, rng)
    messages = random_messages(gp, rng)
    trav_p, trav_o = list(gp.get_traversal_iterator()), go.traversal()
    ref = om.belief_propagation_map(trav_o, True)(tensors, to_oracle_messages(messages))
    _lib.check(_lib.lib().dq_tune_override(0, _lib.DQ_C64, 16, 2, 2, variant))
    try:
        out = dq.get_belief_propagation_map(trav_p, True)({k: v.astype(np.complex64) for k, v in tensors.items()},
                                                          {k: v.astype(np.complex64) for k, v in messages.items()})
        worst = max(np.linalg.norm(host(out[dq.MessageID(k.src, k.dst)]) - v) / np.linalg.norm(v)
                    for k, v in ref.items())
    finally:
        _lib.check(_lib.lib().dq_tune_override(0, _lib.DQ_C64, 16, 2, 2, -1))
    assert worst < 2e-5, (variant, worst)


@pytest.mark.parametrize("dtype", [torch.complex64, torch.complex128])
def test_simple_update_modes_agree(dtype):
    from dequantizer_amd.annealing import Annealer, edge_colored, get_scheduler, linear_schedule, random_regular
    code = _lib.DQ_C64 if dtype == torch.complex64 else _lib.DQ_C128
    ef = edge_colored(random_regular(40, 3, seed=4, couplings="pm1", field_std=0.1))
    out = []
    for mode in (2, 1):
        _lib.check(_lib.lib().dq_tune_override(1, code, 16, 2, 2, mode))
        try:
            ann = Annealer([ef], [dq.PRNGKey(3)], 16, 1, 1e-5, 100, True, "bfs", dtype=dtype, chi=16, fixed_sweeps=4)
            for mix, cpl in get_scheduler(0.1, linear_schedule, 200)[:6]:
                ann.step(mix, cpl)
            out.append(ann.finish(two_site=True)[0].density_matrices)
        finally:
            _lib.check(_lib.lib().dq_tune_override(1, code, 16, 2, 2, -1))
    assert np.array_equal(out[0], out[1])
