"""dq_vidal_distance_noalloc (scratch in a fixed device array, no allocation)
against dq_vidal_distance (the same kernels on the handle's workspace), through
the C ABI: two calls in a row on each kernel path, the second on a smaller
graph, results bitwise equal; a graph whose scratch would exceed the fixed
array is rejected before any launch."""
import numpy as np
import pytest

from tests.conftest import gpu_available

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not gpu_available(), reason="needs a ROCm GPU")]

import torch  # noqa: E402

from dequantizer_amd import _lib  # noqa: E402

DEV = "cuda"


def _graph(n, chi, seed):
    import networkx as nx

    import dequantizer_amd as dq
    from dequantizer_amd.engine import Topology
    g = dq.TensorGraph()
    for _ in range(n):
        g.add_node(2)
    for a, b in nx.random_regular_graph(3, n, seed=seed).edges:
        g.add_edge((a, b), chi)
    topo = Topology([list(g.get_traversal_iterator(ordering="bfs"))])
    return topo, topo.handle(chi, 2, True)


def _inputs(topo, h, chi, dtype, seed):
    gen = torch.Generator(device=DEV).manual_seed(seed)
    gam = torch.randn(h.tensor_elems, dtype=dtype, device=DEV, generator=gen)
    lam = torch.rand((topo.n_edges, chi), dtype=torch.float64, device=DEV, generator=gen) + 0.05
    lam = lam / lam.norm(dim=1, keepdim=True)
    return gam, lam.reshape(-1).contiguous()


def _both(h, code, gam, lam):
    L, st = _lib.lib(), _lib.stream_handle()
    ref = torch.full((1,), -1.0, dtype=torch.float64, device=DEV)
    got = torch.full((1,), -2.0, dtype=torch.float64, device=DEV)
    _lib.check(L.dq_vidal_distance(h.ptr, code, _lib.ptr(gam), _lib.ptr(lam), 1e-10, _lib.ptr(ref), None, st))
    _lib.check(L.dq_vidal_distance_noalloc(h.ptr, code, _lib.ptr(gam), _lib.ptr(lam), 1e-10, _lib.ptr(got), None,
                                           st))
    torch.cuda.synchronize()
    return float(ref.item()), float(got.item())


@pytest.mark.parametrize("chi,dtype,sizes", [
    (16, torch.complex64, (1000, 200)),  # matrix-core Gram path
    (4, torch.complex128, (60, 20)),  # generic per-node path
])
def test_noalloc_vidal_distance_equals_the_allocating_entry_point(chi, dtype, sizes):
    code = _lib.DQ_C64 if dtype == torch.complex64 else _lib.DQ_C128
    for k, n in enumerate(sizes):
        print(f"vidal distance no-alloc: {n} nodes, chi={chi}, {dtype}", flush=True)
        topo, h = _graph(n, chi, seed=3 + k)
        gam, lam = _inputs(topo, h, chi, dtype, seed=10 + k)
        ref, got = _both(h, code, gam, lam)
        assert ref > 0 and np.isfinite(ref)
        assert got == ref, (n, got, ref)


def test_noalloc_vidal_distance_rejects_graphs_beyond_the_scratch():
    """12,000 nodes (18,000 edges): 2 * 18000 * 2056 + 256 bytes of Grams and
    per-bond distances exceed the fixed 64 MiB array."""
    chi = 16
    topo, h = _graph(12000, chi, seed=4)
    gam, lam = _inputs(topo, h, chi, torch.complex64, seed=1)
    out = torch.zeros(1, dtype=torch.float64, device=DEV)
    L = _lib.lib()
    rc = L.dq_vidal_distance_noalloc(h.ptr, _lib.DQ_C64, _lib.ptr(gam), _lib.ptr(lam), 1e-10, _lib.ptr(out), None,
                                     _lib.stream_handle())
    assert rc == 1  # DQ_ERR_INVALID
    assert b"scratch" in L.dq_last_error()
    torch.cuda.synchronize()
    assert float(out.item()) == 0.0  # nothing was launched
