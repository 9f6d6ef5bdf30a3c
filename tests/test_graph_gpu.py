"""Multi-sweep BP through a captured HIP graph (dq_bp_sweeps): every replay
equals the same sweeps launched one by one (identical kernels, so the
results are bitwise equal), with eager work and synchronisations between the
replays; odd and even sweep counts; the matrix-core (chi=16, complex64) and
generic (chi=4, complex128, asynchronous schedule) node kernels."""
import numpy as np
import pytest

from tests.conftest import gpu_available

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not gpu_available(), reason="needs a ROCm GPU")]

import torch  # noqa: E402

from dequantizer_amd import _lib  # noqa: E402

DEV = "cuda"


def _handle(n, chi, sync):
    import networkx as nx

    import dequantizer_amd as dq
    from dequantizer_amd.engine import Topology
    g = dq.TensorGraph()
    for _ in range(n):
        g.add_node(2)
    for a, b in nx.random_regular_graph(3, n, seed=2).edges:
        g.add_edge((a, b), chi)
    topo = Topology([list(g.get_traversal_iterator(ordering="bfs"))])
    return topo, topo.handle(chi, 2, sync)


@pytest.mark.parametrize("chi,dtype,sync", [(16, torch.complex64, True), (4, torch.complex128, False)])
@pytest.mark.parametrize("n_sweeps", [3, 10])
def test_sweep_graph_replays_match_eager_sweeps(chi, dtype, sync, n_sweeps):
    topo, h = _handle(24, chi, sync)
    code = _lib.DQ_C64 if dtype == torch.complex64 else _lib.DQ_C128
    L, st = _lib.lib(), _lib.stream_handle()
    gen = torch.Generator(device=DEV).manual_seed(1)
    T = torch.randn(h.tensor_elems, dtype=dtype, device=DEV, generator=gen)
    nm = 2 * topo.n_edges * chi * chi
    a = torch.randn((2 * topo.n_edges, chi, chi), dtype=dtype, device=DEV, generator=gen)
    m0 = (a @ a.conj().transpose(1, 2)).reshape(-1)
    bufs = [torch.zeros(nm, dtype=dtype, device=DEV) for _ in range(4)]

    def eager(T):
        cur = [m0.clone(), m0.clone()]
        for _ in range(n_sweeps):
            nxt = [torch.zeros_like(m0), torch.zeros_like(m0)]
            _lib.check(L.dq_bp_sweep(h.ptr, code, _lib.ptr(T), _lib.ptr(cur[0]), _lib.ptr(cur[1]), _lib.ptr(nxt[0]),
                                     _lib.ptr(nxt[1]), None, st))
            cur = nxt
        return cur

    for replay in range(3):
        if replay:
            T.mul_(1.0 + 0.1 * replay)  # eager work on the graph's input between replays
        torch.cuda.synchronize()
        bufs[0].copy_(m0)
        bufs[1].copy_(m0)
        bufs[2].zero_()
        bufs[3].zero_()
        _lib.check(L.dq_bp_sweeps(h.ptr, code, _lib.ptr(T), _lib.ptr(bufs[0]), _lib.ptr(bufs[1]), _lib.ptr(bufs[2]),
                                  _lib.ptr(bufs[3]), n_sweeps, None, st))
        got = (bufs[2], bufs[3]) if n_sweeps % 2 else (bufs[0], bufs[1])
        ref = eager(T)
        torch.cuda.synchronize()
        assert torch.equal(got[0], ref[0]) and torch.equal(got[1], ref[1]), (replay, n_sweeps)


def test_sweep_graph_rejects_aliased_buffers():
    topo, h = _handle(12, 4, True)
    L, st = _lib.lib(), _lib.stream_handle()
    T = torch.zeros(h.tensor_elems, dtype=torch.complex128, device=DEV)
    m = torch.zeros(2 * topo.n_edges * 16, dtype=torch.complex128, device=DEV)
    other = torch.zeros_like(m)
    assert L.dq_bp_sweeps(h.ptr, _lib.DQ_C128, _lib.ptr(T), _lib.ptr(m), _lib.ptr(other), _lib.ptr(m),
                          _lib.ptr(other), 2, None, st) != 0


def test_matrix_core_sweep_beyond_the_gathered_metadata():
    """42,000 nodes: more than the 160 nodes per block (x 256 blocks) whose
    graph metadata the complex64 matrix-core node kernel gathers into LDS, so
    the later nodes of each block read it through scalar loads.  One sweep
    against the complex128 matrix-core kernel (metadata through scalar loads
    only) on the same inputs: FP32 accumulation, relative 2e-5 per message."""
    n, chi = 42000, 16
    print(f"BP sweep on {n} nodes", flush=True)
    topo, h = _handle(n, chi, True)
    L, st = _lib.lib(), _lib.stream_handle()
    gen = torch.Generator(device=DEV).manual_seed(5)
    T = torch.randn(h.tensor_elems, dtype=torch.complex128, device=DEV, generator=gen)
    a = torch.randn((2 * topo.n_edges, chi, chi), dtype=torch.complex128, device=DEV, generator=gen)
    m0 = (a @ a.conj().transpose(1, 2)).reshape(-1)
    del a
    out = {}
    for code, dtype in ((_lib.DQ_C64, torch.complex64), (_lib.DQ_C128, torch.complex128)):
        Td, md = T.to(dtype), m0.to(dtype)
        o = [torch.zeros_like(md), torch.zeros_like(md)]
        _lib.check(L.dq_bp_sweep(h.ptr, code, _lib.ptr(Td), _lib.ptr(md), _lib.ptr(md), _lib.ptr(o[0]),
                                 _lib.ptr(o[1]), None, st))
        out[code] = o
        del Td, md
    for k in range(2):
        got = out[_lib.DQ_C64][k].view(-1, chi * chi).to(torch.complex128)
        ref = out[_lib.DQ_C128][k].view(-1, chi * chi)
        rel = (got - ref).abs().pow(2).sum(1).sqrt() / ref.abs().pow(2).sum(1).sqrt()
        assert float(rel.max()) < 2e-5, (k, float(rel.max()))


def _set_schedule(h, sched):
    import ctypes
    L = _lib.lib()
    n_levels, level_ptr, level_nodes, node_src, edge_src = sched
    ip, ucp = ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_ubyte)
    lp = np.ascontiguousarray(level_ptr, dtype=np.int32)
    ln = np.ascontiguousarray(level_nodes, dtype=np.int32)
    ns = np.ascontiguousarray(node_src, dtype=np.uint8)
    es = np.ascontiguousarray(edge_src, dtype=np.uint8)
    _lib.check(L.dq_bp_set_schedule(h.ptr, n_levels, lp.ctypes.data_as(ip), ln.ctypes.data_as(ip),
                                    ns.ctypes.data_as(ucp), es.ctypes.data_as(ucp)))


@pytest.mark.parametrize("chi,dtype", [(16, torch.complex64), (4, torch.complex128)])
def test_sweep_graph_after_schedule_change(chi, dtype):
    """dq_bp_sweeps captured under the synchronous schedule, then the same
    handle switched to the asynchronous (levelled) schedule and dq_bp_sweeps
    called again with the same buffers: the second call must run the new
    schedule (equal to eager dq_bp_sweep under it), not replay the graph
    captured for the old one.  Then back to the synchronous schedule."""
    import dequantizer_amd as dq
    import networkx as nx
    from dequantizer_amd.engine import Topology
    g = dq.TensorGraph()
    for _ in range(24):
        g.add_node(2)
    for a, b in nx.random_regular_graph(3, 24, seed=2).edges:
        g.add_edge((a, b), chi)
    topo = Topology([list(g.get_traversal_iterator(ordering="dfs"))])
    h = topo.handle(chi, 2, True)
    code = _lib.DQ_C64 if dtype == torch.complex64 else _lib.DQ_C128
    L, st = _lib.lib(), _lib.stream_handle()
    gen = torch.Generator(device=DEV).manual_seed(3)
    T = torch.randn(h.tensor_elems, dtype=dtype, device=DEV, generator=gen)
    nm = 2 * topo.n_edges * chi * chi
    a = torch.randn((2 * topo.n_edges, chi, chi), dtype=dtype, device=DEV, generator=gen)
    m0 = (a @ a.conj().transpose(1, 2)).reshape(-1)
    bufs = [torch.zeros(nm, dtype=dtype, device=DEV) for _ in range(4)]
    n_sweeps = 3

    def eager():
        cur = [m0.clone(), m0.clone()]
        for _ in range(n_sweeps):
            nxt = [torch.zeros_like(m0), torch.zeros_like(m0)]
            _lib.check(L.dq_bp_sweep(h.ptr, code, _lib.ptr(T), _lib.ptr(cur[0]), _lib.ptr(cur[1]), _lib.ptr(nxt[0]),
                                     _lib.ptr(nxt[1]), None, st))
            cur = nxt
        return cur

    def graph_run():
        bufs[0].copy_(m0)
        bufs[1].copy_(m0)
        bufs[2].zero_()
        bufs[3].zero_()
        _lib.check(L.dq_bp_sweeps(h.ptr, code, _lib.ptr(T), _lib.ptr(bufs[0]), _lib.ptr(bufs[1]), _lib.ptr(bufs[2]),
                                  _lib.ptr(bufs[3]), n_sweeps, None, st))
        return bufs[2].clone(), bufs[3].clone()

    results = {}
    for name, sched in (("sync", topo.schedule(True)), ("async", topo.schedule(False)), ("sync2", topo.schedule(True))):
        _set_schedule(h, sched)
        got = graph_run()
        ref = eager()
        torch.cuda.synchronize()
        assert torch.equal(got[0], ref[0]) and torch.equal(got[1], ref[1]), name
        results[name] = got
    # the two schedules give different iterates (the test would be vacuous otherwise)
    assert not torch.equal(results["sync"][0], results["async"][0])
    assert torch.equal(results["sync"][0], results["sync2"][0])
