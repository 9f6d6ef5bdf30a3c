"""Tensor-valued methods of the dictionary API of ``TensorGraph``.

``TensorGraph.apply1``, ``apply2``, ``apply2_to_vidal_canonical`` and
``truncate`` (dequantizer/src/graph/tensor_graph.py:235-446) work on dicts of
per-node tensors, each with its own bond shape.  Here they run on the GPU:

* ``apply1`` and the gate halves of ``apply2`` go through the single-tensor
  kernels (``dq_tensor_apply1``, ``dq_tensor_apply_half``);
* ``simple_update`` runs the batched simple-update kernel on a local handle:
  the gate's two nodes plus one degree-1 stub per outer bond, every bond
  padded to the largest bond or to the full singular spectrum, whichever is
  larger, so the kernel keeps every singular value; the reference's shapes
  (its spectrum length min(m1, m2, chi_e r) and its rank rule) are applied
  when the result is sliced back;
* ``truncate`` reads the lambdas to the host for the rank, as the reference
  does (``int(rank)``), and slices on the device.

Returned tensors and lambdas are torch tensors on the GPU (complex64 inputs
stay complex64, everything else becomes complex128; lambdas float64).
"""
import math
from types import SimpleNamespace
from typing import Dict, Optional

import numpy as np
import torch

from . import _lib
from .engine import Handle

DEVICE = "cuda"

_LOCAL_HANDLES: Dict[tuple, Handle] = {}


# ---------------------------------------------------------------- helpers
def _dev(x, dtype: Optional[torch.dtype] = None) -> torch.Tensor:
    """Device complex tensor (complex64 kept, anything else -> complex128)."""
    t = x if isinstance(x, torch.Tensor) else torch.from_numpy(np.array(x, copy=True, order="C"))
    if dtype is None:
        dtype = torch.complex64 if t.dtype == torch.complex64 else torch.complex128
    return t.to(device=DEVICE, dtype=dtype).contiguous()


def _dev_real(x) -> torch.Tensor:
    t = x if isinstance(x, torch.Tensor) else torch.from_numpy(np.array(x, copy=True, order="C"))
    if t.is_complex():
        t = t.real
    return t.to(device=DEVICE, dtype=torch.float64).contiguous()


def _host(x) -> np.ndarray:
    if isinstance(x, torch.Tensor):
        return x.detach().cpu().numpy()
    return np.asarray(x)


def _stream() -> int:
    return _lib.stream_handle()


def find_rank(lmbd, accuracy, rank: Optional[int]) -> int:
    """Kept bond rank, dequantizer/src/tensor_ops.py:7-17: without an
    accuracy, ``rank`` (all values if None); with one, the number of values
    minus the count of tail positions whose cumulative norm (summed from the
    smallest value), relative to the full norm, is below the accuracy."""
    lm = np.asarray(_host(lmbd), dtype=np.float64)
    if accuracy is None:
        return int(lm.shape[0]) if rank is None else int(rank)
    acc = float(_host(accuracy))
    total = math.sqrt(float(np.dot(lm, lm)))
    tail = np.sqrt(np.cumsum(lm[::-1] ** 2)) / total
    return int(lm.shape[0] - np.count_nonzero(tail < acc))


def _gate_halves(gate, eps: float):
    """Device operator-Schmidt split of one gate: (halves (2, p^4) complex128
    on the device, rank (read back: the result shapes depend on it), p)."""
    from .emulator import decompose_gates_device
    g = np.asarray(_host(gate), dtype=np.complex128)
    p = int(g.shape[0])
    d_g = torch.from_numpy(np.array(g.reshape(1, -1), copy=True)).to(DEVICE)
    halves = torch.empty((1, 2 * p ** 4), dtype=torch.complex128, device=DEVICE)
    ranks = torch.zeros(1, dtype=torch.int32, device=DEVICE)
    decompose_gates_device(d_g, p, eps, halves, ranks.data_ptr(), 1, _stream())
    return halves.view(2, p ** 4), int(ranks.item()), p


# ------------------------------------------------------ single-tensor ops
def apply1(tensor, gate) -> torch.Tensor:
    """tensor x_phys gate^T, then Frobenius normalisation
    (tensor_graph.py:268-275)."""
    t = _dev(tensor)
    p = int(t.shape[-1])
    g = _dev(gate, torch.complex128)
    if tuple(g.shape) != (p, p):
        raise ValueError(f"one-site gate of shape {tuple(g.shape)} does not act on a physical dimension {p}")
    out = torch.empty_like(t)
    ws = torch.empty(_lib.DQ_TENSOR_WS, dtype=torch.float64, device=DEVICE)
    L = _lib.lib()
    _lib.check(L.dq_tensor_apply1(_lib.dtype_code(t.dtype), _lib.ptr(t), t.numel() // p, p, _lib.ptr(g),
                                  _lib.ptr(out), _lib.ptr(ws), _stream()))
    return out


def apply_half_gate(tensor, half: torch.Tensor, rank: int, index: int) -> torch.Tensor:
    """Gate half (rank, out, in) on the physical axis, rank fused into bond
    ``index`` as its fast part, then normalisation (tensor_ops.py:221-237)."""
    t = _dev(tensor)
    dims = list(t.shape)
    p = dims[-1]
    lead = int(np.prod(dims[:index], dtype=np.int64))
    trail = int(np.prod(dims[index + 1:-1], dtype=np.int64))
    shape = dims.copy()
    shape[index] *= rank
    out = torch.empty(shape, dtype=t.dtype, device=DEVICE)
    ws = torch.empty(_lib.DQ_TENSOR_WS, dtype=torch.float64, device=DEVICE)
    L = _lib.lib()
    _lib.check(L.dq_tensor_apply_half(_lib.dtype_code(t.dtype), _lib.ptr(t), lead, dims[index], trail, p,
                                      _lib.ptr(half), rank, _lib.ptr(out), _lib.ptr(ws), _stream()))
    return out


def apply_gate_halves(t1, t2, i1: int, i2: int, gate, threshold):
    """Non-truncating two-site gate (tensor_graph.py:387-446): both halves
    applied, the shared bond grows by the gate rank.  Returns (t1, t2, rank)."""
    halves, r, p = _gate_halves(gate, float(_host(threshold)))
    n = r * p * p
    return apply_half_gate(t1, halves[0, :n].contiguous(), r, i1), apply_half_gate(
        t2, halves[1, :n].contiguous(), r, i2), r


# ---------------------------------------------------------- simple update
def _local_handle(d1: int, d2: int, i1: int, i2: int, chi: int, phys: int) -> Handle:
    """Handle of a two-node graph (shared edge 0, node 0 on side 0) with one
    degree-1 stub per outer bond; bonds in the nodes' own orders."""
    key = (d1, d2, i1, i2, chi, phys)
    h = _LOCAL_HANDLES.get(key)
    if h is not None:
        return h
    deg = [d1, d2]
    adj_edge, adj_side, edge_node = [], [], [(0, 1)]
    stubs = []
    for node, d, shared in ((0, d1, i1), (1, d2, i2)):
        for k in range(d):
            if k == shared:
                adj_edge.append(0)
                adj_side.append(node)
            else:
                e = len(edge_node)
                stub = 2 + len(stubs)
                stubs.append(e)
                edge_node.append((node, stub))
                adj_edge.append(e)
                adj_side.append(0)
    for e in stubs:
        deg.append(1)
        adj_edge.append(e)
        adj_side.append(1)
    n_nodes = len(deg)
    topo = SimpleNamespace(deg=np.asarray(deg, np.int32), adj_edge=np.asarray(adj_edge, np.int32),
                           adj_side=np.asarray(adj_side, np.int32),
                           edge_node=np.asarray(edge_node, np.int32).reshape(-1, 2),
                           node_inst=np.zeros(n_nodes, np.int32), n_nodes=n_nodes, n_edges=len(edge_node), n_inst=1)
    h = Handle(topo, chi, phys, True)
    h.slot_edges = [adj_edge[:d1], adj_edge[d1:d1 + d2]]
    _LOCAL_HANDLES[key] = h
    return h


def simple_update(graph, edge, n1, n2, tensors, lambdas, gate, threshold, accuracy, max_rank=None):
    """Two-site gate on a Vidal-canonical pair with truncation
    (tensor_graph.py:294-375 -> simple_update, tensor_ops.py:318-389).
    Returns (Gamma_1, Gamma_2, lambda (unnormalised), error, entropy)."""
    eps = float(_host(threshold))
    halves, r, _ = _gate_halves(gate, eps)
    i1, i2 = n1.indices[edge.id], n2.indices[edge.id]
    lam1 = [_dev_real(lambdas[e.id]) for e in n1.neighbors]
    lam2 = [_dev_real(lambdas[e.id]) for e in n2.neighbors]
    t1, t2 = _dev(tensors[n1.id]), _dev(tensors[n2.id])
    dtype = torch.complex64 if (t1.dtype == torch.complex64 and t2.dtype == torch.complex64) else torch.complex128
    t1, t2 = t1.to(dtype), t2.to(dtype)
    p = int(t1.shape[-1])
    if int(t2.shape[-1]) != p:
        raise ValueError("both nodes of a simple update must have the same physical dimension")
    dims1, dims2 = [int(x) for x in t1.shape[:-1]], [int(x) for x in t2.shape[:-1]]
    de = dims1[i1]
    if dims2[i2] != de or lam1[i1].shape[0] != de:
        raise ValueError(f"bond dimensions of edge {edge.id} disagree: {dims1[i1]}, {dims2[i2]}, "
                         f"{lam1[i1].shape[0]}")
    m1 = p * int(np.prod([x for k, x in enumerate(dims1) if k != i1], dtype=np.int64))
    m2 = p * int(np.prod([x for k, x in enumerate(dims2) if k != i2], dtype=np.int64))
    n_s = min(m1, m2, de * r)  # length of the reference's singular spectrum
    if accuracy is None:
        rank = n_s if max_rank is None else min(int(max_rank), n_s)
        # the kernel computes all chi * r singular values and keeps the top
        # min(max_rank, chi): padding to the kept rank is enough
        chi = max(dims1 + dims2 + [rank])
    else:
        chi = max(dims1 + dims2 + [n_s])
    h = _local_handle(len(dims1), len(dims2), i1, i2, chi, p)
    T = torch.zeros(h.tensor_elems, dtype=dtype, device=DEVICE)
    h.node_view(T, 0)[tuple(slice(0, x) for x in dims1)] = t1
    h.node_view(T, 1)[tuple(slice(0, x) for x in dims2)] = t2
    lam = torch.zeros((h.topo.n_edges, chi), dtype=torch.float64, device=DEVICE)
    for node_lams, slots in ((lam1, h.slot_edges[0]), (lam2, h.slot_edges[1])):
        for lm, e in zip(node_lams, slots):
            lam[e, : lm.shape[0]] = lm
    if accuracy is None:
        kernel_rank = max(1, rank)
    else:
        kernel_rank = chi
    meta = torch.tensor([[0, 0, r, 0]], dtype=torch.int32, device=DEVICE)
    err = torch.zeros(1, dtype=torch.float64, device=DEVICE)
    ent = torch.zeros(1, dtype=torch.float64, device=DEVICE)
    L = _lib.lib()
    _lib.check(L.dq_simple_update(h.ptr, _lib.dtype_code(dtype), _lib.ptr(T), _lib.ptr(lam), _lib.ptr(meta), 1,
                                  _lib.ptr(halves.contiguous()), r, eps, kernel_rank, _lib.ptr(err), _lib.ptr(ent),
                                  _stream()))
    spectrum = lam[0, :n_s]
    if accuracy is not None:
        s = spectrum.cpu().numpy()
        rank = find_rank(s, accuracy, max_rank)
        nrm = math.sqrt(float(np.dot(s, s)))
        err = torch.tensor(math.sqrt(float(np.dot(s[rank:], s[rank:]))) / nrm, dtype=torch.float64, device=DEVICE)
    else:
        err = err[0]
    new1, new2 = dims1.copy(), dims2.copy()
    new1[i1] = new2[i2] = rank
    g1 = h.node_view(T, 0)[tuple(slice(0, x) for x in new1)].clone()
    g2 = h.node_view(T, 1)[tuple(slice(0, x) for x in new2)].clone()
    return g1, g2, spectrum[:rank].clone(), err, ent[0]


# --------------------------------------------------------------- truncate
def truncate(graph, tensors, core_edge_tensors, accuracy, max_rank=None):
    """tensor_graph.py:235-258: rank per edge from its lambdas, edge and node
    dimensions updated in the graph, tensors and lambdas sliced (new dicts)."""
    new_lambdas = {}
    for eid, lm in core_edge_tensors.items():
        lm_d = _dev_real(lm)
        rank = find_rank(lm_d.cpu().numpy(), accuracy, max_rank)
        if rank > lm_d.shape[0]:
            raise ValueError(f"rank {rank} exceeds the {lm_d.shape[0]} values on edge {eid}")
        new_lambdas[eid] = lm_d[:rank].clone()
        graph.edges[eid].dimension = int(rank)
    new_tensors = {}
    for nid, node in graph.nodes.items():
        shape = [nb.dimension for nb in node.neighbors]
        node.bond_shape = tuple(shape)
        t = _dev(tensors[nid])
        if any(s > x for s, x in zip(shape, t.shape)):
            raise ValueError(f"node {nid}: bond shape {tuple(shape)} exceeds the tensor shape {tuple(t.shape)}")
        new_tensors[nid] = t[tuple(slice(0, s) for s in shape)].clone()
    return new_tensors, new_lambdas
