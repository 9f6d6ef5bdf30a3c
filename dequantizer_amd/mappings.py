"""Whole-graph maps with the reference signatures, running on libdq.

Each ``get_*_map(traverser, ...)`` mirrors its counterpart in
dequantizer/src/mappings/ and returns a function over dicts keyed by
NodeID / MessageID / EdgeID.  Inputs may be plain dicts of arrays (NumPy or
torch) or the packed, device-resident dict views these maps return; packed
inputs on the same traverser are used in place, without host traffic.
Values are torch tensors on the GPU.  Arithmetic follows the input dtype
(complex64 -> complex64 kernels, otherwise complex128); decompositions and
lambdas are FP64.
"""
from collections.abc import Mapping
from typing import Dict, Iterable, Optional, Tuple

import numpy as np
import torch

from . import _lib
from .engine import Handle, Topology, topology_for
from .graph import Edge, MessageID, Node

DEVICE = "cuda"


# --------------------------------------------------------------- containers
class PackedTensors(Mapping):
    """Node tensors in padded device storage (view per node, sliced to its
    true bond dimensions)."""

    def __init__(self, handle: Handle, buf: torch.Tensor, dims: np.ndarray):
        self.handle, self.buf, self.dims = handle, buf, dims  # dims[e] = true bond dim

    @property
    def topo(self) -> Topology:
        return self.handle.topo

    def _index(self, node_id) -> int:
        return self.topo.node_index[0][node_id]

    def __getitem__(self, node_id):
        i = self._index(node_id)
        v = self.handle.node_view(self.buf, i)
        sl = tuple(slice(0, int(self.dims[e])) for e in self.topo.adj_edge[self.topo.adj_ptr[i]:self.topo.adj_ptr[i + 1]])
        return v[sl + (slice(None),)]

    def __iter__(self):
        return (n.id for n in self.topo.nodes)

    def __len__(self):
        return self.topo.n_nodes


class PackedMessages(Mapping):
    """All 4|E| messages: n2e[2e+s] = node(side s) -> edge e,
    e2n[2e+s] = edge e -> node(side s)."""

    def __init__(self, handle: Handle, n2e: torch.Tensor, e2n: torch.Tensor, dims: np.ndarray):
        self.handle, self.n2e, self.e2n, self.dims = handle, n2e, e2n, dims
        topo = handle.topo
        idx = {}
        for j, e in enumerate(topo.edges):
            for s, nb in enumerate(e.neighbors):
                idx[MessageID(nb.id, e.id)] = (0, 2 * j + s, j)
                idx[MessageID(e.id, nb.id)] = (1, 2 * j + s, j)
        self._idx = idx

    @property
    def topo(self) -> Topology:
        return self.handle.topo

    def __getitem__(self, key):
        which, slot, j = self._idx[key]
        chi = self.handle.chi
        buf = self.n2e if which == 0 else self.e2n
        d = int(self.dims[j])
        return buf.view(-1, chi, chi)[slot, :d, :d]

    def __iter__(self):
        return iter(self._idx)

    def __len__(self):
        return len(self._idx)


class PackedLambdas(Mapping):
    """Lambda vector per edge (FP64, padded storage n_edges x chi)."""

    def __init__(self, handle: Handle, lam: torch.Tensor, dims: np.ndarray):
        self.handle, self.lam, self.dims = handle, lam, dims

    @property
    def topo(self) -> Topology:
        return self.handle.topo

    def __getitem__(self, edge_id):
        j = self.topo.edge_index[0][edge_id]
        return self.lam.view(-1, self.handle.chi)[j, : int(self.dims[j])]

    def __iter__(self):
        return (e.id for e in self.topo.edges)

    def __len__(self):
        return self.topo.n_edges


# ------------------------------------------------------------------ packing
def _as_tensor(x) -> torch.Tensor:
    if isinstance(x, torch.Tensor):
        return x
    return torch.as_tensor(np.asarray(x))


def _result_dtype(*values) -> torch.dtype:
    for v in values:
        if isinstance(v, torch.Tensor):
            if v.dtype == torch.complex64:
                return torch.complex64
        elif np.asarray(v).dtype == np.complex64:
            return torch.complex64
    return torch.complex128


def _edge_dims_from_tensors(topo: Topology, tensors) -> np.ndarray:
    dims = np.zeros(topo.n_edges, dtype=np.int64)
    phys = None
    for i, node in enumerate(topo.nodes):
        shape = tuple(tensors[node.id].shape)
        d = int(topo.deg[i])
        if len(shape) != d + 1:
            raise ValueError(f"tensor of node {node.id} has {len(shape)} axes, expected {d + 1}")
        if phys is None:
            phys = shape[-1]
        elif phys != shape[-1]:
            raise NotImplementedError("nodes with different physical dimensions in one graph are not supported")
        for k in range(d):
            e = topo.adj_edge[topo.adj_ptr[i] + k]
            if dims[e] and dims[e] != shape[k]:
                raise ValueError(f"bond dimension mismatch on edge {topo.edges[e].id}")
            dims[e] = shape[k]
    return dims, (phys or 1)


def pack_tensors(topo: Topology, tensors, chi: Optional[int] = None, dtype=None) -> PackedTensors:
    """Pack a dict of node tensors into padded device storage."""
    if isinstance(tensors, PackedTensors) and tensors.topo is topo and (chi is None or chi == tensors.handle.chi) \
            and (dtype is None or tensors.buf.dtype == dtype):
        return tensors
    if isinstance(tensors, PackedTensors) and tensors.topo is topo and chi is not None and chi > tensors.handle.chi:
        return _repad(tensors, chi, dtype)
    dims, phys = _edge_dims_from_tensors(topo, tensors)
    chi = int(chi or max(1, int(dims.max()) if len(dims) else 1))
    dtype = dtype or _result_dtype(*[tensors[n.id] for n in topo.nodes])
    h = topo.handle(chi, phys)
    on_device = any(isinstance(tensors[n.id], torch.Tensor) and tensors[n.id].is_cuda for n in topo.nodes)
    if on_device:
        buf = torch.zeros(h.tensor_elems, dtype=dtype, device=DEVICE)
        for i, node in enumerate(topo.nodes):
            t = _as_tensor(tensors[node.id]).to(device=DEVICE, dtype=dtype)
            h.node_view(buf, i)[tuple(slice(0, s) for s in t.shape)] = t
    else:
        npdt = np.complex64 if dtype == torch.complex64 else np.complex128
        host = np.zeros(h.tensor_elems, dtype=npdt)
        for i, node in enumerate(topo.nodes):
            t = np.asarray(tensors[node.id] if not isinstance(tensors[node.id], torch.Tensor)
                           else tensors[node.id].numpy())
            d = int(topo.deg[i])
            off = int(h.offsets[i])
            view = host[off: off + h.node_size(i)].reshape((chi,) * d + (phys,))
            view[tuple(slice(0, s) for s in t.shape)] = t
        buf = torch.from_numpy(host).to(DEVICE)
    return PackedTensors(h, buf, dims)


def _same_layout(a: Handle, b: Handle) -> bool:
    return a.topo is b.topo and a.chi == b.chi and a.phys == b.phys


def _repad(pt: PackedTensors, chi: int, dtype=None) -> PackedTensors:
    L = _lib.lib()
    old = pt.handle
    h = old.topo.handle(chi, old.phys)
    buf = torch.empty(h.tensor_elems, dtype=pt.buf.dtype, device=DEVICE)
    _lib.check(L.dq_repad(old.ptr, h.ptr, _lib.dtype_code(pt.buf.dtype), _lib.ptr(pt.buf), _lib.ptr(buf), _stream()))
    if dtype is not None and dtype != buf.dtype:
        buf = buf.to(dtype)
    return PackedTensors(h, buf, pt.dims)


def pack_messages(h: Handle, messages, dims: np.ndarray, dtype) -> PackedMessages:
    topo = h.topo
    if isinstance(messages, PackedMessages) and _same_layout(messages.handle, h) and messages.n2e.dtype == dtype:
        return messages
    chi = h.chi
    n2e = torch.zeros(2 * topo.n_edges * chi * chi, dtype=dtype, device=DEVICE)
    e2n = torch.zeros_like(n2e)
    vn, ve = n2e.view(-1, chi, chi), e2n.view(-1, chi, chi)
    host_n = np.zeros((2 * topo.n_edges, chi, chi), dtype=np.complex64 if dtype == torch.complex64 else np.complex128)
    host_e = np.zeros_like(host_n)
    device_items = []
    for j, e in enumerate(topo.edges):
        for s, nb in enumerate(e.neighbors):
            for which, key in ((0, MessageID(nb.id, e.id)), (1, MessageID(e.id, nb.id))):
                m = messages[key]
                if isinstance(m, torch.Tensor) and m.is_cuda:
                    device_items.append((which, 2 * j + s, m))
                else:
                    m = np.asarray(m.numpy() if isinstance(m, torch.Tensor) else m)
                    (host_n if which == 0 else host_e)[2 * j + s, : m.shape[0], : m.shape[1]] = m
    if len(device_items) < 4 * topo.n_edges:
        vn.copy_(torch.from_numpy(host_n).to(DEVICE))
        ve.copy_(torch.from_numpy(host_e).to(DEVICE))
    for which, slot, m in device_items:
        (vn if which == 0 else ve)[slot, : m.shape[0], : m.shape[1]] = m.to(dtype)
    return PackedMessages(h, n2e, e2n, dims)


def pack_lambdas(h: Handle, lambdas, dims: np.ndarray) -> PackedLambdas:
    topo = h.topo
    if isinstance(lambdas, PackedLambdas) and _same_layout(lambdas.handle, h):
        return lambdas
    host = np.zeros((topo.n_edges, h.chi), dtype=np.float64)
    for j, e in enumerate(topo.edges):
        v = lambdas[e.id]
        v = v.detach().cpu().numpy() if isinstance(v, torch.Tensor) else np.asarray(v)
        host[j, : v.shape[0]] = np.real(v)
    return PackedLambdas(h, torch.from_numpy(host.reshape(-1)).to(DEVICE), dims)


def _chi_for(*dims_list) -> int:
    chi = 1
    for d in dims_list:
        if len(d):
            chi = max(chi, int(np.max(d)))
    return chi


def _stream():
    return _lib.stream_handle()


# --------------------------------------------------------------------- maps
def get_belief_propagation_map(traverser: Iterable, is_synchronous: bool = False):
    """One BP sweep (mappings/belief_propagation.py:24-64), asynchronous in
    traverser order by default."""
    topo = topology_for(traverser)

    def belief_propagation_map(tensors, messages) -> PackedMessages:
        L = _lib.lib()
        pt = pack_tensors(topo, tensors)
        h = topo.handle(pt.handle.chi, pt.handle.phys, is_synchronous)
        dtype = pt.buf.dtype
        pm = pack_messages(h, messages, pt.dims, dtype)
        n2e = torch.empty_like(pm.n2e)
        e2n = torch.empty_like(pm.e2n)
        _lib.check(L.dq_bp_sweep(h.ptr, _lib.dtype_code(dtype), _lib.ptr(pt.buf), _lib.ptr(pm.n2e), _lib.ptr(pm.e2n),
                                 _lib.ptr(n2e), _lib.ptr(e2n), None, _stream()))
        return PackedMessages(h, n2e, e2n, pt.dims)

    return belief_propagation_map


def get_vidal_gauge_fixing_map(traverser: Iterable, eps):
    """Messages -> Vidal gauge (Gamma, lambda) (mappings/vidal_gauge.py:21-63)."""
    topo = topology_for(traverser)
    eps = float(eps)

    def vidal_gauge_fixing_map(tensors, messages) -> Tuple[PackedTensors, PackedLambdas]:
        L = _lib.lib()
        pt = pack_tensors(topo, tensors)
        h = pt.handle
        dtype = pt.buf.dtype
        pm = pack_messages(h, messages, pt.dims, dtype)
        gam = torch.empty_like(pt.buf)
        lam = torch.empty(topo.n_edges * h.chi, dtype=torch.float64, device=DEVICE)
        _lib.check(L.dq_vidal_gauge(h.ptr, _lib.dtype_code(dtype), _lib.ptr(pt.buf), _lib.ptr(pm.n2e), eps,
                                    _lib.ptr(gam), _lib.ptr(lam), None, _stream()))
        return PackedTensors(h, gam, pt.dims), PackedLambdas(h, lam, pt.dims)

    return vidal_gauge_fixing_map


def get_vidal_gauge_distance_map(traverser: Iterable, eps):
    """Distance to the Vidal gauge (mappings/vidal_distance.py:17-42); a 0-d
    float64 device tensor."""
    topo = topology_for(traverser)
    eps = float(eps)

    def vidal_gauge_distance(tensors, lambdas) -> torch.Tensor:
        L = _lib.lib()
        pt = pack_tensors(topo, tensors)
        h = pt.handle
        pl = pack_lambdas(h, lambdas, pt.dims)
        out = torch.empty(1, dtype=torch.float64, device=DEVICE)
        _lib.check(L.dq_vidal_distance(h.ptr, _lib.dtype_code(pt.buf.dtype), _lib.ptr(pt.buf), _lib.ptr(pl.lam), eps,
                                       _lib.ptr(out), None, _stream()))
        return out[0]

    return vidal_gauge_distance


def get_symmetric_gauge_fixing_map(traverser: Iterable, eps):
    """Vidal -> symmetric gauge (mappings/symmetric_gauge.py:17-39)."""
    topo = topology_for(traverser)
    eps = float(eps)

    def symmetric_gauge(tensors, lambdas) -> PackedTensors:
        L = _lib.lib()
        pt = pack_tensors(topo, tensors)
        h = pt.handle
        pl = pack_lambdas(h, lambdas, pt.dims)
        out = torch.empty_like(pt.buf)
        _lib.check(L.dq_symmetric_gauge(h.ptr, _lib.dtype_code(pt.buf.dtype), _lib.ptr(pt.buf), _lib.ptr(pl.lam), eps,
                                        _lib.ptr(out), _stream()))
        return PackedTensors(h, out, pt.dims)

    return symmetric_gauge


def lambdas2messages(edge_core_diagonal_tensors) -> Mapping:
    """diag(lambda) in all four directions of each edge, not normalised
    (mappings/symmetric_gauge.py:50-61)."""
    lams = edge_core_diagonal_tensors
    if isinstance(lams, PackedLambdas):
        L = _lib.lib()
        h = lams.handle
        dtype = torch.complex128
        n = 2 * h.topo.n_edges * h.chi * h.chi
        n2e = torch.empty(n, dtype=dtype, device=DEVICE)
        e2n = torch.empty(n, dtype=dtype, device=DEVICE)
        _lib.check(L.dq_lambdas_to_messages(h.ptr, _lib.dtype_code(dtype), _lib.ptr(lams.lam), _lib.ptr(n2e),
                                            _lib.ptr(e2n), _stream()))
        return PackedMessages(h, n2e, e2n, lams.dims)
    out = {}
    for edge_id, lm in lams.items():
        m = torch.diag(_as_tensor(lm).to(DEVICE).to(torch.complex128))
        for node_id in edge_id:
            out[MessageID(edge_id, node_id)] = m
            out[MessageID(node_id, edge_id)] = m
    return out


def messages_frob_distance(lhs, rhs) -> torch.Tensor:
    """Mean Frobenius distance of two message dicts (mappings/messages_distance.py:14-22)."""
    if len(lhs) != len(rhs):
        raise AssertionError(f"{len(lhs)}, {len(rhs)}")
    packed = lhs if isinstance(lhs, PackedMessages) else rhs if isinstance(rhs, PackedMessages) else None
    if packed is None:
        raise TypeError("messages_frob_distance needs at least one dict returned by a map of this package")
    L = _lib.lib()
    h = packed.handle
    dtype = packed.n2e.dtype
    a = pack_messages(h, lhs, packed.dims, dtype)
    b = pack_messages(h, rhs, packed.dims, dtype)
    out = torch.empty(1, dtype=torch.float64, device=DEVICE)
    _lib.check(L.dq_messages_distance(h.ptr, _lib.dtype_code(dtype), _lib.ptr(a.n2e), _lib.ptr(a.e2n),
                                      _lib.ptr(b.n2e), _lib.ptr(b.e2n), _lib.ptr(out), _stream()))
    return out[0]


def _rdm_inputs(tensors, messages):
    packed = messages if isinstance(messages, PackedMessages) else (
        tensors if isinstance(tensors, PackedTensors) else None)
    if packed is None:
        raise TypeError("density matrices need tensors or messages returned by a map of this package")
    h = packed.handle
    pt = pack_tensors(h.topo, tensors, chi=h.chi)
    dtype = pt.buf.dtype
    dims = pt.dims
    pm = pack_messages(h, messages, dims, dtype)
    return h, pt, pm, dtype


def get_one_side_density_matrix(node: Node, tensors, messages) -> torch.Tensor:
    """rho of one node from its tensor and incoming messages, trace 1
    (mappings/density_matrix.py:16-37)."""
    L = _lib.lib()
    h, pt, pm, dtype = _rdm_inputs(tensors, messages)
    i = h.topo.node_index[0][node.id]
    idx = torch.tensor([i], dtype=torch.int32, device=DEVICE)
    rho = torch.empty((1, h.phys, h.phys), dtype=torch.complex128, device=DEVICE)
    _lib.check(L.dq_rdm1(h.ptr, _lib.dtype_code(dtype), _lib.ptr(pt.buf), _lib.ptr(pm.e2n), _lib.ptr(idx), 1,
                         _lib.ptr(rho), _stream()))
    return rho[0]


def get_two_side_density_matrix(edge: Edge, tensors, messages) -> torch.Tensor:
    """Two-site BP density matrix of an edge, rows (s_a s_b) for the edge's
    node order (no reference counterpart)."""
    L = _lib.lib()
    h, pt, pm, dtype = _rdm_inputs(tensors, messages)
    j = h.topo.edge_index[0][edge.id]
    idx = torch.tensor([j], dtype=torch.int32, device=DEVICE)
    p2 = h.phys * h.phys
    rho = torch.empty((1, p2, p2), dtype=torch.complex128, device=DEVICE)
    _lib.check(L.dq_rdm2(h.ptr, _lib.dtype_code(dtype), _lib.ptr(pt.buf), _lib.ptr(pm.e2n), _lib.ptr(idx), 1,
                         _lib.ptr(rho), _stream()))
    return rho[0]
