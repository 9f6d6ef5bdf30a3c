"""Shared builders for parity tests: the same graph and inputs for the
product (dequantizer_amd) and the oracle."""
import numpy as np

import dequantizer_amd as dq
from oracle import graph as og


def twin_graphs(build):
    """Run the same construction on a product TensorGraph and an oracle TensorGraph."""
    return build(dq.TensorGraph()), build(og.TensorGraph())


def random_tensors(graph, rng, dtype=np.complex128):
    out = {}
    for el in graph.get_traversal_iterator() or iter([]):
        if isinstance(el, dq.Node):
            shape = (*el.bond_shape, el.dimension)
            out[el.id] = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(dtype)
    return out


def random_messages(graph, rng, dtype=np.complex128):
    out = {}
    for el in graph.get_traversal_iterator() or iter([]):
        for nb in el.neighbors:
            d = nb.dimension if isinstance(el, dq.Node) else el.dimension
            a = rng.standard_normal((d, d)) + 1j * rng.standard_normal((d, d))
            m = a @ a.conj().T
            out[dq.MessageID(el.id, nb.id)] = (m / np.trace(m)).astype(dtype)
    return out


def to_oracle_messages(msgs):
    return {og.MessageID(k.src, k.dst): np.asarray(v) for k, v in msgs.items()}


def host(x):
    import torch
    if isinstance(x, torch.Tensor):
        return x.detach().cpu().numpy()
    return np.asarray(x)
