"""The C ABI: libdq.so loads (no GPU needed) and exports every function
declared in include/dq.h, and the ctypes table declares exactly those."""
import os
import re

from dequantizer_amd import _lib

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "dq.h")


def header_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(dq_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    names = header_functions()
    assert len(names) >= 15
    cdll = _lib.load_library()
    for n in names:
        assert hasattr(cdll, n), n
    assert sorted(_lib.SIGNATURES) == names


def test_version_and_error_without_gpu():
    cdll = _lib.load_library()
    assert cdll.dq_version() >= 100
    # argument validation fails before touching the device
    import ctypes
    out = ctypes.c_void_p()
    rc = cdll.dq_graph_create(ctypes.byref(out), 1, 0, 0, 2, None, None, None, None, None, 1)
    assert rc == 1
    assert b"bad sizes" in cdll.dq_last_error()


def test_compute_paths_fail_loudly_without_a_gpu():
    """No CPU fallback: without a visible GPU every compute entry of the
    package raises instead of computing elsewhere."""
    import numpy as np
    import pytest
    import torch

    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    import dequantizer_amd as dq
    from dequantizer_amd import annealing, exact

    with pytest.raises(RuntimeError, match="needs a ROCm GPU"):
        exact.StateVector(4)
    g = dq.TensorGraph()
    for _ in range(3):
        g.add_node(2)
    g.add_edge((0, 1), 1)
    g.add_edge((1, 2), 1)
    with pytest.raises(RuntimeError, match="needs a ROCm GPU"):
        exact.ExactQuantumEmulator(g, dq.PRNGKey(0))
    ef = annealing.EnergyFunction(np.ones(2), np.array([[0, 1], [1, 2]]), np.zeros(3))
    with pytest.raises(RuntimeError, match="needs a ROCm GPU"):
        annealing.run_quantum_annealer(ef, [(0.1, 0.0)], dq.PRNGKey(0), 4, 2, 1e-5, 10, True, "bfs")
    with pytest.raises(RuntimeError, match="needs a ROCm GPU"):
        exact.run_exact_quantum_annealer(ef, [(0.1, 0.0)], dq.PRNGKey(0), False, False)
