"""ctypes binding of libdq.so (the HIP kernels behind the C ABI in include/dq.h).

The library is built in-tree by ``dequantizer_amd.build``.  There is no
fallback: every compute entry point of this package goes through libdq, and
``lib()`` raises when the shared object is missing or no GPU is visible.
"""
import ctypes
import os
import threading

import torch

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libdq.so")

DQ_C64 = 0
DQ_C128 = 1
DQ_TENSOR_WS = 1024

_lock = threading.Lock()
_lib = None

c_int = ctypes.c_int
c_ll = ctypes.c_longlong
c_dbl = ctypes.c_double
vp = ctypes.c_void_p
ip = ctypes.POINTER(ctypes.c_int)
llp = ctypes.POINTER(ctypes.c_longlong)
ucp = ctypes.POINTER(ctypes.c_ubyte)

# name -> (restype, argtypes); mirrors include/dq.h
SIGNATURES = {
    "dq_version": (c_int, []),
    "dq_last_error": (ctypes.c_char_p, []),
    "dq_graph_create": (c_int, [ctypes.POINTER(vp), c_int, c_int, c_int, c_int, ip, ip, ip, ip, ip, c_int]),
    "dq_graph_destroy": (c_int, [vp]),
    "dq_graph_tensor_elems": (c_ll, [vp]),
    "dq_graph_tensor_offsets": (c_int, [vp, llp]),
    "dq_bp_set_schedule": (c_int, [vp, c_int, ip, ip, ucp, ucp]),
    "dq_bp_sweep": (c_int, [vp, c_int, vp, vp, vp, vp, vp, vp, vp]),
    "dq_messages_distance": (c_int, [vp, c_int, vp, vp, vp, vp, vp, vp]),
    "dq_lambdas_to_messages": (c_int, [vp, c_int, vp, vp, vp, vp]),
    "dq_vidal_gauge": (c_int, [vp, c_int, vp, vp, c_dbl, vp, vp, vp, vp]),
    "dq_vidal_distance": (c_int, [vp, c_int, vp, vp, c_dbl, vp, vp, vp]),
    "dq_vidal_distance_noalloc": (c_int, [vp, c_int, vp, vp, c_dbl, vp, vp, vp]),
    "dq_symmetric_gauge": (c_int, [vp, c_int, vp, vp, c_dbl, vp, vp]),
    "dq_apply1": (c_int, [vp, c_int, vp, vp, c_int, vp, vp]),
    "dq_simple_update": (c_int, [vp, c_int, vp, vp, vp, c_int, vp, c_int, c_dbl, c_int, vp, vp, vp]),
    "dq_decompose_gates": (c_int, [vp, c_int, c_int, c_dbl, vp, vp, c_int, vp]),
    "dq_tensor_apply1": (c_int, [c_int, vp, c_ll, c_int, vp, vp, vp, vp]),
    "dq_tensor_apply_half": (c_int, [c_int, vp, c_ll, c_int, c_ll, c_int, vp, c_int, vp, vp, vp]),
    "dq_su_tape_bytes": (c_ll, [vp, c_int]),
    "dq_simple_update_tape": (c_int, [vp, c_int, vp, vp, vp, c_int, vp, c_int, c_dbl, c_int, vp, vp, vp, vp]),
    "dq_bp_sweeps": (c_int, [vp, c_int, vp, vp, vp, vp, vp, c_int, vp, vp]),
    "dq_autotune": (c_int, [vp, c_int, c_int, c_int, vp, ctypes.c_char_p, c_int]),
    "dq_tune_override": (c_int, [c_int, c_int, c_int, c_int, c_int, c_int]),
    "dq_simple_update_vjp": (c_int, [vp, c_int, vp, vp, vp, c_int, vp, c_int, c_dbl, c_dbl, vp, vp, vp, vp, vp]),
    "dq_edge_svd_tape_bytes": (c_ll, [vp]),
    "dq_vidal_gauge_tape": (c_int, [vp, c_int, vp, vp, c_dbl, vp, vp, vp, vp, vp]),
    "dq_bp_sweep_vjp": (c_int, [vp, c_int, vp, vp, vp, vp, vp, vp]),
    "dq_vidal_gauge_vjp": (c_int, [vp, c_int, vp, vp, vp, c_dbl, c_dbl, vp, vp, vp, vp, vp, vp]),
    "dq_symmetric_gauge_vjp": (c_int, [vp, c_int, vp, vp, c_dbl, vp, vp, vp, vp]),
    "dq_lambdas_to_messages_vjp": (c_int, [vp, c_int, vp, vp, vp, vp]),
    "dq_rdm1_vjp": (c_int, [vp, c_int, vp, vp, vp, c_int, vp, vp, vp, vp]),
    "dq_rdm2_vjp": (c_int, [vp, c_int, vp, vp, vp, c_int, vp, vp, vp, vp]),
    "dq_apply1_vjp": (c_int, [vp, c_int, vp, vp, c_int, vp, vp, vp, vp]),
    "dq_rdm1": (c_int, [vp, c_int, vp, vp, vp, c_int, vp, vp]),
    "dq_rdm2": (c_int, [vp, c_int, vp, vp, vp, c_int, vp, vp]),
    "dq_update_active": (c_int, [vp, c_dbl, vp, c_int, vp, c_int, vp]),
    "dq_repad": (c_int, [vp, vp, c_int, vp, vp, vp]),
    "dq_sv_work_bytes": (c_ll, []),
    "dq_sv_set_basis": (c_int, [vp, c_int, c_ll, vp]),
    "dq_sv_apply1": (c_int, [vp, c_int, c_int, vp, vp]),
    "dq_sv_apply2": (c_int, [vp, c_int, c_int, c_int, vp, vp]),
    "dq_sv_apply_diagonal": (c_int, [vp, c_int, c_int, vp, vp, vp, vp]),
    "dq_sv_apply1_layer": (c_int, [vp, c_int, c_ll, vp, vp]),
    "dq_sv_apply_diagonal_layer1": (c_int, [vp, c_int, c_int, vp, vp, c_ll, vp, vp, vp]),
    "dq_sv_rdm1": (c_int, [vp, c_int, c_ll, vp, vp, vp]),
}


class DQError(RuntimeError):
    """A libdq call returned a non-zero status."""


def load_library(path: str = LIB_PATH) -> ctypes.CDLL:
    """Load libdq.so and declare every exported symbol (no GPU needed)."""
    if not os.path.exists(path):
        raise ImportError(f"libdq.so not found at {path}; build it with `python -m dequantizer_amd.build`")
    cdll = ctypes.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(cdll, name)
        fn.restype = res
        fn.argtypes = args
    return cdll


def lib() -> ctypes.CDLL:
    """The loaded library; raises unless a GPU is visible."""
    global _lib
    if _lib is None:
        with _lock:
            if _lib is None:
                if not torch.cuda.is_available():
                    raise RuntimeError("dequantizer_amd needs a ROCm GPU (torch.cuda.is_available() is False)")
                _lib = load_library()
    return _lib


def autotune(handle_ptr, dtype_code: int, max_gate_rank: int = 2, force: bool = False) -> str:
    """Run the library's autotuner for a graph handle (kernel variants timed
    on synthetic inputs of the handle's shape, fastest kept process-wide per
    shape); returns its timing report."""
    buf = ctypes.create_string_buffer(4096)
    check(lib().dq_autotune(handle_ptr, dtype_code, max_gate_rank, int(force), stream_handle(), buf, len(buf)))
    return buf.value.decode()


def check(status: int) -> None:
    if status != 0:
        msg = _lib.dq_last_error().decode(errors="replace") if _lib is not None else ""
        raise DQError(f"libdq status {status}: {msg}")


def ptr(t) -> int:
    """Device pointer of a torch tensor (None -> NULL)."""
    if t is None:
        return None
    return t.data_ptr()


def stream_handle(stream=None) -> int:
    s = stream if stream is not None else torch.cuda.current_stream()
    return s.cuda_stream


def int_array(values):
    arr = (ctypes.c_int * max(1, len(values)))(*[int(v) for v in values])
    return arr


def ubyte_array(values):
    return (ctypes.c_ubyte * max(1, len(values)))(*[int(v) for v in values])


def dtype_code(dtype: torch.dtype) -> int:
    if dtype == torch.complex64:
        return DQ_C64
    if dtype == torch.complex128:
        return DQ_C128
    raise TypeError(f"unsupported dtype {dtype}: use torch.complex64 or torch.complex128")
