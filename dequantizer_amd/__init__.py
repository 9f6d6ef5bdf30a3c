"""MI355X-native belief-propagation engine for graph tensor networks.

Drop-in for the public API of LuchnikovI/dequantizer (dequantizer/__init__.py:5-29):
graph container and initialisers, the whole-graph BP / Vidal / symmetric
gauge maps, and the BP quantum emulator.  All tensor arithmetic runs in the
HIP kernels of libdq.so (see include/dq.h); there is no CPU fallback.
"""
from .graph import (
    Edge,
    EdgeID,
    Element,
    ElementID,
    MessageID,
    Node,
    NodeID,
    TensorGraph,
    get_heavy_hex_ibm_eagle_lattice,
    get_heavy_hex_ibm_eagle_lattice_infinite,
    get_message_random_nonnegative_initializer,
    get_nd_lattice,
    get_potts_initializer,
    get_random_tree_tensor_graph,
    get_tensor_bloated_ghz_initializer,
    get_tensor_random_normal_initializer,
    get_tensor_std_state_initializer,
)
from .mappings import (
    get_belief_propagation_map,
    get_one_side_density_matrix,
    get_symmetric_gauge_fixing_map,
    get_two_side_density_matrix,
    get_vidal_gauge_distance_map,
    get_vidal_gauge_fixing_map,
    lambdas2messages,
    messages_frob_distance,
)
from ._lib import DQError
from .emulator import BatchedBPEmulator, BPQuantumEmulator
from .exact import ExactQuantumEmulator
from .random import PRNGKey, split

__all__ = [
    "TensorGraph", "NodeID", "EdgeID", "Node", "Edge", "Element", "ElementID", "MessageID",
    "get_nd_lattice", "get_random_tree_tensor_graph", "get_tensor_std_state_initializer",
    "get_tensor_random_normal_initializer", "get_tensor_bloated_ghz_initializer", "get_potts_initializer",
    "get_message_random_nonnegative_initializer", "get_heavy_hex_ibm_eagle_lattice",
    "get_heavy_hex_ibm_eagle_lattice_infinite", "get_belief_propagation_map", "get_vidal_gauge_fixing_map",
    "get_vidal_gauge_distance_map", "get_symmetric_gauge_fixing_map", "lambdas2messages",
    "messages_frob_distance", "get_one_side_density_matrix", "get_two_side_density_matrix",
    "PRNGKey", "split", "BPQuantumEmulator", "BatchedBPEmulator", "ExactQuantumEmulator", "DQError",
]
