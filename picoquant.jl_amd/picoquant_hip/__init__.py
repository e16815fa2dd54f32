"""picoquant_hip: PicoQuant's pairwise tensor-contraction path on MI355X.

Host-side mirror of the reference's backend API (src/backends.jl,
src/backends/interactive.jl, src/layer2.jl, src/layer3.jl, src/layer2/slicing.jl)
over the C-ABI library libpq_hip.so (include/pq_hip.h).
"""
from . import _lib
from .backend import AbstractBackend, HIPBackend, Metrics, record_compute_costs, record_memory_costs
from .network import (Edge, Node, TensorNetworkCircuit, add_gate, add_input, add_output,
                      decompose_gate, gate_matrix_to_tensor, gate_tensor, network_from_gates,
                      new_label)
from .contraction import (contract_network, contract_pair, create_ncon_indices,
                          full_wavefunction_contraction, random_contraction_plan, sort_indices)
from .slicing import (multi_index_partition, partition_network_on_virtual_bonds,
                      replace_with_view, slice_tensor_network)
from .optimal_order import netcon, netcon_contraction
from .network import network_from_json, to_json
from .qasm import convert_qiskit_circ_to_network, load_qasm_as_circuit, load_qasm_as_circuit_from_file
from .dsl import DSLBackend, execute_dsl_file
from .mps import (MPSState, calculate_mps_amplitudes, compress_bond, compress_tensor_chain,
                  contract_mps_tensor_network_circuit, contract_tensor_network_circuit_with_compression,
                  decompose_tensor)
from . import circuits

__all__ = [
    "AbstractBackend", "HIPBackend", "Metrics", "record_compute_costs", "record_memory_costs",
    "Edge", "Node", "TensorNetworkCircuit", "add_gate", "add_input", "add_output",
    "decompose_gate", "gate_matrix_to_tensor", "gate_tensor", "network_from_gates", "new_label",
    "contract_network", "contract_pair", "create_ncon_indices", "full_wavefunction_contraction",
    "random_contraction_plan", "sort_indices", "multi_index_partition",
    "partition_network_on_virtual_bonds", "replace_with_view", "slice_tensor_network", "circuits",
    "netcon", "netcon_contraction", "network_from_json", "to_json", "convert_qiskit_circ_to_network",
    "load_qasm_as_circuit", "load_qasm_as_circuit_from_file", "DSLBackend", "execute_dsl_file",
    "MPSState", "calculate_mps_amplitudes", "compress_bond", "compress_tensor_chain",
    "contract_mps_tensor_network_circuit", "contract_tensor_network_circuit_with_compression",
    "decompose_tensor",
]
