"""TensorNetworkCircuit: PicoQuant's circuit -> tensor-network graph (src/layer3.jl).

Same data model and labelling as the reference: nodes `node_<k>`, edges
`index_<k>`, one layer per gate, gate tensors indexed [in_1..in_k, out_1..out_k]
in column-major order.  Gate data goes to the backend through
save_tensor_data exactly where the reference calls it.
"""
from __future__ import annotations

import json

import math
from collections import OrderedDict
from dataclasses import dataclass, field

import numpy as np

from .backend import AbstractBackend, HIPBackend

SQ2 = 1.0 / math.sqrt(2.0)


@dataclass
class Node:
    """src/layer3.jl:23-30."""
    indices: list
    dims: list
    data_label: str


@dataclass
class Edge:
    """src/layer3.jl:43-53."""
    src: str | None = None
    dst: str | None = None
    qubit: int | None = None
    virtual: bool = False


_VERBS = ("save_tensor_data", "load_tensor_data", "contract_tensors", "save_output",
          "reshape_tensor", "permute_tensor", "decompose_tensor", "delete_tensor", "view_tensor")


@dataclass
class TensorNetworkCircuit:
    """src/layer3.jl:56-118."""
    backend: AbstractBackend
    number_qubits: int
    input_qubits: list
    output_qubits: list
    nodes: OrderedDict = field(default_factory=OrderedDict)
    edges: OrderedDict = field(default_factory=OrderedDict)
    qubit_ordering: list = field(default_factory=list)
    counters: dict = field(default_factory=dict)
    node_layers: dict = field(default_factory=dict)

    @classmethod
    def create(cls, qubits: int, backend: AbstractBackend | None = None):
        """TensorNetworkCircuit(qubits, backend) (src/layer3.jl:93-118)."""
        backend = backend if backend is not None else HIPBackend()
        labels = [f"index_{i}" for i in range(1, qubits + 1)]
        edges = OrderedDict((labels[i], Edge(None, None, i + 1, False)) for i in range(qubits))
        return cls(backend, qubits, list(labels), list(labels), OrderedDict(), edges,
                   list(range(1, qubits + 1)), {"index": qubits, "node": 0, "layer": 0}, {})

    # The nine backend verbs are forwarded (src/layer3.jl:124-134).
    def __getattr__(self, name):
        if name in _VERBS:
            return getattr(self.backend, name)
        raise AttributeError(name)


def new_label(network, label_str):
    """src/layer3.jl:145-148."""
    network.counters[label_str] += 1
    return f"{label_str}_{network.counters[label_str]}"


def decompose_gate(gate_data, threshold=1e-15):
    """src/layer3.jl:715-739: split a 2-qubit gate [in_a,in_b,out_a,out_b] into
    B[in_a,out_a,v] and C[v,in_b,out_b] by SVD of the (in_a,out_a)x(in_b,out_b)
    matrix, keeping singular values above threshold."""
    d = gate_data.shape
    A = np.reshape(np.transpose(gate_data, (0, 2, 1, 3)), (d[0] * d[2], d[1] * d[3]), order="F")
    U, S, Vh = np.linalg.svd(A, full_matrices=False)
    chi = int(np.sum(S > threshold))
    s = np.sqrt(S[:chi])
    B = np.reshape(U[:, :chi] * s[None, :], (d[0], d[2], chi), order="F")
    C = np.reshape(s[:, None] * Vh[:chi, :], (chi, d[1], d[3]), order="F")
    return B, C


def _rewire(network, node_label, input_index, output_index, qubit):
    network.edges[output_index] = Edge(node_label, network.edges[input_index].dst, qubit)
    dst = network.edges[input_index].dst
    if dst is not None:
        out_node = network.nodes[dst]
        for i, x in enumerate(out_node.indices):
            if x == input_index:
                out_node.indices[i] = output_index
    network.edges[input_index].dst = node_label


def add_gate(network, gate_data, target_qubits, decompose=False):
    """add_gate! (src/layer3.jl:158-249); target qubits are 1-based."""
    gate_data = np.asarray(gate_data)
    n = len(target_qubits)
    input_indices = [network.output_qubits[q - 1] for q in target_qubits]
    output_indices = [new_label(network, "index") for _ in range(n)]
    for i, q in enumerate(target_qubits):
        network.output_qubits[q - 1] = output_indices[i]
    network.counters["layer"] += 1
    layer = network.counters["layer"]
    if decompose and n == 2:
        parts = decompose_gate(gate_data)
        v = new_label(network, "index")
        labels = [new_label(network, "node") for _ in range(2)]
        network.edges[v] = Edge(labels[0], labels[1], None, True)
        for i in range(2):
            lab = labels[i]
            idx = ([input_indices[0], output_indices[0], v] if i == 0
                   else [v, input_indices[1], output_indices[1]])
            network.nodes[lab] = Node(idx, list(parts[i].shape), lab)
            network.node_layers[lab] = layer
            network.backend.save_tensor_data(lab, parts[i])
            _rewire(network, lab, input_indices[i], output_indices[i], target_qubits[i])
        return labels
    lab = new_label(network, "node")
    network.nodes[lab] = Node(input_indices + output_indices, list(gate_data.shape), lab)
    network.node_layers[lab] = layer
    network.backend.save_tensor_data(lab, gate_data)
    for i in range(n):
        _rewire(network, lab, input_indices[i], output_indices[i], target_qubits[i])
    return lab


_INPUTS = {"0": [1.0, 0.0], "1": [0.0, 1.0], "+": [SQ2, SQ2], "-": [SQ2, -SQ2]}


def add_input(network, config: str):
    """add_input! (src/layer3.jl:257-284): '0', '1', '+', '-' per qubit."""
    assert len(config) == network.number_qubits
    for idx, ch in zip(network.input_qubits, config):
        if network.edges[idx].src is None:
            lab = new_label(network, "node")
            network.nodes[lab] = Node([idx], [2], lab)
            network.node_layers[lab] = 0
            network.backend.save_tensor_data(lab, np.array(_INPUTS[ch], dtype=np.complex128))
            network.edges[idx].src = lab


def add_output(network, config: str):
    """add_output! (src/layer3.jl:292-313): '0' -> [1,0], otherwise [0,1].

    Extension: 'x' leaves that qubit's output open, so one contraction yields
    the amplitudes of every bitstring over the open qubits (a batch of output
    bitstrings); the reference always closes every qubit.
    """
    assert len(config) == network.number_qubits
    for i in range(network.number_qubits):
        pos = network.qubit_ordering[i]
        oidx, ch = network.output_qubits[pos - 1], config[i]
        if ch == "x":
            continue
        if network.edges[oidx].dst is None:
            lab = new_label(network, "node")
            network.nodes[lab] = Node([oidx], [2], lab)
            network.node_layers[lab] = -1
            data = np.array([1.0, 0.0] if ch == "0" else [0.0, 1.0], dtype=np.complex128)
            network.backend.save_tensor_data(lab, data)
            network.edges[oidx].dst = lab


def gate_matrix_to_tensor(U):
    """src/layer3.jl:566-568: reshape(permutedims(U, (2,1)), 2, ..., 2) -- for a
    qiskit-ordered matrix U (first target = least significant bit) the tensor
    is G[in_1..in_k, out_1..out_k] = U[out, in]."""
    U = np.asarray(U, dtype=np.complex128)
    k = int(round(math.log2(U.shape[0])))
    return np.reshape(np.ascontiguousarray(U.T), (2,) * (2 * k), order="F")


def network_from_gates(n_qubits, gates, backend=None, decompose=False):
    """convert_qiskit_circ_to_network (src/layer3.jl:544-573) for a gate list
    [(U, [q0, q1, ...]), ...] with 0-based qubits, without qiskit."""
    tn = TensorNetworkCircuit.create(n_qubits, backend)
    for U, qs in gates:
        add_gate(tn, gate_matrix_to_tensor(U), [q + 1 for q in qs], decompose=decompose)
    return tn


# -- accessors (src/layer3.jl:324-455) ---------------------------------------

def inneighbours(network, node_label):
    out = []
    for ix in network.nodes[node_label].indices:
        e = network.edges[ix]
        if e.src is not None and e.src != node_label and not e.virtual:
            out.append(e.src)
    return out


def outneighbours(network, node_label):
    out = []
    for ix in network.nodes[node_label].indices:
        e = network.edges[ix]
        if e.dst is not None and e.dst != node_label and not e.virtual:
            out.append(e.dst)
    return out


def virtualneighbours(network, node_label):
    out = []
    for ix in network.nodes[node_label].indices:
        e = network.edges[ix]
        if e.dst is not None and e.dst != node_label and e.virtual:
            out.append(e.dst)
        elif e.src is not None and e.src != node_label and e.virtual:
            out.append(e.src)
    return out


def virtualedges(network, node_label):
    return [x for x in network.nodes[node_label].indices if network.edges[x].virtual]


# -- GATE_TENSORS (src/layer3.jl:746-759) --------------------------------------

GATE_TENSORS = {
    "I": np.array([[1, 0], [0, 1]], dtype=np.complex128),
    "X": np.array([[0, 1], [1, 0]], dtype=np.complex128),
    "Y": np.array([[0, -1j], [1j, 0]], dtype=np.complex128),
    "Z": np.array([[1, 0], [0, -1]], dtype=np.complex128),
    "H": np.array([[1, 1], [1, -1]], dtype=np.complex128) * SQ2,
    "S": np.array([[1, 0], [0, 1j]], dtype=np.complex128),
    "T": np.array([[1, 0], [0, (1 + 1j) * SQ2]], dtype=np.complex128),
    "CX": np.reshape(np.array([[1, 0, 0, 0], [0, 0, 0, 1], [0, 0, 1, 0], [0, 1, 0, 0]],
                              dtype=np.complex128), (2, 2, 2, 2), order="F"),
    "CZ": np.reshape(np.diag([1, 1, 1, -1]).astype(np.complex128), (2, 2, 2, 2), order="F"),
    "SWAP": np.reshape(np.array([[1, 0, 0, 0], [0, 0, 1, 0], [0, 1, 0, 0], [0, 0, 0, 1]],
                                dtype=np.complex128), (2, 2, 2, 2), order="F"),
}


def gate_tensor(gate: str):
    """src/layer3.jl:789-802."""
    if gate not in GATE_TENSORS:
        raise KeyError(f"Invalid input to gate_tensor: {gate}")
    return GATE_TENSORS[gate].copy()


# -- serialisation (src/layer3.jl:578-707) -----------------------------------
# The structure only (nodes, edges, qubit bookkeeping), in the reference's
# schema, so files written by either side load in the other; tensor data
# stays in the backend under each node's data_label.

def edge_to_dict(edge):
    """to_dict(edge::Edge) (src/layer3.jl:584-587)."""
    return {"src": edge.src, "dst": edge.dst, "virtual": edge.virtual, "qubit": edge.qubit}


def edge_from_dict(d):
    """src/layer3.jl:594-599."""
    return Edge(d["src"], d["dst"], d["qubit"], d["virtual"])


def node_to_dict(node):
    """to_dict(node::Node) (src/layer3.jl:606-611): dims as strings."""
    return {"indices": list(node.indices), "dims": [str(x) for x in node.dims], "data_label": node.data_label}


def node_from_dict(d):
    """src/layer3.jl:618-623."""
    return Node(list(d["indices"]), [int(x) for x in d["dims"]], d["data_label"])


def to_dict(network):
    """to_dict(network) (src/layer3.jl:630-647)."""
    return OrderedDict([
        ("number_qubits", network.number_qubits),
        ("edges", OrderedDict((k, edge_to_dict(v)) for k, v in network.edges.items())),
        ("nodes", OrderedDict((k, node_to_dict(v)) for k, v in network.nodes.items())),
        ("input_qubits", list(network.input_qubits)),
        ("output_qubits", list(network.output_qubits)),
        ("qubit_ordering", list(network.qubit_ordering)),
        ("node_layers", OrderedDict((k, v) for k, v in network.node_layers.items())),
    ])


def network_from_dict(d, backend=None):
    """network_from_dict (src/layer3.jl:654-685): counters from the largest
    index / node numbers, backend tensors not touched."""
    backend = backend if backend is not None else HIPBackend()
    counters = {"index": 0, "node": 0, "layer": 0}
    edges = OrderedDict()
    for k, v in d["edges"].items():
        counters["index"] = max(counters["index"], int(k.split("_")[-1]))
        edges[k] = edge_from_dict(v)
    nodes = OrderedDict()
    for k, v in d["nodes"].items():
        counters["node"] = max(counters["node"], int(k.split("_")[-1]))
        nodes[k] = node_from_dict(v)
    layers = {k: v for k, v in d["node_layers"].items()}
    counters["layer"] = max(layers.values(), default=0)
    return TensorNetworkCircuit(backend, d["number_qubits"], list(d["input_qubits"]), list(d["output_qubits"]),
                                nodes, edges, list(d["qubit_ordering"]), counters, layers)


def to_json(network, indent=0):
    """to_json (src/layer3.jl:690-696)."""
    return json.dumps(to_dict(network), indent=indent or None)


def network_from_json(json_str, backend=None):
    """network_from_json (src/layer3.jl:704-707)."""
    return network_from_dict(json.loads(json_str, object_pairs_hook=OrderedDict), backend)
