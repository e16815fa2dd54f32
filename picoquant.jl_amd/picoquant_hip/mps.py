"""MPS contraction and bond compression (src/layer2.jl:423-703, src/mps.jl).

The drivers keep the reference's labels, layer order and rewiring; their
work runs on the GPU through the backend verbs: contract_tensors
(pq_contract) and decompose_tensor (pq_decompose, rocSOLVER's SVD with the
reference's threshold / max_rank rule for the bond dimension).
"""
from __future__ import annotations

import math

import numpy as np

from .contraction import contract_pair, create_ncon_indices, sort_indices
from .network import Edge, Node, inneighbours, new_label


def decompose_tensor(network, node_label, left_indices, right_indices, threshold=1e-13, max_rank=0,
                     left_label=None, right_label=None):
    """decompose_tensor! (src/layer2.jl:482-560): split node_label into B (the
    left indices and a new virtual bond) and C (the bond and the right
    indices), rewire the edges, and return (B label, C label)."""
    node = network.nodes[node_label]
    pos = {v: k + 1 for k, v in enumerate(node.indices)}
    lp = [pos[x] for x in left_indices]
    rp = [pos[x] for x in right_indices]
    b_label = new_label(network, "node") if left_label is None else left_label
    c_label = new_label(network, "node") if right_label is None else right_label
    chi = network.backend.decompose_tensor(node_label, lp, rp, threshold=threshold, max_rank=max_rank,
                                           left_label=b_label, right_label=c_label)
    b_dims = [node.dims[i - 1] for i in lp]
    c_dims = [node.dims[i - 1] for i in rp]
    if chi > 0:
        vdim = chi
    else:   # a backend that records instead of computing (src/layer2.jl:517-524)
        vdim = min(math.prod(b_dims), math.prod(c_dims))
        if max_rank > 0:
            vdim = min(vdim, max_rank)
    index_label = new_label(network, "index")
    network.nodes[b_label] = Node(list(left_indices) + [index_label], b_dims + [vdim], b_label)
    network.nodes[c_label] = Node([index_label] + list(right_indices), [vdim] + c_dims, c_label)
    for ix in left_indices:
        e = network.edges[ix]
        if e.src == node_label:
            e.src = b_label
        elif e.dst == node_label:
            e.dst = b_label
    for ix in right_indices:
        e = network.edges[ix]
        if e.src == node_label:
            e.src = c_label
        elif e.dst == node_label:
            e.dst = c_label
    network.edges[index_label] = Edge(b_label, c_label, None, True)
    if node_label not in (b_label, c_label):
        del network.nodes[node_label]
    return b_label, c_label


def compress_bond(network, node_1, node_2, threshold=1e-13, max_rank=0):
    """compress_bond! (src/layer2.jl:451-468): contract the two nodes and split
    the result again along the same partition (the bond is truncated by the
    threshold / max_rank rule)."""
    left = network.nodes[node_1]
    right = network.nodes[node_2]
    left_indices = [x for x in left.indices if x not in right.indices]
    right_indices = [x for x in right.indices if x not in left.indices]
    combined = contract_pair(network, node_1, node_2)
    return decompose_tensor(network, combined, left_indices, right_indices, threshold=threshold,
                            max_rank=max_rank, left_label=node_1, right_label=node_2)


def compress_tensor_chain(network, nodes, threshold=1e-13, max_rank=0):
    """compress_tensor_chain! (src/layer2.jl:423-441): a forward pass over the
    chain's bonds, then a backward pass."""
    for i in range(len(nodes) - 1):
        compress_bond(network, nodes[i], nodes[i + 1], threshold, max_rank)
    for i in range(len(nodes) - 2, -1, -1):
        compress_bond(network, nodes[i], nodes[i + 1], threshold, max_rank)


def _gate_layers(network):
    layer_nodes: dict[int, list] = {}
    for k, v in network.node_layers.items():
        layer_nodes.setdefault(v, []).append(k)
    layers = sorted(k for k in layer_nodes if k > 0)
    if -1 in layer_nodes:
        layers.append(-1)
    return layers, layer_nodes


def _absorb(network, nodes, layer_nodes, layer):
    updated = []
    for node in layer_nodes[layer]:
        input_node = inneighbours(network, node)[0]
        if input_node not in nodes:
            raise ValueError(f"{input_node} not in the state's nodes")
        idx = nodes.index(input_node)
        nodes[idx] = contract_pair(network, input_node, node)
        updated.append(idx)
    return updated


def contract_mps_tensor_network_circuit(network, threshold=1e-13, max_rank=0, pbc=False):
    """contract_mps_tensor_network_circuit! (src/layer2.jl:570-625): the input
    nodes become the MPS; each gate layer is absorbed node by node, then the
    bonds between consecutive updated sites are compressed (gates must act
    on neighbouring qubits).  Each MPS tensor is saved under its own label;
    returns the MPS node labels."""
    mps_nodes = [network.edges[x].src for x in network.input_qubits]
    if any(x is None for x in mps_nodes):
        raise ValueError("Input qubit values must be set")
    layers, layer_nodes = _gate_layers(network)
    for layer in layers:
        updated = sorted(_absorb(network, mps_nodes, layer_nodes, layer))
        for i in range(len(updated) - 1):
            dist = updated[i + 1] - updated[i]
            if pbc:
                dist = min(dist, updated[i] + len(mps_nodes) - updated[i + 1])
            if dist != 1:
                raise ValueError("Gates between non-neighboring qubits")
            compress_bond(network, mps_nodes[updated[i]], mps_nodes[updated[i + 1]], threshold, max_rank)
    for node in mps_nodes:
        network.backend.save_output(node, node)
    return mps_nodes


def calculate_mps_amplitudes(network, mps_nodes, result="result"):
    """calculate_mps_amplitudes! (src/layer2.jl:634-645): contract the chain,
    permute to qubit_ordering, reshape to a vector, save as `result`."""
    out = mps_nodes[0]
    for node in mps_nodes[1:]:
        out = contract_pair(network, out, node)
    network.backend.permute_tensor(out, list(network.qubit_ordering))
    network.backend.reshape_tensor(out, [list(range(1, len(mps_nodes) + 1))])
    network.backend.save_output(out, result)
    return out


def contract_tensor_network_circuit_with_compression(network, threshold=1e-13, max_rank=0):
    """contract_tensor_network_circuit_with_compression! (src/layer2.jl:655-703):
    as the MPS driver, compressing the bonds between consecutive updated nodes
    of each layer in the order they were absorbed, with no neighbour check."""
    nodes = [network.edges[x].src for x in network.input_qubits]
    if any(x is None for x in nodes):
        raise ValueError("Input qubit values must be set")
    layers, layer_nodes = _gate_layers(network)
    for layer in layers:
        updated = _absorb(network, nodes, layer_nodes, layer)
        for i in range(len(updated) - 1):
            compress_bond(network, nodes[updated[i]], nodes[updated[i + 1]], threshold, max_rank)
    for node in nodes:
        network.backend.save_output(node, node)
    return nodes


class MPSState:
    """MPSState (src/mps.jl:9-110): single amplitudes of an MPS by selecting
    each site's output index and contracting the chain.  The site tensors are
    read once (load_tensor_data) and the selections contracted on the host:
    each amplitude is a chain of chi x chi products.  The configuration bit
    read for site k is the one the reference's `ordering` names
    (ordering[qubit_ordering[k]] = position of site k's output index within
    its node, src/mps.jl:51-54), kept as written."""

    def __init__(self, network, mps_nodes, dtype=np.complex64):
        self.n = len(mps_nodes)
        self.nodes, self.data = [], []
        out_pos = []
        for label in mps_nodes:
            node = network.nodes[label]
            p = next(k for k, x in enumerate(node.indices) if x in network.output_qubits)
            other = [k for k in range(len(node.indices)) if k != p]
            t = np.transpose(np.asarray(network.backend.load_tensor_data(label)), [p] + other).astype(dtype)
            self.nodes.append(Node([node.indices[k] for k in other], [node.dims[k] for k in other], label))
            self.data.append((t[0], t[1]))
            out_pos.append(p + 1)
        self.ordering = [0] * self.n
        for i, e in zip(out_pos, network.qubit_ordering):
            self.ordering[e - 1] = i

    @property
    def shape(self):
        return (2,) * self.n

    def __len__(self):
        return 2 ** self.n

    def __getitem__(self, conf):
        bits = [0 if ch == "0" else 1 for ch in conf] if isinstance(conf, str) else [int(b) - 1 for b in conf]
        n1 = self.nodes[0]
        d1 = self.data[0][bits[self.ordering[0] - 1]]
        for idx in range(1, self.n):
            n2 = self.nodes[idx]
            common, remaining = sort_indices(n1, n2)
            ia, ib = create_ncon_indices(n1, n2, common, remaining)
            d1 = _ncon(d1, ia, self.data[idx][bits[self.ordering[idx] - 1]], ib)
            n1 = Node(remaining, [1] * len(remaining), "n1")
        return complex(np.asarray(d1).reshape(-1)[0])


def _ncon(A, ia, B, ib):
    """Pairwise ncon on the host for MPSState's small selections: labels > 0
    are summed, labels < 0 open, the output in symdiff(ia, ib) order (A's
    open labels, then B's) as src/layer1.jl:85-92."""
    out = [x for x in list(ia) + list(ib) if x < 0]
    return np.einsum(f"{_letters(ia)},{_letters(ib)}->{_letters(out)}", A, B)


def _letters(labels):
    alpha = "abcdefghijklmnopqrstuvwxyzABCDEFGHIJKLMNOPQRSTUVWXYZ"
    if any(abs(x) > 26 or x == 0 for x in labels):
        raise ValueError("MPSState selections support labels 1..26 and -1..-26")
    return "".join(alpha[(x - 1) if x > 0 else 26 + (-x - 1)] for x in labels)
