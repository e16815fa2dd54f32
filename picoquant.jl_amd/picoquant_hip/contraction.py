"""Contraction drivers (src/layer2.jl:114-405): host bookkeeping around the
backend's contract_tensors, with the reference's labels, index order and cost
ledger."""
from __future__ import annotations

from .backend import record_compute_costs
from .network import Node, new_label


def _unique(seq):
    seen = set()
    out = []
    for x in seq:
        if x not in seen:
            seen.add(x)
            out.append(x)
    return out


def sort_indices(A: Node, B: Node):
    """src/layer2.jl:209-216: common = A ∩ B (A's order); remaining =
    (A \\ common) ∪ (B \\ common), first occurrence order."""
    bset = set(B.indices)
    common = _unique([x for x in A.indices if x in bset])
    cset = set(common)
    rem = _unique([x for x in A.indices if x not in cset] + [x for x in B.indices if x not in cset])
    return common, rem


def create_ncon_indices(A: Node, B: Node, common, uncommon):
    """src/layer2.jl:226-240: summed indices get +1.., open ones -1.."""
    m = {x: i + 1 for i, x in enumerate(common)}
    m.update({x: -(i + 1) for i, x in enumerate(uncommon)})
    return [m[x] for x in A.indices], [m[x] for x in B.indices]


def contract_pair(network, a, b=None):
    """contract_pair! (src/layer2.jl:329-405).  With b=None, `a` is an edge
    label and the edge's end nodes are contracted (a no-op if the edge is gone).
    Returns the new node's label."""
    if b is None:
        e = network.edges.get(a)
        if e is None:
            return None
        return contract_pair(network, e.src, e.dst)
    if a == b:
        return None
    A = network.nodes[a]
    B = network.nodes[b]
    common, remaining = sort_indices(A, B)
    ia, ib = create_ncon_indices(A, B, common, remaining)
    dims = dict(zip(A.indices + B.indices, A.dims + B.dims))
    contracted_dims = [dims[x] for x in common]
    c_dims = [dims[x] for x in remaining]
    record_compute_costs(network.backend, c_dims, contracted_dims)
    c = new_label(network, "node")
    network.nodes[c] = Node(list(remaining), c_dims, c)
    for x in common:
        del network.edges[x]
    for x in remaining:
        e = network.edges[x]
        if e.src == a or e.src == b:
            e.src = c
        elif e.dst == a or e.dst == b:
            e.dst = c
    del network.nodes[a]
    del network.nodes[b]
    network.backend.contract_tensors(a, ia, b, ib, c)
    return c


def _shape_output(network, label, node, order_source, output_shape):
    order = [node.indices.index(x) + 1 for x in order_source if x in node.indices]
    network.backend.permute_tensor(label, order)
    if output_shape == "vector":
        network.backend.reshape_tensor(label, [list(range(1, len(node.indices) + 1))])
    elif output_shape != "":
        network.backend.reshape_tensor(label, output_shape)


def full_wavefunction_contraction(network, output_shape=""):
    """full_wavefunction_contraction! (src/layer2.jl:132-195): inputs first (outer
    products in qubit order), then each layer in ascending order (the layer's
    nodes together, then into the state), output layer last; the result is
    permuted to output_qubits[qubit_ordering] and saved as :result."""
    inputs = [network.edges[e].src for e in network.input_qubits]
    if any(x is None for x in inputs):
        raise ValueError("Please create nodes for each input before using wavefunction contraction")
    wf = inputs[0]
    for x in inputs[1:]:
        wf = contract_pair(network, wf, x)
    layer_nodes: dict[int, list] = {}
    for k, v in network.node_layers.items():
        layer_nodes.setdefault(v, []).append(k)
    layers = sorted(k for k in layer_nodes if k > 0)
    if -1 in layer_nodes:
        layers.append(-1)
    for layer in layers:
        nodes = layer_nodes[layer]
        n = nodes[0]
        for n2 in nodes[1:]:
            n = contract_pair(network, n, n2)
        wf = contract_pair(network, wf, n)
    node = network.nodes[wf]
    if len(node.indices) != 0:
        order_src = [network.output_qubits[q - 1] for q in network.qubit_ordering]
        _shape_output(network, wf, node, order_src, output_shape)
    network.backend.save_output(wf)
    return wf


def contract_network(network, plan, output_shape=""):
    """contract_network! (src/layer2.jl:250-321).  `plan` is either a list of
    edge labels (then disjoint leftovers are outer-producted) or a list of
    [A, B] node pairs.  Open outputs are put in output_qubits order; when
    some outputs are closed (a bitstring batch) the open ones keep that order."""
    if plan and isinstance(plan[0], (list, tuple)):
        for a, b in plan:
            contract_pair(network, a, b)
    else:
        for e in plan:
            contract_pair(network, e)
        while len(network.nodes) > 1:
            it = iter(network.nodes)
            n1 = next(it)
            n2 = next(it)
            contract_pair(network, n1, n2)
    out = f"node_{network.counters['node']}"
    node = network.nodes[out]
    if len(node.indices) != 0:
        src = [x for x in network.output_qubits if x in node.indices]
        order = [node.indices.index(x) + 1 for x in src]
        node.indices[:] = src
        network.backend.permute_tensor(out, order)
        if output_shape == "vector":
            network.backend.reshape_tensor(out, [list(range(1, len(node.indices) + 1))])
        elif output_shape != "":
            network.backend.reshape_tensor(out, output_shape)
    network.backend.save_output(out)
    return out


def random_contraction_plan(network, rng):
    """src/layer2.jl:119-122: the closed edges in random order (numpy Generator)."""
    closed = [k for k, v in network.edges.items() if v.src is not None and v.dst is not None]
    return [closed[i] for i in rng.permutation(len(closed))]
