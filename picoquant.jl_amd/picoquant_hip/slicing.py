"""Slicing (src/layer2/slicing.jl): fix the values of chosen bonds so that each
slice is an independent, smaller contraction; the slices sum to the result."""
from __future__ import annotations

import math

from .network import Node, new_label


def multi_index_partition(dims, number_partitions, partition):
    """src/layer2/slicing.jl:12-26: the 1-based multi-index of `partition` over
    the leading dims whose product equals number_partitions (first index fastest)."""
    total, num = 1, 1
    while num < len(dims):
        total *= dims[num - 1]
        if total == number_partitions:
            break
        if total > number_partitions:
            raise ValueError("Partitions and product of dimensions must match")
        num += 1
    p = partition - 1
    out = []
    for d in dims[:num]:
        out.append(p % d + 1)
        p //= d
    return tuple(out)


def partition_network_on_virtual_bonds(network, number_partitions, partition):
    """src/layer2/slicing.jl:37-57: virtual bonds sorted (stably) by dimension;
    the dimension comes from the stored tensor's shape, with no data transfer."""
    vb = [(k, e) for k, e in network.edges.items() if e.virtual]
    dims = []
    for k, e in vb:
        shp = network.backend.tensor_shape(e.src)
        pos = [i for i, x in enumerate(network.nodes[e.src].indices) if x == k]
        dims.append(math.prod(shp[i] for i in pos))
    order = sorted(range(len(vb)), key=lambda i: (dims[i], i))
    bonds = [vb[i][0] for i in order]
    dims = [dims[i] for i in order]
    if not dims:
        raise AssertionError("There must be some virtual bonds, try turning on decompose")
    ci = multi_index_partition(dims, number_partitions, partition)
    return bonds[: len(ci)], ci


def replace_with_view(network, node_label, bond_label, bond_range):
    """src/layer2/slicing.jl:67-93.  As in the reference the new Node keeps the
    old dims, so the cost ledger counts sliced steps at full size."""
    node = network.nodes[node_label]
    bond_idx = node.indices.index(bond_label) + 1
    view = new_label(network, "node")
    network.nodes[view] = Node(node.indices, node.dims, view)
    network.node_layers[view] = network.node_layers[node_label]
    for x in node.indices:
        e = network.edges[x]
        if e.src == node_label:
            e.src = view
        else:
            e.dst = view
    network.backend.view_tensor(view, node_label, bond_idx, bond_range)
    network.backend.delete_tensor(node_label)
    del network.nodes[node_label]
    del network.node_layers[node_label]


def slice_tensor_network(network, bond_labels, bond_values):
    """src/layer2/slicing.jl:100-110: both end nodes of each bond become views
    at the given 1-based value."""
    for b, v in zip(bond_labels, bond_values):
        n1, n2 = network.edges[b].src, network.edges[b].dst
        replace_with_view(network, n1, b, (v, v))
        replace_with_view(network, n2, b, (v, v))


def slice_with_map(network, bond_labels, bond_values):
    """slice_tensor_network with the same operations, also returning the map
    from each replaced node to the view node that now stands for it (so a
    pair plan made for the unsliced network can be translated)."""
    label_map = {}

    def current(lab):
        while lab in label_map:
            lab = label_map[lab]
        return lab

    views = set()
    for b, v in zip(bond_labels, bond_values):
        n1, n2 = network.edges[b].src, network.edges[b].dst
        for n in (n1, n2):
            view = f"node_{network.counters['node'] + 1}"
            replace_with_view(network, n, b, (v, v))
            label_map[n] = view
            views.add(view)
    # a node with several sliced bonds is replaced several times: map the
    # original to the last view (the intermediate views are gone)
    originals = {k: current(k) for k in label_map if k not in views}
    return originals


def translate_plan(plan, label_map, old_counter, new_counter):
    """Rename a pair plan for a network whose leaves were replaced (label_map:
    original leaf -> its view) and whose node counter moved from old_counter
    to new_counter.  Labels above old_counter are the plan's intermediates
    and only shift."""
    shift = new_counter - old_counter

    def tr(lab):
        k = int(lab.rsplit("_", 1)[1])
        if k > old_counter:
            return f"node_{k + shift}"
        return label_map.get(lab, lab)

    return [[tr(a), tr(b)] for a, b in plan]
