"""netcon: optimal contraction plans for small networks
(src/layer2/netcon_contraction.jl).

The reference hands (tensor indices, index dims) to
TensorOperations.optimaltree; here the search runs natively in libpq_hip
(pq_netcon: cost-capped breadth-first search over subnetworks, exact in the
total multiply-add count that Metrics.flops records).  The tree and plan
conversions follow the reference's helpers.
"""
from __future__ import annotations

import ctypes as ct

from . import _lib
from .contraction import contract_network

MAX_NODES = 36     # src/layer2/netcon_contraction.jl:118-120


def create_netcon_input(network):
    """src/layer2/netcon_contraction.jl:49-73: per-tensor indices, the index
    dims and the tensors' data labels, in node order."""
    tensor_indices, tensor_dims, tensor_labels = [], [], []
    for lab, node in network.nodes.items():
        tensor_indices.append(list(node.indices))
        tensor_dims.append(list(node.dims))
        tensor_labels.append(node.data_label or lab)
    index_dims = {}
    for idx, dims in zip(tensor_indices, tensor_dims):
        index_dims.update(zip(idx, dims))
    return tensor_indices, index_dims, tensor_labels


def optimal_tree(tensor_indices, index_dims, max_work=0):
    """(tree, cost): tree is a nested pair list over tensor positions
    (0-based), cost the total multiply-adds.  Raises PQError when the search
    exceeds max_work pair evaluations (0 = 2^32)."""
    lib = _lib.load()
    n = len(tensor_indices)
    ids = {}
    labels, dims, rank = [], [], []
    for idx in tensor_indices:
        rank.append(len(idx))
        for i in idx:
            labels.append(ids.setdefault(i, len(ids)))
            dims.append(index_dims[i])
    plan = (ct.c_int32 * max(1, 2 * (n - 1)))()
    cost = ct.c_double()
    rc = lib.pq_netcon(n, (ct.c_int32 * max(1, n))(*rank), _lib.i64arr(labels), _lib.i64arr(dims),
                       int(max_work), plan, ct.byref(cost))
    if rc != _lib.PQ_OK:
        raise _lib.PQError(rc, "netcon: invalid network" if rc == _lib.PQ_ERR_ARG
                           else "netcon: search too large for an exact order")
    trees = list(range(n))
    for s in range(n - 1):
        trees.append([trees[plan[2 * s]], trees[plan[2 * s + 1]]])
    return trees[-1], cost.value


def convert_tree_to_plan(tree, last_node_label, tensor_labels, plan=None):
    """src/layer2/netcon_contraction.jl:95-123: post-order pairs; the pair
    that completes a subtree is named node_<last + 1>.  Returns
    (label, last_node_label, plan)."""
    plan = [] if plan is None else plan
    if isinstance(tree, int):
        return tensor_labels[tree], last_node_label, plan
    a, last_node_label, _ = convert_tree_to_plan(tree[0], last_node_label, tensor_labels, plan)
    b, last_node_label, _ = convert_tree_to_plan(tree[1], last_node_label, tensor_labels, plan)
    plan.append([a, b])
    last_node_label += 1
    return f"node_{last_node_label}", last_node_label, plan


def replace_leaf_integers_with_labels(tree, labels):
    """src/layer2/netcon_contraction.jl:131-143."""
    if isinstance(tree, int):
        return labels[tree]
    return [replace_leaf_integers_with_labels(tree[0], labels), replace_leaf_integers_with_labels(tree[1], labels)]


def netcon(network, return_type="plan", max_work=0):
    """src/layer2/netcon_contraction.jl:14-41: "plan" -> [[A, B], ...],
    "tree" -> nested label pairs, anything else -> (integer tree, labels)."""
    tensor_indices, index_dims, tensor_labels = create_netcon_input(network)
    tree, _ = optimal_tree(tensor_indices, index_dims, max_work)
    if return_type == "tree":
        return replace_leaf_integers_with_labels(tree, tensor_labels)
    if return_type == "plan":
        return convert_tree_to_plan(tree, network.counters["node"], tensor_labels)[2]
    return tree, tensor_labels


def netcon_contraction(network, output_shape=""):
    """netcon_contraction! (src/layer2/netcon_contraction.jl:116-123)."""
    if len(network.nodes) > MAX_NODES:
        raise ValueError("The number of nodes in the network exceed the maximum for netcon")
    plan = netcon(network)
    contract_network(network, plan, output_shape)
