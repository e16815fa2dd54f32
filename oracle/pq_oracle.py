"""CPU oracle: a numpy restatement of PicoQuant's tensor-network contraction path.

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import this module, and only as the
checker (or the timed CPU proxy).  The product path (``picoquant.jl_amd``) never
imports it and never falls back to it.

Every function restates the reference's behaviour and cites the file:line it
follows (paths relative to the PicoQuant.jl source tree).  Storage convention:
a tensor is a numpy array whose axis ``i`` is the Julia index ``i``; Julia's
column-major reshapes map to ``order='F'``.

Pins (what checks each part; see tests/test_oracle.py):
  * contract semantics / index order ......... examples/ghz_3.json -> ghz_3_contracted.json
  * cost ledger (Metrics) ..................... test/layer2_tests.jl:127-132 (124 / 44 / 8)
  * full_wavefunction_contraction ............ brute-force state vector + QFT closed form
  * contract_network, single amplitudes ...... brute-force state vector; plan invariance
  * slicing .................................. sum over slices == unsliced
  * decompose_gate ............................ round trip (test/layer3_tests.jl:57-71)
  * bgreedy tau->0 ............................ hand-worked known answers (alpha = 1 and 0: every
                                                pair energy, the plan, its costs), planted energy
                                                mistakes rejected, plan validity + ledger re-count
                                                (tests/test_greedy_parity.py, tests/test_oracle.py)
  * optimal_contraction_cost .................. hand-counted costs of tiny chains / stars
  * gate tensors .............................. closed forms + unitarity
  * decompose_tensor / MPS drivers ............ exact (untruncated) MPS == brute-force
                                                state vector; GHZ-5 closed form and
                                                MPSState amplitudes (test/layer2_tests.jl:
                                                300-347); Eckart-Young residual of a
                                                truncated split; chain compression to one
                                                bond of dimension 2 (:241-295); threshold /
                                                max_rank bond dimensions (:349-416)
"""

from __future__ import annotations

import cmath
import math
from collections import OrderedDict
from dataclasses import dataclass, field

import numpy as np

# --------------------------------------------------------------------------- #
# Gates (qiskit little-endian matrices U: row = output bits, col = input bits,
# first target qubit = least significant bit)
# --------------------------------------------------------------------------- #

SQ2 = 1.0 / math.sqrt(2.0)
I2 = np.eye(2, dtype=np.complex128)
X = np.array([[0, 1], [1, 0]], dtype=np.complex128)
Y = np.array([[0, -1j], [1j, 0]], dtype=np.complex128)
Z = np.array([[1, 0], [0, -1]], dtype=np.complex128)
H = np.array([[1, 1], [1, -1]], dtype=np.complex128) * SQ2
W = (X + Y) * SQ2


def sqrt_pauli(P):
    """Principal square root of an involutory Hermitian P: (1+i)/2 I + (1-i)/2 P."""
    return 0.5 * (1 + 1j) * I2 + 0.5 * (1 - 1j) * P


SQRT_X = sqrt_pauli(X)
SQRT_Y = sqrt_pauli(Y)
SQRT_W = sqrt_pauli(W)
CZ = np.diag([1, 1, 1, -1]).astype(np.complex128)
SWAP = np.array([[1, 0, 0, 0], [0, 0, 1, 0], [0, 1, 0, 0], [0, 0, 0, 1]],
                dtype=np.complex128)
CX = np.array([[1, 0, 0, 0], [0, 0, 0, 1], [0, 0, 1, 0], [0, 1, 0, 0]],
              dtype=np.complex128)  # control = first qarg (LSB), qiskit order


def cphase(theta):
    return np.diag([1, 1, 1, cmath.exp(1j * theta)]).astype(np.complex128)


def gate_matrix_to_tensor(U):
    """src/layer3.jl:566-568: data = reshape(permutedims(U,(2,1)), 2,...,2).

    Column-major reshape of U^T, so G[in_1..in_k, out_1..out_k] = U[out, in].
    """
    k = int(round(math.log2(U.shape[0])))
    return np.reshape(np.ascontiguousarray(U.T), (2,) * (2 * k), order="F")


# --------------------------------------------------------------------------- #
# Backend (InteractiveBackend{Array{ComplexF64}}) and cost ledger
# --------------------------------------------------------------------------- #


@dataclass
class Metrics:
    """src/backends.jl:8-20."""
    max_tensor_size: int = 0
    total_space_allocated: int = 0
    flops: int = 0


def record_memory_costs(backend, tensor_memory):
    """src/backends.jl:50-57."""
    backend.metrics.total_space_allocated += tensor_memory
    if tensor_memory > backend.metrics.max_tensor_size:
        backend.metrics.max_tensor_size = tensor_memory


def record_compute_costs(backend, open_dims, contracted_dims):
    """src/backends.jl:30-41 (prod of an empty list is 1)."""
    space = int(np.prod(open_dims, dtype=np.int64)) if open_dims else 1
    record_memory_costs(backend, space)
    fl = space * (int(np.prod(contracted_dims, dtype=np.int64)) if contracted_dims else 1)
    backend.metrics.flops += fl
    return fl, space


def contract_tensors_data(A, ia, B, ib):
    """src/layer1.jl:85-92: tensorcontract(A, ia, B, ib, symdiff(ia, ib)).

    Positive labels are summed, negative labels are open; the output order is
    symdiff(ia, ib) = A's open labels in A order, then B's open labels in B order.
    """
    ia = list(ia)
    ib = list(ib)
    common = [l for l in ia if l in ib]
    ax_a = [ia.index(l) for l in common]
    ax_b = [ib.index(l) for l in common]
    out = np.tensordot(A, B, axes=(ax_a, ax_b))
    # tensordot emits A's free axes then B's free axes, each in input order,
    # which is exactly symdiff(ia, ib).
    return out


class InteractiveBackend:
    """src/backends/interactive.jl:10-172 with T = Array{ComplexF64}."""

    def __init__(self, dtype=np.complex128):
        self.tensors: dict[str, np.ndarray] = {}
        self.metrics = Metrics()
        self.dtype = dtype

    def save_tensor_data(self, label, data):          # :32-36
        self.tensors[label] = np.array(data, dtype=self.dtype)

    def load_tensor_data(self, label):                 # :44-49
        return self.tensors.get(label)

    def contract_tensors(self, a, ia, b, ib, c):       # :60-75
        A = self.tensors[a]
        B = self.tensors[b]
        self.save_tensor_data(c, contract_tensors_data(A, ia, B, ib))
        self.delete_tensor(a)
        self.delete_tensor(b)

    def save_output(self, node, name="result"):        # :84-88
        self.save_tensor_data(name, self.tensors[node])

    def reshape_tensor(self, label, groups):           # :97-102, src/layer1.jl:100-103
        t = self.tensors[label]
        dims = [int(np.prod([t.shape[g - 1] for g in grp])) for grp in groups]
        self.tensors[label] = np.reshape(t, dims, order="F")

    def permute_tensor(self, label, axes):             # :111-115, src/layer1.jl:111-114
        self.tensors[label] = np.transpose(self.tensors[label], [a - 1 for a in axes])

    def delete_tensor(self, label):                    # :159-161
        self.tensors.pop(label, None)

    def view_tensor(self, view_node, node, bond_idx, bond_range):  # :169-172
        t = self.tensors[node]
        sl = [slice(None)] * t.ndim
        lo, hi = bond_range
        sl[bond_idx - 1] = slice(lo - 1, hi)          # Julia lo:hi, 1-based inclusive
        self.tensors[view_node] = np.array(t[tuple(sl)])   # src/layer1.jl:191-194

    def decompose_tensor(self, label, left, right, threshold=1e-13, max_rank=0,
                         left_label=None, right_label=None):
        """src/backends/interactive.jl:130-152 -> src/layer1.jl:146-184."""
        t = self.tensors[label]
        B, C, chi = decompose_tensor_data(t, left, right, threshold, max_rank)
        self.tensors[left_label] = B
        self.tensors[right_label] = C
        self.delete_tensor(label)
        return chi


def decompose_tensor_data(t, left, right, threshold=1e-13, max_rank=0):
    """src/layer1.jl:146-184 (SVD split; chi from the normalised spectrum)."""
    dims = t.shape
    ld = [dims[x - 1] for x in left]
    rd = [dims[x - 1] for x in right]
    A = np.transpose(t, [x - 1 for x in left + right])
    A = np.reshape(A, (int(np.prod(ld)), int(np.prod(rd))), order="F")
    U, S, Vh = np.linalg.svd(A, full_matrices=False)
    eps = np.finfo(np.float64 if t.dtype == np.complex128 else np.float32).eps
    threshold = max(threshold, math.sqrt(eps))
    s_norm = math.sqrt(float(np.sum(S ** 2)))
    chi = int(np.sum(S / s_norm > threshold))
    if max_rank > 0:
        chi = min(max_rank, chi)
    s = np.sqrt(S[:chi])
    B = np.reshape(U[:, :chi] * s[None, :], tuple(ld + [chi]), order="F")
    C = np.reshape(s[:, None] * Vh[:chi, :], tuple([chi] + rd), order="F")
    return B, C, chi


def decompose_gate(gate_data, threshold=1e-15):
    """src/layer3.jl:715-739: SVD of permutedims(G,(1,3,2,4)) as 4x4."""
    dims = gate_data.shape
    A = np.transpose(gate_data, (0, 2, 1, 3))
    A = np.reshape(A, (dims[0] * dims[2], dims[1] * dims[3]), order="F")
    U, S, Vh = np.linalg.svd(A, full_matrices=False)
    chi = int(np.sum(S > threshold))
    s = np.sqrt(S[:chi])
    B = np.reshape(U[:, :chi] * s[None, :], (dims[0], dims[2], chi), order="F")
    C = np.reshape(s[:, None] * Vh[:chi, :], (chi, dims[1], dims[3]), order="F")
    return B, C


# --------------------------------------------------------------------------- #
# Tensor network circuit (src/layer3.jl:22-313)
# --------------------------------------------------------------------------- #


@dataclass
class Node:
    indices: list
    dims: list
    data_label: str


@dataclass
class Edge:
    src: str | None
    dst: str | None
    qubit: int | None
    virtual: bool = False


@dataclass
class TensorNetworkCircuit:
    backend: object
    number_qubits: int
    input_qubits: list
    output_qubits: list
    nodes: OrderedDict = field(default_factory=OrderedDict)
    edges: OrderedDict = field(default_factory=OrderedDict)
    qubit_ordering: list = field(default_factory=list)
    counters: dict = field(default_factory=dict)
    node_layers: dict = field(default_factory=dict)


def new_network(qubits, backend=None):
    """src/layer3.jl:93-118."""
    backend = backend if backend is not None else InteractiveBackend()
    labels = [f"index_{i}" for i in range(1, qubits + 1)]
    edges = OrderedDict((labels[i], Edge(None, None, i + 1, False)) for i in range(qubits))
    return TensorNetworkCircuit(backend, qubits, list(labels), list(labels), OrderedDict(),
                                edges, list(range(1, qubits + 1)),
                                {"index": qubits, "node": 0, "layer": 0}, {})


def new_label(tn, s):
    """src/layer3.jl:145-148."""
    tn.counters[s] += 1
    return f"{s}_{tn.counters[s]}"


def add_gate(tn, gate_data, target_qubits, decompose=False):
    """src/layer3.jl:158-249 (target qubits are 1-based)."""
    n = len(target_qubits)
    input_indices = [tn.output_qubits[q - 1] for q in target_qubits]
    output_indices = [new_label(tn, "index") for _ in range(n)]
    for i, q in enumerate(target_qubits):
        tn.output_qubits[q - 1] = output_indices[i]
    tn.counters["layer"] += 1
    layer = tn.counters["layer"]

    def rewire(node_label, i):
        inp, outp = input_indices[i], output_indices[i]
        tn.edges[outp] = Edge(node_label, tn.edges[inp].dst, target_qubits[i])
        if tn.edges[inp].dst is not None:
            on = tn.nodes[tn.edges[inp].dst]
            on.indices[:] = [outp if x == inp else x for x in on.indices]
        tn.edges[inp].dst = node_label

    if decompose and n == 2:
        gd = decompose_gate(gate_data)
        v = new_label(tn, "index")
        labels = [new_label(tn, "node") for _ in range(2)]
        tn.edges[v] = Edge(labels[0], labels[1], None, True)
        for i in range(2):
            lab = labels[i]
            if i == 0:
                idx = [input_indices[0], output_indices[0], v]
            else:
                idx = [v, input_indices[1], output_indices[1]]
            tn.nodes[lab] = Node(idx, list(gd[i].shape), lab)
            tn.node_layers[lab] = layer
            tn.backend.save_tensor_data(lab, gd[i])
            rewire(lab, i)
        return labels
    lab = new_label(tn, "node")
    tn.nodes[lab] = Node(input_indices + output_indices, list(gate_data.shape), lab)
    tn.node_layers[lab] = layer
    tn.backend.save_tensor_data(lab, gate_data)
    for i in range(n):
        rewire(lab, i)
    return lab


INPUT_VECTORS = {"0": [1.0, 0.0], "1": [0.0, 1.0],
                 "+": [SQ2, SQ2], "-": [SQ2, -SQ2]}


def add_input(tn, config):
    """src/layer3.jl:257-284."""
    assert len(config) == tn.number_qubits
    for idx, ch in zip(tn.input_qubits, config):
        if tn.edges[idx].src is None:
            lab = new_label(tn, "node")
            data = np.array(INPUT_VECTORS[ch], dtype=np.complex128)
            tn.nodes[lab] = Node([idx], [2], lab)
            tn.node_layers[lab] = 0
            tn.backend.save_tensor_data(lab, data)
            tn.edges[idx].src = lab


def add_output(tn, config):
    """src/layer3.jl:292-313 ('0' -> [1,0], anything else -> [0,1]).

    A character 'x' leaves that qubit open (an extension used for batches of
    output bitstrings; the reference closes every qubit).
    """
    assert len(config) == tn.number_qubits
    for i in range(tn.number_qubits):
        pos = tn.qubit_ordering[i]
        oidx, ch = tn.output_qubits[pos - 1], config[i]
        if ch == "x":
            continue
        if tn.edges[oidx].dst is None:
            lab = new_label(tn, "node")
            data = np.array([1.0, 0.0] if ch == "0" else [0.0, 1.0], dtype=np.complex128)
            tn.nodes[lab] = Node([oidx], [2], lab)
            tn.node_layers[lab] = -1
            tn.backend.save_tensor_data(lab, data)
            tn.edges[oidx].dst = lab


def network_from_gates(n, gates, decompose=False, backend=None):
    """src/layer3.jl:544-573 without qiskit: gates = [(U, [q0, q1, ...] 0-based)]."""
    tn = new_network(n, backend)
    for U, qs in gates:
        add_gate(tn, gate_matrix_to_tensor(U), [q + 1 for q in qs], decompose=decompose)
    return tn


# --------------------------------------------------------------------------- #
# Contraction drivers (src/layer2.jl)
# --------------------------------------------------------------------------- #


def _unique(seq):
    out = []
    for x in seq:
        if x not in out:
            out.append(x)
    return out


def sort_indices(A, B):
    """src/layer2.jl:209-216 (Julia intersect / setdiff / union keep first order)."""
    common = _unique([x for x in A.indices if x in B.indices])
    rem_a = _unique([x for x in A.indices if x not in common])
    rem_b = _unique([x for x in B.indices if x not in common])
    uncommon = _unique(rem_a + rem_b)
    return common, uncommon


def create_ncon_indices(A, B, common, uncommon):
    """src/layer2.jl:226-240."""
    m = {x: i + 1 for i, x in enumerate(common)}
    m.update({x: -(i + 1) for i, x in enumerate(uncommon)})
    return [m[x] for x in A.indices], [m[x] for x in B.indices]


def contract_pair(tn, a, b=None):
    """src/layer2.jl:329-405 (edge form when b is None)."""
    if b is None:
        if a in tn.edges:
            e = tn.edges[a]
            return contract_pair(tn, e.src, e.dst)
        return None
    if a == b:
        return None
    A, B = tn.nodes[a], tn.nodes[b]
    common, remaining = sort_indices(A, B)
    ia, ib = create_ncon_indices(A, B, common, remaining)
    dmap = dict(zip(A.indices + B.indices, A.dims + B.dims))
    cdims = [dmap[x] for x in common]
    odims = [dmap[x] for x in remaining]
    record_compute_costs(tn.backend, odims, cdims)
    c = new_label(tn, "node")
    tn.nodes[c] = Node(list(remaining), odims, c)
    for x in common:
        del tn.edges[x]
    for x in remaining:
        e = tn.edges[x]
        if e.src in (a, b):
            e.src = c
        elif e.dst in (a, b):
            e.dst = c
    del tn.nodes[a]
    del tn.nodes[b]
    tn.backend.contract_tensors(a, ia, b, ib, c)
    return c


def _finish(tn, label, output_shape, order_source):
    node = tn.nodes[label]
    if len(node.indices) != 0:
        order = [node.indices.index(x) + 1 for x in order_source if x in node.indices]
        tn.backend.permute_tensor(label, order)
        if output_shape == "vector":
            tn.backend.reshape_tensor(label, [list(range(1, len(node.indices) + 1))])
        elif output_shape != "":
            tn.backend.reshape_tensor(label, output_shape)
    tn.backend.save_output(label)
    return label


def full_wavefunction_contraction(tn, output_shape=""):
    """src/layer2.jl:132-195."""
    inputs = [tn.edges[e].src for e in tn.input_qubits]
    if any(x is None for x in inputs):
        raise ValueError("Please create nodes for each input before using wavefunction contraction")
    wf = inputs[0]
    for x in inputs[1:]:
        wf = contract_pair(tn, wf, x)
    layer_nodes: dict[int, list] = {}
    for k, v in tn.node_layers.items():
        layer_nodes.setdefault(v, []).append(k)
    layers = sorted(k for k in layer_nodes if k > 0)
    if -1 in layer_nodes:
        layers.append(-1)
    for layer in layers:
        nodes = layer_nodes[layer]
        n = nodes[0]
        for n2 in nodes[1:]:
            n = contract_pair(tn, n, n2)
        wf = contract_pair(tn, wf, n)
    node = tn.nodes[wf]
    order_src = [tn.output_qubits[q - 1] for q in tn.qubit_ordering]
    if len(node.indices) != 0:
        order = [node.indices.index(x) + 1 for x in order_src]
        tn.backend.permute_tensor(wf, order)
        if output_shape == "vector":
            tn.backend.reshape_tensor(wf, [list(range(1, len(node.indices) + 1))])
        elif output_shape != "":
            tn.backend.reshape_tensor(wf, output_shape)
    tn.backend.save_output(wf)
    return wf


def contract_network(tn, plan, output_shape=""):
    """src/layer2.jl:250-321 (edge plan: list of str; pair plan: list of pairs).

    Open outputs are permuted into output_qubits order; a partially closed
    output (some qubits open) keeps the open ones in that order.
    """
    if plan and isinstance(plan[0], (list, tuple)):
        for a, b in plan:
            contract_pair(tn, a, b)
    else:
        for e in plan:
            contract_pair(tn, e)
        while len(tn.nodes) > 1:
            it = iter(tn.nodes)
            n1 = next(it)
            n2 = next(it)
            contract_pair(tn, n1, n2)
    out = f"node_{tn.counters['node']}"
    node = tn.nodes[out]
    if len(node.indices) != 0:
        order = [node.indices.index(x) + 1 for x in tn.output_qubits if x in node.indices]
        node.indices[:] = [x for x in tn.output_qubits if x in node.indices]
        tn.backend.permute_tensor(out, order)
        if output_shape == "vector":
            tn.backend.reshape_tensor(out, [list(range(1, len(node.indices) + 1))])
        elif output_shape != "":
            tn.backend.reshape_tensor(out, output_shape)
    tn.backend.save_output(out)
    return out


def random_contraction_plan(tn, rng):
    """src/layer2.jl:119-122 (closed edges, shuffled)."""
    closed = [k for k, v in tn.edges.items() if v.src is not None and v.dst is not None]
    perm = rng.permutation(len(closed))
    return [closed[i] for i in perm]


# --------------------------------------------------------------------------- #
# Slicing (src/layer2/slicing.jl)
# --------------------------------------------------------------------------- #


def multi_index_partition(dims, number_partitions, partition):
    """src/layer2/slicing.jl:12-26; returns the 1-based CartesianIndex tuple."""
    total, num = 1, 1
    while num < len(dims):
        total *= dims[num - 1]
        if total == number_partitions:
            break
        elif total > number_partitions:
            raise ValueError("Partitions and product of dimensions must match")
        num += 1
    sub = dims[:num]
    p = partition - 1
    out = []
    for d in sub:                       # CartesianIndices: first index fastest
        out.append(p % d + 1)
        p //= d
    return tuple(out)


def partition_network_on_virtual_bonds(tn, number_partitions, partition):
    """src/layer2/slicing.jl:37-57 (stable sort of virtual bonds by dimension)."""
    vb = [(k, e) for k, e in tn.edges.items() if e.virtual]
    dims = []
    for k, e in vb:
        data = tn.backend.load_tensor_data(e.src)
        pos = [i for i, x in enumerate(tn.nodes[e.src].indices) if x == k]
        dims.append(int(np.prod([data.shape[i] for i in pos])))
    order = sorted(range(len(vb)), key=lambda i: (dims[i], i))
    bonds = [vb[i][0] for i in order]
    dims = [dims[i] for i in order]
    assert len(dims) > 0, "There must be some virtual bonds, try turning on decompose"
    ci = multi_index_partition(dims, number_partitions, partition)
    return bonds[: len(ci)], ci


def replace_with_view(tn, node_label, bond_label, bond_range):
    """src/layer2/slicing.jl:67-93 (Node.dims copied unchanged, as the reference does)."""
    node = tn.nodes[node_label]
    bond_idx = node.indices.index(bond_label) + 1
    view = new_label(tn, "node")
    tn.nodes[view] = Node(node.indices, node.dims, view)
    tn.node_layers[view] = tn.node_layers[node_label]
    for x in node.indices:
        e = tn.edges[x]
        if e.src == node_label:
            e.src = view
        else:
            e.dst = view
    tn.backend.view_tensor(view, node_label, bond_idx, bond_range)
    tn.backend.delete_tensor(node_label)
    del tn.nodes[node_label]
    del tn.node_layers[node_label]


def slice_tensor_network(tn, bond_labels, bond_values):
    """src/layer2/slicing.jl:100-110."""
    for b, v in zip(bond_labels, bond_values):
        n1, n2 = tn.edges[b].src, tn.edges[b].dst
        replace_with_view(tn, n1, b, (v, v))
        replace_with_view(tn, n2, b, (v, v))


# --------------------------------------------------------------------------- #
# Boltzmann greedy planner, tau -> 0 limit (src/layer2/bgreedy_contraction.jl)
# --------------------------------------------------------------------------- #


def contraction_cost(A, B, alpha):
    """src/layer2/bgreedy_contraction.jl:76-94."""
    _, C_idx = sort_indices(A, B)
    a_open = [A.dims[i] for i in range(len(A.dims)) if A.indices[i] in C_idx]
    b_open = [B.dims[i] for i in range(len(B.dims)) if B.indices[i] in C_idx]
    C_dims = a_open + b_open
    sC = int(np.prod(C_dims)) if C_dims else 1
    sA = int(np.prod(A.dims)) if A.dims else 1
    sB = int(np.prod(B.dims)) if B.dims else 1
    return sC - alpha * (sA + sB), C_idx, C_dims, (math.sqrt(sA * sB * sC), sC)


def greedy_plan(tn, alpha=1.0):
    """bgreedy (src/layer2/bgreedy_contraction.jl:17-45, :105-131) with tau -> 0:
    the minimum-energy pair, first in (i < j) scan order on ties.  O(N^3)."""
    tensors = OrderedDict((k, Node(list(v.indices), list(v.dims), k)) for k, v in tn.nodes.items())
    plan = []
    counter = tn.counters["node"]
    time_cost, space_cost = 0.0, 0
    while len(tensors) > 1:
        items = list(tensors.items())
        best = None
        for i in range(len(items)):
            for j in range(i + 1, len(items)):
                e, ci, cd, costs = contraction_cost(items[i][1], items[j][1], alpha)
                if best is None or e < best[0]:
                    best = (e, items[i][0], items[j][0], ci, cd, costs)
        _, a, b, ci, cd, costs = best
        counter += 1
        c = f"node_{counter}"
        del tensors[a]
        del tensors[b]
        tensors[c] = Node(ci, cd, c)
        plan.append([a, b])
        time_cost += costs[0]
        space_cost = max(space_cost, costs[1])
    return plan, time_cost, space_cost


# --------------------------------------------------------------------------- #
# Optimal contraction cost, exhaustive (the quantity netcon minimises:
# src/layer2/netcon_contraction.jl:14-41 via TensorOperations.optimaltree)
# --------------------------------------------------------------------------- #


def optimal_contraction_cost(tensors, outer=True):
    """Minimum over ALL binary contraction trees of the total multiply-adds,
    each step costing prod(dims of A's and B's indices, shared ones once) --
    Metrics.flops (src/backends.jl:37-38).  outer=False admits only steps
    whose operands share an index (for connected networks).  tensors:
    [(indices, dims)].  Dynamic programming over every subset and every
    split; exponential, for a handful of tensors only."""
    n = len(tensors)
    dim = {}
    count = {}
    for idx, dims in tensors:
        for i, d in zip(idx, dims):
            dim[i] = d
            count[i] = count.get(i, 0) + 1
    members = {}
    for t, (idx, _) in enumerate(tensors):
        for i in idx:
            members.setdefault(i, set()).add(t)

    def open_set(mask):
        ts = {t for t in range(n) if mask >> t & 1}
        return {i for t in ts for i in tensors[t][0] if not members[i] <= ts or count[i] == 1}

    opened = {m: open_set(m) for m in range(1, 1 << n)}
    best = {1 << t: 0 for t in range(n)}
    for m in range(1, 1 << n):   # a subset's splits are smaller subsets: ascending order works
        if m & (m - 1) == 0:
            continue
        low = m & -m
        b = None
        sub = (m - 1) & m
        while sub:
            rest = m ^ sub
            if sub & low and best[sub] is not None and best[rest] is not None and (
                    outer or opened[sub] & opened[rest]):     # each split once
                macs = math.prod(dim[i] for i in opened[sub] | opened[rest])
                c = best[sub] + best[rest] + macs
                if b is None or c < b:
                    b = c
            sub = (sub - 1) & m
        best[m] = b
    return best[(1 << n) - 1]


# --------------------------------------------------------------------------- #
# Circuits (gate lists) and brute-force state-vector simulation
# --------------------------------------------------------------------------- #


def qft_gates(n):
    """src/algorithms.jl:14-32: H, CP(pi/2^(j-i+1)) on (j, i-1), then SWAPs (0-based qubits)."""
    gates = []
    for i in range(1, n + 1):
        gates.append((H, [i - 1]))
        for j in range(i, n):
            gates.append((cphase(math.pi / 2 ** (j - i + 1)), [j, i - 1]))
    for i in range(1, n // 2 + 1):
        gates.append((SWAP, [i - 1, n - i]))
    return gates


GRID_PATTERNS = "ACBD"


def grid_couplers(rows, cols, pattern):
    """Nearest-neighbour CZ couplers of one pattern; qubit q = r*cols + c.
    A/B: horizontal pairs starting at even/odd column; C/D: vertical pairs
    starting at even/odd row."""
    out = []
    if pattern in "AB":
        par = 0 if pattern == "A" else 1
        for r in range(rows):
            for c in range(par, cols - 1, 2):
                out.append((r * cols + c, r * cols + c + 1))
    else:
        par = 0 if pattern == "C" else 1
        for r in range(par, rows - 1, 2):
            for c in range(cols):
                out.append((r * cols + c, (r + 1) * cols + c))
    return out


def grid_circuit_gates(rows, cols, depth, seed):
    """Random grid circuit: per cycle one of {sqrtX, sqrtY, sqrtW} on every qubit
    (drawn with numpy PCG64, never the gate the qubit had in the previous cycle),
    then CZ on the couplers of pattern 'ACBD'[cycle % 4]."""
    rng = np.random.Generator(np.random.PCG64(seed))
    n = rows * cols
    sq = [SQRT_X, SQRT_Y, SQRT_W]
    prev = [-1] * n
    gates = []
    for t in range(depth):
        for q in range(n):
            if prev[q] < 0:
                g = int(rng.integers(0, 3))
            else:
                g = [x for x in range(3) if x != prev[q]][int(rng.integers(0, 2))]
            prev[q] = g
            gates.append((sq[g], [q]))
        for a, b in grid_couplers(rows, cols, GRID_PATTERNS[t % 4]):
            gates.append((CZ, [a, b]))
    return gates


def apply_gate_sv(psi, n, U, qs):
    """Dense state vector; qubit q is bit q of the amplitude index."""
    k = len(qs)
    t = psi.reshape((2,) * n, order="F")          # axis q <-> qubit q
    Ut = U.reshape((2,) * (2 * k), order="F")      # Ut[o_0..o_{k-1}, i_0..i_{k-1}]
    t = np.tensordot(Ut, t, axes=(list(range(k, 2 * k)), qs))
    # result axes: o_0..o_{k-1}, then remaining qubits in order
    rest = [q for q in range(n) if q not in qs]
    perm_src = list(qs) + rest
    inv = [perm_src.index(q) for q in range(n)]
    t = np.transpose(t, inv)
    return t.reshape(-1, order="F")


def statevector(n, gates, init_bits=None):
    psi = np.zeros(2 ** n, dtype=np.complex128)
    idx = 0
    if init_bits is not None:
        for q, b in enumerate(init_bits):
            idx |= int(b) << q
    psi[idx] = 1.0
    for U, qs in gates:
        psi = apply_gate_sv(psi, n, U, qs)
    return psi


def bit_reverse(v, n):
    return int(format(v, f"0{n}b")[::-1], 2) if n > 0 else 0


def qft_closed_form(n, x):
    """QFT of basis |x> (qubit 1 = least significant bit of x and of y), with
    the SWAP layer: amp[y] = exp(+2 pi i rev(x) rev(y) / 2^n) / 2^(n/2), where
    rev reverses n bits -- the textbook QFT read with qubit 1 as the most
    significant bit, as the reference's QFT-vs-ifft test reads it
    (test/algorithms_tests.jl:39-82, big_endian=false)."""
    N = 2 ** n
    yr = np.array([bit_reverse(v, n) for v in range(N)], dtype=np.int64)
    xr = bit_reverse(x, n)
    ph = (xr * yr) % N
    return np.exp(2j * np.pi * ph / N) / math.sqrt(N)


def bits_to_config(bits):
    return "".join("1" if b else "0" for b in bits)


# -- MPS and compression (src/layer2.jl:423-703, src/mps.jl) -------------------

def inneighbours(tn, node_label):
    """src/layer3.jl:324-336: nodes feeding node_label through non-virtual edges."""
    out = []
    for ix in tn.nodes[node_label].indices:
        e = tn.edges[ix]
        if e.src is not None and e.src != node_label and not e.virtual:
            out.append(e.src)
    return out


def decompose_tensor_network(tn, node_label, left_indices, right_indices, threshold=1e-13, max_rank=0,
                             left_label=None, right_label=None):
    """decompose_tensor! (src/layer2.jl:482-560): split a node into B (left
    indices + new virtual bond) and C (bond + right indices) and rewire."""
    node = tn.nodes[node_label]
    pos = {v: k + 1 for k, v in enumerate(node.indices)}
    lp = [pos[x] for x in left_indices]
    rp = [pos[x] for x in right_indices]
    b_label = new_label(tn, "node") if left_label is None else left_label
    c_label = new_label(tn, "node") if right_label is None else right_label
    chi = tn.backend.decompose_tensor(node_label, lp, rp, threshold=threshold, max_rank=max_rank,
                                      left_label=b_label, right_label=c_label)
    b_dims = [node.dims[i - 1] for i in lp]
    c_dims = [node.dims[i - 1] for i in rp]
    if chi > 0:
        vdim = chi
    else:
        vdim = min(math.prod(b_dims), math.prod(c_dims))
        if max_rank > 0:
            vdim = min(vdim, max_rank)
    index_label = new_label(tn, "index")
    tn.nodes[b_label] = Node(list(left_indices) + [index_label], b_dims + [vdim], b_label)
    tn.nodes[c_label] = Node([index_label] + list(right_indices), [vdim] + c_dims, c_label)
    for ix in left_indices:
        e = tn.edges[ix]
        if e.src == node_label:
            e.src = b_label
        elif e.dst == node_label:
            e.dst = b_label
    for ix in right_indices:
        e = tn.edges[ix]
        if e.src == node_label:
            e.src = c_label
        elif e.dst == node_label:
            e.dst = c_label
    tn.edges[index_label] = Edge(b_label, c_label, None, True)
    # the backend verb has already deleted the split tensor's data
    if node_label not in (b_label, c_label):
        del tn.nodes[node_label]
    return b_label, c_label


def compress_bond(tn, node_1, node_2, threshold=1e-13, max_rank=0):
    """compress_bond! (src/layer2.jl:451-468): contract the pair, split it again."""
    left = tn.nodes[node_1]
    right = tn.nodes[node_2]
    left_indices = [x for x in left.indices if x not in right.indices]
    right_indices = [x for x in right.indices if x not in left.indices]
    combined = contract_pair(tn, node_1, node_2)
    return decompose_tensor_network(tn, combined, left_indices, right_indices, threshold=threshold,
                                    max_rank=max_rank, left_label=node_1, right_label=node_2)


def compress_tensor_chain(tn, nodes, threshold=1e-13, max_rank=0):
    """compress_tensor_chain! (src/layer2.jl:423-441): forward then backward pass."""
    for i in range(len(nodes) - 1):
        compress_bond(tn, nodes[i], nodes[i + 1], threshold, max_rank)
    for i in range(len(nodes) - 2, -1, -1):
        compress_bond(tn, nodes[i], nodes[i + 1], threshold, max_rank)


def _gate_layers(tn):
    layer_nodes = {}
    for k, v in tn.node_layers.items():
        layer_nodes.setdefault(v, []).append(k)
    layers = sorted(k for k in layer_nodes if k > 0)
    if -1 in layer_nodes:
        layers.append(-1)
    return layers, layer_nodes


def contract_mps_tensor_network_circuit(tn, threshold=1e-13, max_rank=0, pbc=False):
    """contract_mps_tensor_network_circuit! (src/layer2.jl:570-625): absorb each
    gate layer into the MPS nodes, then compress the bonds between the
    updated neighbours."""
    mps_nodes = [tn.edges[x].src for x in tn.input_qubits]
    if any(x is None for x in mps_nodes):
        raise ValueError("Input qubit values must be set")
    layers, layer_nodes = _gate_layers(tn)
    for layer in layers:
        updated = []
        for node in layer_nodes[layer]:
            input_node = inneighbours(tn, node)[0]
            idx = mps_nodes.index(input_node)
            mps_nodes[idx] = contract_pair(tn, input_node, node)
            updated.append(idx)
        updated.sort()
        for i in range(len(updated) - 1):
            dist = updated[i + 1] - updated[i]
            if pbc:
                dist = min(dist, updated[i] + len(mps_nodes) - updated[i + 1])
            if dist != 1:
                raise ValueError("Gates between non-neighboring qubits")
            compress_bond(tn, mps_nodes[updated[i]], mps_nodes[updated[i + 1]], threshold, max_rank)
    for node in mps_nodes:
        tn.backend.save_output(node, node)
    return mps_nodes


def calculate_mps_amplitudes(tn, mps_nodes, result="result"):
    """calculate_mps_amplitudes! (src/layer2.jl:634-645)."""
    out = mps_nodes[0]
    for node in mps_nodes[1:]:
        out = contract_pair(tn, out, node)
    tn.backend.permute_tensor(out, list(tn.qubit_ordering))
    tn.backend.reshape_tensor(out, [list(range(1, len(mps_nodes) + 1))])
    tn.backend.save_output(out, result)
    return out


def contract_tensor_network_circuit_with_compression(tn, threshold=1e-13, max_rank=0):
    """contract_tensor_network_circuit_with_compression! (src/layer2.jl:655-703):
    as the MPS driver, compressing consecutive updated nodes without the
    neighbour check."""
    nodes = [tn.edges[x].src for x in tn.input_qubits]
    if any(x is None for x in nodes):
        raise ValueError("Input qubit values must be set")
    layers, layer_nodes = _gate_layers(tn)
    for layer in layers:
        updated = []
        for node in layer_nodes[layer]:
            input_node = inneighbours(tn, node)[0]
            idx = nodes.index(input_node)
            nodes[idx] = contract_pair(tn, input_node, node)
            updated.append(idx)
        for i in range(len(updated) - 1):
            compress_bond(tn, nodes[updated[i]], nodes[updated[i + 1]], threshold, max_rank)
    for node in nodes:
        tn.backend.save_output(node, node)
    return nodes


def mps_amplitude(tn, mps_nodes, conf):
    """MPSState(...)[conf] (src/mps.jl:25-55, 87-110): select each node's output
    index at a configuration bit and contract the chain left to right.  The
    bit read for node k is the one the reference's ordering names:
    ordering[qubit_ordering[k]] = (position of node k's output index within
    node k) (src/mps.jl:51-54), kept as written for parity; it is node k's
    own bit only where that position equals k."""
    n = len(mps_nodes)
    sel_nodes, sel_data, out_pos = [], [], []
    for label in mps_nodes:
        node = tn.nodes[label]
        p = next(k for k, x in enumerate(node.indices) if x in tn.output_qubits)
        other = [k for k in range(len(node.indices)) if k != p]
        t = np.transpose(np.asarray(tn.backend.load_tensor_data(label)), [p] + other)
        sel_nodes.append(Node([node.indices[k] for k in other], [node.dims[k] for k in other], label))
        sel_data.append((t[0], t[1]))
        out_pos.append(p)
    ordering = [0] * n
    for i, e in zip(out_pos, tn.qubit_ordering):
        ordering[e - 1] = i + 1          # 1-based position, as in the reference
    bits = [0 if ch == "0" else 1 for ch in conf]
    n1 = sel_nodes[0]
    d1 = sel_data[0][bits[ordering[0] - 1]]
    for idx in range(1, n):
        n2 = sel_nodes[idx]
        common, remaining = sort_indices(n1, n2)
        ia, ib = create_ncon_indices(n1, n2, common, remaining)
        d1 = contract_tensors_data(d1, ia, sel_data[idx][bits[ordering[idx] - 1]], ib)
        n1 = Node(remaining, [1] * len(remaining), "n1")
    return complex(np.asarray(d1).reshape(-1)[0])
