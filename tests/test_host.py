"""Host-side bookkeeping of the product (picoquant_hip) against the oracle,
with a recording backend that keeps shapes only (no device, no arithmetic)."""
import math

import numpy as np
import pytest

from oracle import pq_oracle as O
import picoquant_hip as P
from picoquant_hip import circuits as C
from picoquant_hip.backend import Metrics


class RecordingBackend(P.AbstractBackend):
    """Keeps each tensor's shape and logs contract_tensors calls."""

    def __init__(self):
        self.shapes = {}
        self.metrics = Metrics()
        self.calls = []

    def save_tensor_data(self, label, data):
        self.shapes[label] = tuple(np.asarray(data).shape)

    def load_tensor_data(self, label):
        return None

    def tensor_shape(self, label):
        return self.shapes[label]

    def contract_tensors(self, a, ia, b, ib, c):
        sa, sb = self.shapes.pop(a), self.shapes.pop(b)
        out = [d for d, l in zip(sa, ia) if l < 0] + [d for d, l in zip(sb, ib) if l < 0]
        self.shapes[c] = tuple(out)
        self.calls.append((a, list(ia), b, list(ib), c))

    def save_output(self, node, name="result"):
        self.shapes[name] = self.shapes[node]

    def permute_tensor(self, label, axes):
        s = self.shapes[label]
        self.shapes[label] = tuple(s[a - 1] for a in axes)

    def reshape_tensor(self, label, groups):
        s = self.shapes[label]
        self.shapes[label] = tuple(math.prod(s[g - 1] for g in grp) for grp in groups)

    def delete_tensor(self, label):
        self.shapes.pop(label, None)

    def view_tensor(self, view_node, node, bond_idx, bond_range):
        s = list(self.shapes[node])
        s[bond_idx - 1] = bond_range[1] - bond_range[0] + 1
        self.shapes[view_node] = tuple(s)


class OracleRecording(O.InteractiveBackend):
    def __init__(self):
        super().__init__()
        self.calls = []

    def contract_tensors(self, a, ia, b, ib, c):
        self.calls.append((a, list(ia), b, list(ib), c))
        super().contract_tensors(a, ia, b, ib, c)


def _pair(gates, n, decompose, inp, out=None):
    pb = RecordingBackend()
    tp = P.network_from_gates(n, gates, backend=pb, decompose=decompose)
    P.add_input(tp, inp)
    ob = OracleRecording()
    to = O.network_from_gates(n, gates, decompose=decompose, backend=ob)
    O.add_input(to, inp)
    if out is not None:
        P.add_output(tp, out)
        O.add_output(to, out)
    return tp, pb, to, ob


def test_metrics_pin_product():
    """test/layer2_tests.jl:127-132 on the product's host code."""
    g = [(C.H, [0]), (O.CX, [0, 1]), (O.CX, [0, 1]), (O.CX, [0, 2])]
    tp, pb, _, _ = _pair(g, 3, False, "000")
    P.full_wavefunction_contraction(tp, "vector")
    m = pb.metrics
    assert (m.max_tensor_size, m.total_space_allocated, m.flops) == (8, 44, 124)
    assert pb.shapes["result"] == (8,)


@pytest.mark.parametrize("decompose", [False, True])
def test_full_wavefunction_calls_match_oracle(decompose):
    gates = C.qft_gates(5) + C.grid_circuit_gates(1, 5, 3, seed=3)
    tp, pb, to, ob = _pair(gates, 5, decompose, "01101")
    P.full_wavefunction_contraction(tp, "vector")
    O.full_wavefunction_contraction(to, "vector")
    assert pb.calls == ob.calls
    assert pb.metrics == Metrics(**vars(ob.metrics))
    assert list(tp.nodes) == list(to.nodes)
    assert pb.shapes["result"] == ob.tensors["result"].shape


def test_contract_network_and_slicing_match_oracle():
    gates = C.grid_circuit_gates(2, 3, 5, seed=9)
    tp, pb, to, ob = _pair(gates, 6, True, "000000", "10x1x0")
    bp, vp = P.partition_network_on_virtual_bonds(tp, 8, 5)
    bo, vo = O.partition_network_on_virtual_bonds(to, 8, 5)
    assert (bp, vp) == (bo, vo)
    P.slice_tensor_network(tp, bp, vp)
    O.slice_tensor_network(to, bo, vo)
    plan = O.random_contraction_plan(to, np.random.default_rng(4))
    P.contract_network(tp, plan, "vector")
    O.contract_network(to, plan, "vector")
    assert pb.calls == ob.calls
    assert pb.metrics == Metrics(**vars(ob.metrics))
    assert pb.shapes["result"] == ob.tensors["result"].shape == (4,)


def test_edge_and_node_bookkeeping():
    tn = P.TensorNetworkCircuit.create(3, RecordingBackend())
    P.add_gate(tn, C.H.T.reshape(2, 2), [1])
    assert len(tn.nodes) == 1 and len(tn.edges) == 4          # test/layer3_tests.jl:103-120
    P.add_input(tn, "000")
    P.add_output(tn, "000")
    assert len(tn.nodes) == 7
    P.add_input(tn, "000")
    P.add_output(tn, "000")
    assert len(tn.nodes) == 7                                  # test/layer3_tests.jl:170-189


def test_gate_tensor_conversion_matches_oracle():
    for U in (C.SQRT_X, C.SQRT_Y, C.SQRT_W, C.CZ, C.SWAP, C.cphase(0.7), O.CX):
        np.testing.assert_array_equal(P.gate_matrix_to_tensor(U), O.gate_matrix_to_tensor(U))
    g1 = C.grid_circuit_gates(4, 4, 7, seed=123)
    g2 = O.grid_circuit_gates(4, 4, 7, seed=123)
    assert [q for _, q in g1] == [q for _, q in g2]
    assert all(np.array_equal(a, b) for (a, _), (b, _) in zip(g1, g2))
    assert [q for _, q in C.qft_gates(9)] == [q for _, q in O.qft_gates(9)]


def test_multi_index_partition_product():
    assert P.multi_index_partition((2, 2, 2), 4, 3) == O.multi_index_partition((2, 2, 2), 4, 3)
    with pytest.raises(ValueError):
        P.multi_index_partition((2, 3, 2), 4, 1)
