"""Pins for the CPU oracle (oracle/pq_oracle.py) -- each check is chosen so a
plausible mistake (wrong index order, transposed gate, dropped term, wrong
sign/bit convention) fails it.  Sources: the reference's fixtures and
known-answer tests, closed forms, and brute-force state vectors."""
import json
import math
from collections import OrderedDict

import numpy as np
import pytest

from oracle import pq_oracle as O


def _load_old_schema(path):
    """examples/ghz_3*.json use the early schema: node data inline as
    data_re/data_im (column-major) with data_dims; edges without qubit/virtual."""
    d = json.load(open(path), object_pairs_hook=OrderedDict)
    tn = O.new_network(d["number_qubits"])
    tn.edges = OrderedDict((k, O.Edge(v["src"], v["dst"], None, False)) for k, v in d["edges"].items())
    tn.nodes = OrderedDict()
    for k, v in d["nodes"].items():
        dims = v["data_dims"]
        data = (np.array(v["data_re"]) + 1j * np.array(v["data_im"])).reshape(dims, order="F")
        tn.nodes[k] = O.Node(list(v["indices"]), list(dims), k)
        tn.backend.save_tensor_data(k, data)
        tn.counters["node"] = max(tn.counters["node"], int(k.split("_")[-1]))
    tn.input_qubits = list(d["input_qubits"])
    tn.output_qubits = list(d["output_qubits"])
    return tn, d


def test_ghz3_fixture_contraction(golden_dir):
    """examples/ghz_3.json contracted along examples/ghz_3_plan.json must equal
    examples/ghz_3_contracted.json: index order and every value."""
    tn, _ = _load_old_schema(golden_dir / "ghz_3.json")
    plan = json.load(open(golden_dir / "ghz_3_plan.json"))
    for e in plan:
        O.contract_pair(tn, e)
    ref = json.load(open(golden_dir / "ghz_3_contracted.json"), object_pairs_hook=OrderedDict)
    (lab, node), = ref["nodes"].items()
    assert list(tn.nodes.keys()) == [lab]
    assert tn.nodes[lab].indices == node["indices"]
    want = (np.array(node["data_re"]) + 1j * np.array(node["data_im"])).reshape(node["data_dims"], order="F")
    got = tn.backend.load_tensor_data(lab)
    assert got.shape == tuple(node["data_dims"])
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-15)


def test_ghz3_fixture_rejects_transposed_result(golden_dir):
    """The fixture distinguishes index orders: swapping A and B changes it."""
    tn, _ = _load_old_schema(golden_dir / "ghz_3.json")
    O.contract_pair(tn, "node_2", "node_1")
    O.contract_pair(tn, "node_3", "node_4")
    ref = json.load(open(golden_dir / "ghz_3_contracted.json"), object_pairs_hook=OrderedDict)
    node = ref["nodes"]["node_5"]
    assert tn.nodes["node_5"].indices != node["indices"]


def test_metrics_pin():
    """test/layer2_tests.jl:127-132: h q0; cx 0,1; cx 0,1; cx 0,2 -> 8 / 44 / 124."""
    g = [(O.H, [0]), (O.CX, [0, 1]), (O.CX, [0, 1]), (O.CX, [0, 2])]
    tn = O.network_from_gates(3, g)
    O.add_input(tn, "000")
    O.full_wavefunction_contraction(tn, "vector")
    m = tn.backend.metrics
    assert (m.max_tensor_size, m.total_space_allocated, m.flops) == (8, 44, 124)
    assert len(tn.nodes) == 1


def test_hh_vector_pin():
    """test/layer2_tests.jl:70-100: h q0; h q1 from |00>, random edge plan, every amplitude 1/2."""
    tn = O.network_from_gates(2, [(O.H, [0]), (O.H, [1])])
    O.add_input(tn, "00")
    plan = O.random_contraction_plan(tn, np.random.default_rng(3))
    O.contract_network(tn, plan, "vector")
    r = tn.backend.load_tensor_data("result")
    assert r.ndim == 1 and len(tn.nodes) == 1
    np.testing.assert_allclose(r, 0.5, atol=1e-15)


def test_ghz3_amplitude_pin():
    """test/layer1_tests.jl:11-46: <000| CX(1,2) CX(0,1) H(0) |000> = 1/sqrt(2)."""
    g = [(O.H, [0]), (O.CX, [0, 1]), (O.CX, [1, 2])]
    for seed in range(4):
        tn = O.network_from_gates(3, g)
        O.add_input(tn, "000")
        O.add_output(tn, "000")
        O.contract_network(tn, O.random_contraction_plan(tn, np.random.default_rng(seed)))
        r = tn.backend.load_tensor_data("result")
        assert r.shape == ()
        assert abs(r - 1 / math.sqrt(2)) < 1e-15


def test_decompose_round_trip():
    """test/layer3_tests.jl:57-71: decompose a random 2-qubit tensor, re-contract,
    permute (1,3,2,4) back: equal to the original."""
    rng = np.random.default_rng(0)
    gd = rng.standard_normal((2, 2, 2, 2)) + 1j * rng.standard_normal((2, 2, 2, 2))
    tn = O.new_network(2)
    labels = O.add_gate(tn, gd, [1, 2], decompose=True)
    res = O.contract_pair(tn, *labels)
    data = np.transpose(tn.backend.load_tensor_data(res), (0, 2, 1, 3))
    np.testing.assert_allclose(data, gd, atol=1e-13)


def test_gate_tensors_closed_forms():
    for U in (O.SQRT_X, O.SQRT_Y, O.SQRT_W, O.H, O.CZ, O.SWAP, O.cphase(0.3)):
        np.testing.assert_allclose(U @ U.conj().T, np.eye(U.shape[0]), atol=1e-15)
    np.testing.assert_allclose(O.SQRT_X @ O.SQRT_X, O.X, atol=1e-15)
    np.testing.assert_allclose(O.SQRT_Y @ O.SQRT_Y, O.Y, atol=1e-15)
    np.testing.assert_allclose(O.SQRT_W @ O.SQRT_W, O.W, atol=1e-15)
    # G[in, out] = U[out, in] for a one-qubit gate (src/layer3.jl:566-568)
    G = O.gate_matrix_to_tensor(O.SQRT_Y)
    assert G[0, 1] == O.SQRT_Y[1, 0] and G[1, 0] == O.SQRT_Y[0, 1]
    # CX with control = first target: in (c=1, t=0) -> out (1, 1)
    G = O.gate_matrix_to_tensor(O.CX)
    assert G[1, 0, 1, 1] == 1 and G[1, 0, 1, 0] == 0


@pytest.mark.parametrize("n", [3, 5, 8])
def test_qft_closed_form_and_brute_force(n):
    """QFT (src/algorithms.jl:14-32) of a basis state: the TN state equals the
    brute-force state vector and the closed form exp(2 pi i rev(x) rev(y)/2^n)/2^(n/2)."""
    rng = np.random.default_rng(n)
    x = int(rng.integers(0, 2 ** n))
    bits = [(x >> q) & 1 for q in range(n)]
    gates = O.qft_gates(n)
    tn = O.network_from_gates(n, gates)
    O.add_input(tn, O.bits_to_config(bits))
    O.full_wavefunction_contraction(tn, "vector")
    psi = tn.backend.load_tensor_data("result")
    sv = O.statevector(n, gates, bits)
    np.testing.assert_allclose(psi, sv, atol=1e-13)
    np.testing.assert_allclose(psi, O.qft_closed_form(n, x), atol=1e-13)


def test_qft_closed_form_rejects_conjugate():
    n, x = 4, 3
    bits = [(x >> q) & 1 for q in range(n)]
    sv = O.statevector(n, O.qft_gates(n), bits)
    assert np.abs(sv - np.conj(O.qft_closed_form(n, x))).max() > 0.1


def test_qft12_metrics_closed_form():
    """QFT-12 full wavefunction: 95 contractions; MACs = sum_{k=2..n} 2^k +
    sum_gates 2^(n + arity) (every gate is contracted into the full state)."""
    n = 12
    gates = O.qft_gates(n)
    tn = O.network_from_gates(n, gates)
    O.add_input(tn, "0" * n)
    O.full_wavefunction_contraction(tn, "vector")
    want = sum(2 ** k for k in range(2, n + 1)) + sum(2 ** (n + len(q)) for _, q in gates)
    assert tn.backend.metrics.flops == want == 1286140
    assert tn.backend.metrics.max_tensor_size == 4096


def test_statevector_rejects_transposed_gate():
    """Brute force is sensitive to U vs U^T (sqrtY is not symmetric)."""
    g = [(O.H, [0]), (O.SQRT_Y, [0])]
    a = O.statevector(1, g)
    b = O.statevector(1, [(O.H, [0]), (O.SQRT_Y.T, [0])])
    assert np.abs(a - b).max() > 0.1


def test_grid_amplitudes_brute_force_and_plan_invariance():
    """Closed 3x3 grid circuit amplitudes: two different plans (random edge
    order and tau->0 greedy) agree with the brute-force state vector."""
    gates = O.grid_circuit_gates(3, 3, 6, seed=7)
    sv = O.statevector(9, gates)
    rng = np.random.default_rng(1)
    for trial in range(3):
        bits = [int(b) for b in rng.integers(0, 2, 9)]
        idx = sum(b << q for q, b in enumerate(bits))
        tn = O.network_from_gates(9, gates, decompose=True)
        O.add_input(tn, "0" * 9)
        O.add_output(tn, O.bits_to_config(bits))
        O.contract_network(tn, O.random_contraction_plan(tn, rng))
        a1 = tn.backend.load_tensor_data("result")
        tn = O.network_from_gates(9, gates, decompose=False)
        O.add_input(tn, "0" * 9)
        O.add_output(tn, O.bits_to_config(bits))
        plan, _, _ = O.greedy_plan(tn)
        O.contract_network(tn, plan)
        a2 = tn.backend.load_tensor_data("result")
        assert abs(a1 - sv[idx]) < 1e-13 and abs(a2 - sv[idx]) < 1e-13


def test_open_outputs_batch():
    """Leaving qubits open ('x') gives the amplitudes of all completions, in
    output-qubit order (qubit with the lowest number = fastest index)."""
    gates = O.grid_circuit_gates(2, 3, 5, seed=2)
    sv = O.statevector(6, gates)
    cfg = "1x0x1x"
    tn = O.network_from_gates(6, gates, decompose=True)
    O.add_input(tn, "0" * 6)
    O.add_output(tn, cfg)
    O.contract_network(tn, O.random_contraction_plan(tn, np.random.default_rng(0)), "vector")
    r = tn.backend.load_tensor_data("result")
    open_q = [i for i, ch in enumerate(cfg) if ch == "x"]
    for j in range(8):
        bits = [int(ch) if ch != "x" else 0 for ch in cfg]
        for t, q in enumerate(open_q):
            bits[q] = (j >> t) & 1
        assert abs(r[j] - sv[sum(b << q for q, b in enumerate(bits))]) < 1e-13


def test_slicing_sum_invariance():
    """test/layer2_tests.jl:419-455: sum of the 4 partitions == unsliced state."""
    n = 4
    gates = O.grid_circuit_gates(2, 2, 4, seed=11) + O.qft_gates(n)
    tn = O.network_from_gates(n, gates, decompose=True)
    O.add_input(tn, "0" * n)
    O.full_wavefunction_contraction(tn, "vector")
    wf = tn.backend.load_tensor_data("result")
    np.testing.assert_allclose(wf, O.statevector(n, gates), atol=1e-13)
    parts = []
    for p in range(1, 5):
        tn = O.network_from_gates(n, gates, decompose=True)
        O.add_input(tn, "0" * n)
        bonds, vals = O.partition_network_on_virtual_bonds(tn, 4, p)
        O.slice_tensor_network(tn, bonds, vals)
        O.full_wavefunction_contraction(tn, "vector")
        parts.append(tn.backend.load_tensor_data("result"))
    np.testing.assert_allclose(sum(parts), wf, atol=1e-13)
    assert np.abs(parts[0] - wf).max() > 1e-3   # a single slice is not the answer


def test_multi_index_partition():
    """src/layer2/slicing.jl:12-26: CartesianIndices order, first index fastest."""
    assert [O.multi_index_partition((2, 2, 2), 4, p) for p in range(1, 5)] == \
        [(1, 1), (2, 1), (1, 2), (2, 2)]
    assert O.multi_index_partition((2, 3, 2), 6, 6) == (2, 3)
    with pytest.raises(ValueError):
        O.multi_index_partition((2, 3, 2), 4, 1)


def test_greedy_plan_validity_and_ledger():
    """tau->0 bgreedy plan contracts every node exactly once; its sqrt-cost
    equals a re-count from the executed contraction."""
    gates = O.grid_circuit_gates(2, 2, 3, seed=5)
    tn = O.network_from_gates(4, gates, decompose=True)
    O.add_input(tn, "0000")
    O.add_output(tn, "0110")
    plan, tcost, scost = O.greedy_plan(tn)
    leaves = set(tn.nodes)
    used = [x for p in plan for x in p]
    assert len(used) == len(set(used))
    assert leaves <= set(used)
    # re-count sqrt(|A||B||C|) while executing
    recount = 0.0
    for a, b in plan:
        A, B = tn.nodes[a], tn.nodes[b]
        _, rem = O.sort_indices(A, B)
        d = dict(zip(A.indices + B.indices, A.dims + B.dims))
        sc = math.prod(d[x] for x in rem)
        recount += math.sqrt(math.prod(A.dims) * math.prod(B.dims) * sc)
        O.contract_pair(tn, a, b)
    assert abs(recount - tcost) < 1e-9
    assert len(tn.nodes) == 1


def test_grid_generator_spec():
    gates = O.grid_circuit_gates(3, 4, 9, seed=1)
    n = 12
    prev = {}
    sq = [id(O.SQRT_X), id(O.SQRT_Y), id(O.SQRT_W)]
    cycle = 0
    count1 = 0
    for U, qs in gates:
        if len(qs) == 1:
            k = sq.index(id(U))
            if qs[0] in prev:
                assert prev[qs[0]] != k
            prev[qs[0]] = k
            count1 += 1
    assert count1 == n * 9
    assert O.grid_couplers(3, 4, "A") == [(0, 1), (2, 3), (4, 5), (6, 7), (8, 9), (10, 11)]
    assert O.grid_couplers(3, 4, "D") == [(4, 8), (5, 9), (6, 10), (7, 11)]
