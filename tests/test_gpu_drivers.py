"""The reference's drivers (full_wavefunction_contraction!, contract_network!,
slicing) running on HIPBackend, against the oracle, closed forms and the
committed brute-force amplitudes.

Tolerances (relative amplitude error |a - a_ref| / |a_ref|, worst case):
ComplexF64 -- 1e-12; ComplexF32 -- 1e-5."""
import json
import math

import numpy as np
import pytest

from oracle import pq_oracle as O
import picoquant_hip as P
from picoquant_hip import circuits as C
from picoquant_hip import planner as PL

pytestmark = pytest.mark.gpu

RTOL = {np.complex128: 1e-12, np.complex64: 1e-5}


@pytest.fixture(scope="module", params=[np.complex128, np.complex64], ids=["c128", "c64"])
def dtype(request):
    return request.param


def rel_err(got, want):
    got = np.asarray(got, dtype=np.complex128)
    want = np.asarray(want, dtype=np.complex128)
    return float(np.max(np.abs(got - want) / np.abs(want)))


def test_qft12_full_wavefunction(dtype):
    """Config 1: QFT-12 from a seeded basis state, full_wavefunction_contraction!
    to a 4096-amplitude vector; checked against the oracle and the closed form."""
    n = 12
    x = int(np.random.Generator(np.random.PCG64(12)).integers(0, 2 ** n))
    bits = [(x >> q) & 1 for q in range(n)]
    gates = C.qft_gates(n)
    tn = P.network_from_gates(n, gates, backend=P.HIPBackend(dtype))
    P.add_input(tn, "".join(map(str, bits)))
    P.full_wavefunction_contraction(tn, "vector")
    psi = tn.backend.load_tensor_data("result")
    assert psi.shape == (4096,)
    m = tn.backend.metrics
    assert (m.flops, m.max_tensor_size) == (1286140, 4096)
    assert rel_err(psi, O.qft_closed_form(n, x)) < RTOL[dtype]
    to = O.network_from_gates(n, O.qft_gates(n))
    O.add_input(to, "".join(map(str, bits)))
    O.full_wavefunction_contraction(to, "vector")
    assert rel_err(psi, to.backend.load_tensor_data("result")) < RTOL[dtype]


def test_ghz3_fixture_on_gpu(golden_dir, dtype):
    d = json.load(open(golden_dir / "ghz_3.json"))
    ref = json.load(open(golden_dir / "ghz_3_contracted.json"))
    tn = P.TensorNetworkCircuit.create(3, P.HIPBackend(dtype))
    tn.edges.clear()
    for k, v in d["edges"].items():
        tn.edges[k] = P.Edge(v["src"], v["dst"], None, False)
    for k, v in d["nodes"].items():
        data = (np.array(v["data_re"]) + 1j * np.array(v["data_im"])).reshape(v["data_dims"], order="F")
        tn.nodes[k] = P.Node(list(v["indices"]), list(v["data_dims"]), k)
        tn.backend.save_tensor_data(k, data)
    tn.counters["node"] = 3
    for e in json.load(open(golden_dir / "ghz_3_plan.json")):
        P.contract_pair(tn, e)
    node = ref["nodes"]["node_5"]
    assert tn.nodes["node_5"].indices == node["indices"]
    want = (np.array(node["data_re"]) + 1j * np.array(node["data_im"])).reshape(node["data_dims"], order="F")
    got = tn.backend.load_tensor_data("node_5")
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-15 if dtype == np.complex128 else 1e-7)


def test_reference_pins_on_gpu(dtype):
    """test/layer2_tests.jl:97-100 (H x H -> 1/2), :127-132 (metrics),
    test/layer1_tests.jl:34-39 (GHZ-3 amplitude 1/sqrt(2))."""
    tn = P.network_from_gates(2, [(C.H, [0]), (C.H, [1])], backend=P.HIPBackend(dtype))
    P.add_input(tn, "00")
    P.contract_network(tn, P.random_contraction_plan(tn, np.random.default_rng(0)), "vector")
    r = tn.backend.load_tensor_data("result")
    assert r.shape == (4,) and abs(r[0].real - 0.5) < 1e-7
    g = [(C.H, [0]), (O.CX, [0, 1]), (O.CX, [0, 1]), (O.CX, [0, 2])]
    tn = P.network_from_gates(3, g, backend=P.HIPBackend(dtype))
    P.add_input(tn, "000")
    P.full_wavefunction_contraction(tn, "vector")
    m = tn.backend.metrics
    assert (m.max_tensor_size, m.total_space_allocated, m.flops) == (8, 44, 124)
    tn = P.network_from_gates(3, [(C.H, [0]), (O.CX, [0, 1]), (O.CX, [1, 2])], backend=P.HIPBackend(dtype))
    P.add_input(tn, "000")
    P.add_output(tn, "000")
    P.contract_network(tn, P.random_contraction_plan(tn, np.random.default_rng(1)))
    assert abs(tn.backend.load_tensor_data("result") - 1 / math.sqrt(2)) < 1e-7


def test_slicing_sum_on_gpu(dtype):
    """test/layer2_tests.jl:419-455 on the GPU: sum over 8 slices == unsliced == brute force."""
    n = 5
    gates = C.grid_circuit_gates(1, 5, 6, seed=3) + C.qft_gates(n)
    sv = O.statevector(n, O.grid_circuit_gates(1, 5, 6, seed=3) + O.qft_gates(n))
    tn = P.network_from_gates(n, gates, backend=P.HIPBackend(dtype), decompose=True)
    P.add_input(tn, "0" * n)
    P.full_wavefunction_contraction(tn, "vector")
    wf = tn.backend.load_tensor_data("result")
    assert rel_err(wf, sv) < 10 * RTOL[dtype]
    parts = []
    for p in range(1, 9):
        tn = P.network_from_gates(n, gates, backend=P.HIPBackend(dtype), decompose=True)
        P.add_input(tn, "0" * n)
        b, v = P.partition_network_on_virtual_bonds(tn, 8, p)
        P.slice_tensor_network(tn, b, v)
        P.full_wavefunction_contraction(tn, "vector")
        parts.append(tn.backend.load_tensor_data("result").astype(np.complex128))
    assert rel_err(sum(parts), sv) < 10 * RTOL[dtype]


def test_grid5x5_d12_amplitudes(golden_dir, dtype):
    """Config 2: 5x5 depth-12 grid, single amplitudes <x|C|0> with the sweep
    plan, against the committed brute-force values; plus the open-column batch."""
    g = json.load(open(golden_dir / "grid5x5_d12.json"))
    rows, cols, depth, seed = g["rows"], g["cols"], g["depth"], g["seed"]
    gates = C.grid_circuit_gates(rows, cols, depth, seed)
    errs = []
    for bits, (re, im) in list(zip(g["bitstrings"], g["amplitudes"]))[:3]:
        tn = P.network_from_gates(rows * cols, gates, backend=P.HIPBackend(dtype), decompose=True)
        P.add_input(tn, "0" * (rows * cols))
        P.add_output(tn, bits)
        plan = PL.sweep_plan(tn, PL.grid_site_order(rows, cols))
        assert PL.plan_ledger(tn, plan)["max_size"] <= 2 ** 20
        P.contract_network(tn, plan)
        a = complex(tn.backend.load_tensor_data("result"))
        errs.append(abs(a - complex(re, im)) / abs(complex(re, im)))
    assert max(errs) < RTOL[dtype], errs
    cfg = list(g["batch_base"])
    for q in g["open_qubits_1based"]:
        cfg[q - 1] = "x"
    tn = P.network_from_gates(rows * cols, gates, backend=P.HIPBackend(dtype), decompose=True)
    P.add_input(tn, "0" * (rows * cols))
    P.add_output(tn, "".join(cfg))
    plan = PL.sweep_plan(tn, PL.grid_site_order(rows, cols, open_last=g["open_qubits_1based"]))
    P.contract_network(tn, plan, "vector")
    r = tn.backend.load_tensor_data("result")
    want = np.array([complex(a, b) for a, b in g["batch_amplitudes"]])
    assert rel_err(r, want) < RTOL[dtype]


def test_backend_verbs(dtype):
    be = P.HIPBackend(dtype)
    rng = np.random.default_rng(3)
    T = rng.standard_normal((2, 3, 4)) + 1j * rng.standard_normal((2, 3, 4))
    be.save_tensor_data("t", T)
    np.testing.assert_allclose(be.load_tensor_data("t"), T, rtol=1e-6)
    assert be.load_tensor_data("missing") is None
    be.permute_tensor("t", [3, 1, 2])
    np.testing.assert_allclose(be.load_tensor_data("t"), np.transpose(T, (2, 0, 1)), rtol=1e-6)
    be.reshape_tensor("t", [[1], [2, 3]])
    np.testing.assert_allclose(be.load_tensor_data("t"),
                               np.reshape(np.transpose(T, (2, 0, 1)), (4, 6), order="F"), rtol=1e-6)
    be.view_tensor("v", "t", 2, (2, 4))
    np.testing.assert_allclose(be.load_tensor_data("v"),
                               np.reshape(np.transpose(T, (2, 0, 1)), (4, 6), order="F")[:, 1:4], rtol=1e-6)
    be.save_output("v", "out")
    assert be.load_tensor_data("out").shape == (4, 3)
    G = rng.standard_normal((2, 2, 2, 2)) + 1j * rng.standard_normal((2, 2, 2, 2))
    be.save_tensor_data("g", G)
    chi = be.decompose_tensor("g", [1, 3], [2, 4], left_label="l", right_label="r")
    assert chi == 4
    ob = O.InteractiveBackend()
    ob.save_tensor_data("g", G)
    ob.decompose_tensor("g", [1, 3], [2, 4], left_label="l", right_label="r")
    L, R = be.load_tensor_data("l"), be.load_tensor_data("r")
    back = np.einsum("acx,xbd->abcd", L, R)
    np.testing.assert_allclose(back, G, atol=1e-5)
    be.delete_tensor("g")
    with pytest.raises(KeyError):
        be.contract_tensors("nope", [1], "l", [1, -1, -2], "c")


def test_netcon_contraction(dtype):
    """netcon_contraction! on the reference's netcon test network (H, CX, CX;
    input and output 000, test/layer2_tests.jl:145-170): one node is left and
    <000|GHZ> = 1/sqrt(2); a 3x3 grid amplitude through netcon's plan matches
    the brute-force state vector."""
    from picoquant_hip import optimal_order as NC
    tn = P.network_from_gates(3, [(C.H, [0]), (C.CX, [0, 1]), (C.CX, [0, 2])], backend=P.HIPBackend(dtype))
    P.add_input(tn, "000")
    P.add_output(tn, "000")
    NC.netcon_contraction(tn)
    assert len(tn.nodes) == 1
    amp = complex(np.asarray(tn.backend.load_tensor_data("result")).reshape(-1)[0])
    assert abs(amp - 1 / math.sqrt(2)) < RTOL[dtype]
    bits = "101100110"
    tn = P.network_from_gates(9, C.grid_circuit_gates(3, 3, 1, 7), backend=P.HIPBackend(dtype), decompose=True)
    P.add_input(tn, "0" * 9)
    P.add_output(tn, bits)
    assert len(tn.nodes) <= NC.MAX_NODES
    NC.netcon_contraction(tn)
    amp = complex(np.asarray(tn.backend.load_tensor_data("result")).reshape(-1)[0])
    sv = O.statevector(9, O.grid_circuit_gates(3, 3, 1, 7))
    want = sv[sum(int(b) << q for q, b in enumerate(bits))]
    assert rel_err([amp], [want]) < RTOL[dtype]


def test_qasm_qft10_on_gpu(golden_dir, dtype):
    """examples/qft_10.qasm (qiskit-transpiled u1/u2/cx) through the QASM
    parser and full_wavefunction_contraction! on HIPBackend: the DFT."""
    from picoquant_hip import qasm as Q
    circ = Q.load_qasm_as_circuit_from_file(golden_dir / "qft_10.qasm")
    tn = Q.convert_qiskit_circ_to_network(circ, P.HIPBackend(dtype))
    x = 613
    P.add_input(tn, "".join(str((x >> q) & 1) for q in range(10)))
    P.full_wavefunction_contraction(tn, "vector")
    psi = tn.backend.load_tensor_data("result")
    y = np.arange(1024)
    assert rel_err(psi, np.exp(2j * np.pi * x * y / 1024) / 32) < RTOL[dtype]


@pytest.mark.parametrize("ext", ["npz", "h5"])
def test_dsl_replay_on_gpu(tmp_path, dtype, ext):
    """A contraction recorded by DSLBackend and replayed by execute_dsl_file
    through libpq_hip: the reference's DSL test amplitude (1/sqrt 2,
    test/layer1_tests.jl:11-44) and a QFT-6 wavefunction, with the tensor
    store as .npz and as HDF5 in the reference's layout."""
    from picoquant_hip import dsl as D
    tl, store = tmp_path / "contract_network.tl", tmp_path / f"tensor_data.{ext}"
    be = D.DSLBackend(tl, store, "", True)
    tn = P.network_from_gates(3, [(C.H, [0]), (C.CX, [0, 1]), (C.CX, [1, 2])], backend=be)
    P.add_input(tn, "000")
    P.add_output(tn, "000")
    P.contract_network(tn, P.random_contraction_plan(tn, np.random.default_rng(1)))
    assert D.execute_dsl_file(dtype, tl, store) == [f"node_{tn.counters['node']}"]   # the final node
    amp = complex(np.asarray(be.load_tensor_data("result")).reshape(-1)[0])
    assert abs(amp - 1 / math.sqrt(2)) < RTOL[dtype]
    tl2, store2, out2 = tmp_path / "qft.tl", tmp_path / f"qft.{ext}", tmp_path / f"out.{ext}"
    be = D.DSLBackend(tl2, store2, out2, True)
    n, x = 6, 45
    tn = P.network_from_gates(n, C.qft_gates(n), backend=be)
    P.add_input(tn, "".join(str((x >> q) & 1) for q in range(n)))
    P.full_wavefunction_contraction(tn, "vector")
    D.execute_dsl_file(dtype, tl2, store2)
    psi = be.load_tensor_data("result")
    assert rel_err(psi, O.qft_closed_form(n, x)) < RTOL[dtype]


def test_fp64_norm_of_full_wavefunctions():
    """The FP64 norm check: the full wavefunction of a unitary circuit from a
    basis state has norm 1 (QFT-12, config 1) -- 1e-12 in ComplexF64 -- and a
    5x5 grid circuit's wavefunction contracted through the planned executor in
    the bench's precision split keeps its norm to 1e-5, summed in FP64."""
    from picoquant_hip.executor import compile_plan
    n, x = 12, 2741
    tn = P.network_from_gates(n, C.qft_gates(n), backend=P.HIPBackend(np.complex128))
    P.add_input(tn, "".join(str((x >> q) & 1) for q in range(n)))
    P.full_wavefunction_contraction(tn, "vector")
    psi = np.asarray(tn.backend.load_tensor_data("result"), dtype=np.complex128)
    assert abs(np.vdot(psi, psi).real - 1.0) < 1e-12
    # every output open: 2^16 amplitudes of a 4x4 grid circuit
    tn = P.network_from_gates(16, C.grid_circuit_gates(4, 4, 8, 7), backend=P.HIPBackend(np.complex128),
                              decompose=True)
    P.add_input(tn, "0" * 16)
    plan = PL.reconfigure(tn, PL.sweep_plan(tn, PL.grid_site_order(4, 4, "col")), windows=(8,))
    for prec, tol in (("uniform", 1e-12), ("mixed", 1e-5)):
        p = compile_plan(tn, plan, precision=prec, small_elems=64)
        p.execute(True)
        amps = p.result().reshape(-1)
        p.close()
        assert amps.size == 2 ** 16
        assert abs(np.vdot(amps, amps).real - 1.0) < tol, prec
