"""Network ingest: the qiskit-free QASM 2.0 parser and
the JSON structure format of src/layer3.jl:459-707.  The QASM fixtures
(qiskit-transpiled QFTs from the reference's examples/) must reproduce the
discrete Fourier transform exactly, which pins gate matrices, parameter
parsing and qubit order independently of this code."""
import json
import math

import numpy as np
import pytest

from conftest import GOLDEN
from oracle import pq_oracle as O
import picoquant_hip as P
from picoquant_hip import circuits as C, network as NW, planner as PL, qasm as Q
from test_host import RecordingBackend


def test_parameter_expressions():
    assert Q.eval_param("pi/2") == math.pi / 2
    assert Q.eval_param("-pi/4") == -math.pi / 4
    assert Q.eval_param("2*pi/3") == 2 * math.pi / 3
    assert Q.eval_param("1.5e-3") == 1.5e-3
    assert Q.eval_param("-(pi+1)/2") == -(math.pi + 1) / 2
    assert Q.eval_param("sin(pi/2)") == 1.0
    assert Q.eval_param("2^3^2") == 2 ** 9
    assert Q.eval_param("3.141592653589793") == math.pi
    for bad in ("pi/", "foo", "(1", "1 2", "__import__"):
        with pytest.raises(Q.QasmError):
            Q.eval_param(bad)


@pytest.mark.parametrize("n", [2, 3, 5, 10])
def test_qasm_qft_fixtures_are_the_dft(n):
    circ = Q.load_qasm_as_circuit_from_file(GOLDEN / f"qft_{n}.qasm")
    assert circ.n_qubits == n and set(circ.names) <= {"u1", "u2", "cx"}
    N = 2 ** n
    y = np.arange(N)
    for x in (0, 1, N - 1, N // 3):
        sv = np.zeros(N, dtype=np.complex128)
        sv[x] = 1
        for U, qs in circ.gates:
            sv = O.apply_gate_sv(sv, n, U, qs)
        np.testing.assert_allclose(sv, np.exp(2j * np.pi * x * y / N) / math.sqrt(N), atol=1e-12)


def test_qasm_ghz_and_network_conversion():
    circ = Q.load_qasm_as_circuit((GOLDEN / "ghz_3.qasm").read_text())
    tn = Q.convert_qiskit_circ_to_network(circ, RecordingBackend())
    assert len(tn.nodes) == 3 and tn.number_qubits == 3
    to = O.network_from_gates(3, circ.gates)
    O.add_input(to, "000")
    O.full_wavefunction_contraction(to, "vector")
    psi = to.backend.load_tensor_data("result")
    want = np.zeros(8, dtype=np.complex128)
    want[0] = want[7] = 1 / math.sqrt(2)
    np.testing.assert_allclose(psi, want, atol=1e-15)
    assert Q.load_qasm_as_circuit_from_file(GOLDEN / "missing.qasm") is False
    with pytest.raises(NotImplementedError):
        Q.convert_qiskit_circ_to_network(circ, RecordingBackend(), transpile=True)


def test_gate_matrices():
    rng = np.random.default_rng(3)
    for name, (npar, nq, build) in Q.GATES.items():
        U = np.asarray(build(*rng.uniform(-3, 3, npar)))
        assert U.shape == (2 ** nq, 2 ** nq), name
        np.testing.assert_allclose(U.conj().T @ U, np.eye(2 ** nq), atol=1e-14, err_msg=name)
    g = {k: v[2] for k, v in Q.GATES.items()}
    np.testing.assert_allclose(g["u3"](math.pi / 2, 0.3, -1.1), g["u2"](0.3, -1.1))
    np.testing.assert_allclose(g["rz"](0.7), np.exp(-0.35j) * g["u1"](0.7), atol=1e-15)
    np.testing.assert_allclose(g["cx"](), C.CX)
    np.testing.assert_allclose(g["cz"](), C.CZ)
    np.testing.assert_allclose(g["cu1"](0.4), C.cphase(0.4))
    np.testing.assert_allclose(g["h"](), C.H)
    np.testing.assert_allclose(g["swap"](), C.SWAP)
    # ccx flips the target (third qubit) when both controls are set
    e = np.eye(8)
    assert np.array_equal(g["ccx"]() @ e[:, 3], e[:, 7]) and np.array_equal(g["ccx"]() @ e[:, 5], e[:, 5])


@pytest.mark.parametrize("src", [
    "qreg q[2]; foo q[0];",
    "qreg q[2]; gate g a { h a; } g q[0];",
    "qreg q[2]; cx q[0];",
    "qreg q[2]; h q[2];",
    "qreg q[2]; u1 q[0];",
    "h q[0];",
    "qreg q[2]; cx q[1],q[1];",
    "qreg q[2]; qreg r[2];",
])
def test_qasm_errors(src):
    with pytest.raises(Q.QasmError):
        Q.parse_qasm('OPENQASM 2.0;\ninclude "qelib1.inc";\n' + src)


def test_qasm_skips_registers_and_measurement():
    c = Q.parse_qasm('OPENQASM 2.0; include "qelib1.inc"; qreg q[2]; creg c[2];\n'
                     "h q[0]; // comment\nbarrier q[0],q[1]; cx q[0],q[1]; measure q[0] -> c[0];")
    assert c.names == ["h", "cx"]


def test_json_round_trip():
    """to_json / network_from_json keep the reference's schema (dims as
    strings) and every field; a plan found on the copy has the same ledger."""
    gates = C.grid_circuit_gates(2, 3, 4, 5)
    tn = P.network_from_gates(6, gates, backend=RecordingBackend(), decompose=True)
    P.add_input(tn, "010101")
    P.add_output(tn, "0x1x00")
    s = NW.to_json(tn)
    d = json.loads(s)
    assert d["number_qubits"] == 6 and all(isinstance(x, str) for n in d["nodes"].values() for x in n["dims"])
    assert set(d) == {"number_qubits", "edges", "nodes", "input_qubits", "output_qubits", "qubit_ordering",
                      "node_layers"}
    tn2 = NW.network_from_json(s, RecordingBackend())
    for f in ("number_qubits", "input_qubits", "output_qubits", "qubit_ordering", "node_layers"):
        assert getattr(tn2, f) == getattr(tn, f), f
    assert list(tn2.nodes.items()) == list(tn.nodes.items())
    assert list(tn2.edges.items()) == list(tn.edges.items())
    assert tn2.counters["node"] == tn.counters["node"] and tn2.counters["index"] == tn.counters["index"]
    plan = PL.bgreedy_single(tn2, 1.0, 0.0)[0]
    assert PL.plan_ledger(tn2, plan) == PL.plan_ledger(tn, plan)
    assert NW.to_json(tn2, 2) == json.dumps(json.loads(s), indent=2)
