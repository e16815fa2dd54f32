"""MPS contraction and bond compression in the oracle (src/layer2.jl:423-703,
src/mps.jl), pinned by mathematics and by the reference's own tests:

* an MPS contraction without truncation is exact: it equals the brute-force
  state vector of the same circuit;
* GHZ-5 by MPS gives (|0..0> + |1..1>)/sqrt(2), and MPSState's selections
  of it (test/layer2_tests.jl:300-347);
* a split truncated to chi singular values leaves exactly the discarded
  singular values as residual (Eckart-Young), and the bond dimension follows
  the threshold / max_rank rule on a tensor with a known spectrum
  (test/layer2_tests.jl:349-416);
* the reference's chain-compression example (H, CX, CX, CX; each CX split)
  ends with one bond of dimension 2 (test/layer2_tests.jl:241-295).
"""
import math

import numpy as np
import pytest

from oracle import pq_oracle as O


def ghz_gates(n):
    return [(O.H, [0])] + [(O.CX, [q, q + 1]) for q in range(n - 1)]


def _mps_vector(n, gates, max_rank=0, threshold=1e-13):
    tn = O.network_from_gates(n, gates, decompose=True)
    O.add_input(tn, "0" * n)
    nodes = O.contract_mps_tensor_network_circuit(tn, threshold=threshold, max_rank=max_rank)
    O.calculate_mps_amplitudes(tn, nodes)
    return np.asarray(tn.backend.load_tensor_data("result")).reshape(-1, order="F")


def test_ghz5_mps_closed_form_and_mpsstate():
    n = 5
    psi = _mps_vector(n, ghz_gates(n))
    ref = np.zeros(2 ** n, dtype=np.complex128)
    ref[[0, -1]] = 1 / math.sqrt(2)
    np.testing.assert_allclose(psi, ref, atol=1e-12)
    tn = O.network_from_gates(n, ghz_gates(n), decompose=True)
    O.add_input(tn, "0" * n)
    nodes = O.contract_mps_tensor_network_circuit(tn)
    a1, a0, a5 = (O.mps_amplitude(tn, nodes, c) for c in ("11111", "00000", "10101"))
    assert abs(a1 - 1 / math.sqrt(2)) < 1e-12 and abs(a0 - 1 / math.sqrt(2)) < 1e-12
    assert abs(a5) < 1e-12


@pytest.mark.parametrize("seed", [3, 4])
def test_exact_mps_equals_state_vector(seed):
    """A 1 x 6 grid circuit (neighbour couplers only), depth 6, no truncation."""
    n = 6
    gates = O.grid_circuit_gates(1, n, 6, seed)
    psi = _mps_vector(n, gates)
    sv = O.statevector(n, gates)
    np.testing.assert_allclose(psi, sv, atol=1e-9)
    # the compression driver (no neighbour check) on the same circuit
    tn = O.network_from_gates(n, gates, decompose=True)
    O.add_input(tn, "0" * n)
    nodes = O.contract_tensor_network_circuit_with_compression(tn)
    O.calculate_mps_amplitudes(tn, nodes)
    got = np.asarray(tn.backend.load_tensor_data("result")).reshape(-1, order="F")
    np.testing.assert_allclose(got, sv, atol=1e-9)


def test_truncated_mps_differs_and_keeps_bond_cap():
    n = 6
    gates = O.grid_circuit_gates(1, n, 8, 9)
    sv = O.statevector(n, gates)
    psi = _mps_vector(n, gates, max_rank=2)
    assert np.linalg.norm(psi - sv) > 1e-6          # truncation changed the state
    tn = O.network_from_gates(n, gates, decompose=True)
    O.add_input(tn, "0" * n)
    nodes = O.contract_mps_tensor_network_circuit(tn, max_rank=2)
    for a, b in zip(nodes, nodes[1:]):
        bond = [x for x in tn.nodes[a].indices if x in tn.nodes[b].indices]
        assert len(bond) == 1
        assert tn.nodes[a].dims[tn.nodes[a].indices.index(bond[0])] <= 2


def test_split_residual_is_the_discarded_spectrum():
    rng = np.random.default_rng(5)
    t = rng.standard_normal((4, 3, 4, 2)) + 1j * rng.standard_normal((4, 3, 4, 2))
    left, right = [1, 3], [2, 4]
    A = np.transpose(t, [0, 2, 1, 3]).reshape(16, 6, order="F")
    s = np.linalg.svd(A, compute_uv=False)
    for rank in (1, 2, 4, 6):
        B, C, chi = O.decompose_tensor_data(t, left, right, max_rank=rank)
        assert chi == rank and B.shape == (4, 4, chi) and C.shape == (chi, 3, 2)
        BC = B.reshape(16, chi, order="F") @ C.reshape(chi, 6, order="F")
        assert abs(np.linalg.norm(A - BC) - math.sqrt(float(np.sum(s[chi:] ** 2)))) < 1e-10


def test_bond_dimension_rule_on_a_known_spectrum():
    """A = U diag(s) V^H with s = (1, 0.5, 1e-3, 1e-9): the ratios to ||s||
    decide chi against max(threshold, sqrt(eps)); max_rank caps it."""
    rng = np.random.default_rng(6)
    U, _ = np.linalg.qr(rng.standard_normal((8, 4)) + 1j * rng.standard_normal((8, 4)))
    V, _ = np.linalg.qr(rng.standard_normal((4, 4)) + 1j * rng.standard_normal((4, 4)))
    s = np.array([1.0, 0.5, 1e-3, 1e-9])
    t = (U @ np.diag(s) @ V.conj().T).reshape(2, 4, 4, order="F")
    for thr, rank, want in [(0.4, 0, 2), (1e-2, 0, 2), (1e-4, 0, 3), (1e-13, 0, 3), (1e-13, 1, 1)]:
        assert O.decompose_tensor_data(t, [1, 2], [3], threshold=thr, max_rank=rank)[2] == want


def test_chain_compression_reference_example():
    """H, CX, CX, CX on 2 qubits, each CX split (three bonds of dimension 2);
    contracting the non-virtual edges leaves two nodes joined by those bonds;
    compressing the chain leaves one bond of dimension 2 and the state
    (a Bell pair: CX^3 = CX) unchanged."""
    gates = [(O.H, [0]), (O.CX, [0, 1]), (O.CX, [0, 1]), (O.CX, [0, 1])]
    tn = O.network_from_gates(2, gates)
    O.add_input(tn, "00")
    O.decompose_tensor_network(tn, "node_2", ["index_3", "index_4"], ["index_2", "index_5"])
    O.decompose_tensor_network(tn, "node_3", ["index_4", "index_6"], ["index_5", "index_7"])
    O.decompose_tensor_network(tn, "node_4", ["index_6", "index_8"], ["index_7", "index_9"])
    plan = [k for k, v in tn.edges.items() if not v.virtual and v.src is not None and v.dst is not None]
    for e in plan:
        O.contract_pair(tn, e)
    assert len(tn.nodes) == 2
    a, b = list(tn.nodes)
    shared = [x for x in tn.nodes[a].indices if x in tn.nodes[b].indices]
    assert len(shared) == 3
    O.compress_tensor_chain(tn, [a, b])
    shared = [x for x in tn.nodes[a].indices if x in tn.nodes[b].indices]
    assert len(shared) == 1 and tn.nodes[a].dims[tn.nodes[a].indices.index(shared[0])] == 2
    out = O.contract_pair(tn, a, b)
    node = tn.nodes[out]
    order = [node.indices.index(tn.output_qubits[q - 1]) + 1 for q in tn.qubit_ordering]
    tn.backend.permute_tensor(out, order)
    psi = np.asarray(tn.backend.load_tensor_data(out)).reshape(-1, order="F")
    np.testing.assert_allclose(psi, O.statevector(2, gates), atol=1e-12)
