"""MPS contraction and compression on the GPU (pq_decompose: rocSOLVER SVD;
pq_contract for the pairs) against the oracle's restatement of
src/layer2.jl:423-703 / src/mps.jl and against the state vector.  The SVD's
gauge (a phase per singular vector) is free, so splits are compared through
gauge-invariant quantities: the bond dimension, the product B C, the
truncation residual and the final amplitudes."""
import math

import numpy as np
import pytest

from oracle import pq_oracle as O
import picoquant_hip as P
from picoquant_hip import _lib, circuits as C

pytestmark = pytest.mark.gpu

TOL = {np.complex128: 1e-10, np.complex64: 2e-5}


@pytest.fixture(scope="module", params=[np.complex128, np.complex64], ids=["c128", "c64"])
def dtype(request):
    return request.param


def ghz_gates(n):
    return [(C.H, [0])] + [(C.CX, [q, q + 1]) for q in range(n - 1)]


def _gpu_mps_vector(dtype, n, gates, max_rank=0):
    tn = P.network_from_gates(n, gates, backend=P.HIPBackend(dtype), decompose=True)
    P.add_input(tn, "0" * n)
    nodes = P.contract_mps_tensor_network_circuit(tn, max_rank=max_rank)
    P.calculate_mps_amplitudes(tn, nodes)
    return np.asarray(tn.backend.load_tensor_data("result")).reshape(-1, order="F")


@pytest.mark.parametrize("rank,thr", [(0, 1e-13), (3, 1e-13), (1, 1e-13), (0, 0.3)])
def test_decompose_matches_oracle(dtype, rank, thr):
    rng = np.random.default_rng(rank + 7)
    t = rng.standard_normal((4, 3, 4, 2)) + 1j * rng.standard_normal((4, 3, 4, 2))
    ctx = _lib.Context(dtype)
    try:
        h = ctx.upload(t)
        hb, hc, chi = ctx.decompose(h, [1, 3], [2, 4], thr, rank)
        B, Cm = ctx.download(hb), ctx.download(hc)
        for x in (h, hb, hc):
            ctx.free(x)
    finally:
        ctx.close()
    Bo, Co, chio = O.decompose_tensor_data(t, [1, 3], [2, 4], thr, rank)
    assert chi == chio and B.shape == Bo.shape and Cm.shape == Co.shape
    A = np.transpose(t, [0, 2, 1, 3]).reshape(16, 6, order="F")
    BC = B.reshape(16, chi, order="F").astype(np.complex128) @ Cm.reshape(chi, 6, order="F")
    BCo = Bo.reshape(16, chi, order="F") @ Co.reshape(chi, 6, order="F")
    scale = np.linalg.norm(A)
    assert np.linalg.norm(BC - BCo) / scale < TOL[dtype]
    s = np.linalg.svd(A, compute_uv=False)
    resid = math.sqrt(float(np.sum(s[chi:] ** 2)))
    assert abs(np.linalg.norm(A - BC) - resid) / scale < 10 * TOL[dtype]


def test_bond_dimension_rule_on_a_known_spectrum(dtype):
    rng = np.random.default_rng(6)
    U, _ = np.linalg.qr(rng.standard_normal((8, 4)) + 1j * rng.standard_normal((8, 4)))
    V, _ = np.linalg.qr(rng.standard_normal((4, 4)) + 1j * rng.standard_normal((4, 4)))
    s = np.array([1.0, 0.5, 1e-3, 1e-9])
    t = (U @ np.diag(s) @ V.conj().T).reshape(2, 4, 4, order="F")
    ctx = _lib.Context(dtype)
    try:
        h = ctx.upload(t)
        for thr, rank, want in [(0.4, 0, 2), (1e-2, 0, 2), (1e-13, 0, 3), (1e-13, 1, 1)]:
            hb, hc, chi = ctx.decompose(h, [1, 2], [3], thr, rank)
            ctx.free(hb)
            ctx.free(hc)
            assert chi == want, (thr, rank, chi)
        ctx.free(h)
    finally:
        ctx.close()


def test_ghz5_mps_and_mpsstate(dtype):
    n = 5
    psi = _gpu_mps_vector(dtype, n, ghz_gates(n))
    ref = np.zeros(2 ** n, dtype=np.complex128)
    ref[[0, -1]] = 1 / math.sqrt(2)
    assert np.abs(psi - ref).max() < TOL[dtype]
    tn = P.network_from_gates(n, ghz_gates(n), backend=P.HIPBackend(dtype), decompose=True)
    P.add_input(tn, "0" * n)
    nodes = P.contract_mps_tensor_network_circuit(tn)
    st = P.MPSState(tn, nodes)
    assert st.shape == (2,) * n and len(st) == 2 ** n
    assert abs(st["11111"] - 1 / math.sqrt(2)) < 1e-5 and abs(st["00000"] - 1 / math.sqrt(2)) < 1e-5
    assert abs(st["10101"]) < 1e-5


@pytest.mark.parametrize("seed", [3, 4])
def test_exact_mps_equals_state_vector(dtype, seed):
    n = 6
    gates = C.grid_circuit_gates(1, n, 6, seed)
    psi = _gpu_mps_vector(dtype, n, gates)
    sv = O.statevector(n, O.grid_circuit_gates(1, n, 6, seed))
    assert np.linalg.norm(psi - sv) / np.linalg.norm(sv) < (1e-8 if dtype == np.complex128 else 1e-5)


def test_truncated_mps_matches_oracle():
    """max_rank = 2 on a deeper chain circuit: the same truncation as the
    oracle's (gauge-invariant amplitudes), different from the exact state."""
    n = 6
    gates = C.grid_circuit_gates(1, n, 8, 9)
    psi = _gpu_mps_vector(np.complex128, n, gates, max_rank=2)
    tn = O.network_from_gates(n, O.grid_circuit_gates(1, n, 8, 9), decompose=True)
    O.add_input(tn, "0" * n)
    nodes = O.contract_mps_tensor_network_circuit(tn, max_rank=2)
    O.calculate_mps_amplitudes(tn, nodes)
    ref = np.asarray(tn.backend.load_tensor_data("result")).reshape(-1, order="F")
    assert np.linalg.norm(psi - ref) / np.linalg.norm(ref) < 1e-8
    assert np.linalg.norm(ref - O.statevector(n, O.grid_circuit_gates(1, n, 8, 9))) > 1e-6


def test_chain_compression_reference_example(dtype):
    gates = [(C.H, [0]), (C.CX, [0, 1]), (C.CX, [0, 1]), (C.CX, [0, 1])]
    tn = P.network_from_gates(2, gates, backend=P.HIPBackend(dtype))
    P.add_input(tn, "00")
    P.decompose_tensor(tn, "node_2", ["index_3", "index_4"], ["index_2", "index_5"])
    P.decompose_tensor(tn, "node_3", ["index_4", "index_6"], ["index_5", "index_7"])
    P.decompose_tensor(tn, "node_4", ["index_6", "index_8"], ["index_7", "index_9"])
    plan = [k for k, v in tn.edges.items() if not v.virtual and v.src is not None and v.dst is not None]
    for e in plan:
        P.contract_pair(tn, e)
    a, b = list(tn.nodes)
    P.compress_tensor_chain(tn, [a, b])
    shared = [x for x in tn.nodes[a].indices if x in tn.nodes[b].indices]
    assert len(shared) == 1
    assert tn.backend.load_tensor_data(a).shape[tn.nodes[a].indices.index(shared[0])] == 2
    out = P.contract_pair(tn, a, b)
    node = tn.nodes[out]
    order = [node.indices.index(tn.output_qubits[q - 1]) + 1 for q in tn.qubit_ordering]
    tn.backend.permute_tensor(out, order)
    psi = np.asarray(tn.backend.load_tensor_data(out)).reshape(-1, order="F")
    sv = O.statevector(2, [(O.H, [0]), (O.CX, [0, 1]), (O.CX, [0, 1]), (O.CX, [0, 1])])
    assert np.abs(psi - sv).max() < TOL[dtype]


def test_mps_recorded_and_replayed(tmp_path, dtype):
    """DSLBackend records the MPS contraction of GHZ-4 (decompose commands
    included); execute_dsl_file replays it with the GPU SVD."""
    from picoquant_hip import dsl as D
    n = 4
    tl, store = tmp_path / "mps.tl", tmp_path / "mps.npz"
    be = D.DSLBackend(tl, store, "", True)
    tn = P.network_from_gates(n, ghz_gates(n), backend=be, decompose=True)
    P.add_input(tn, "0" * n)
    nodes = P.contract_mps_tensor_network_circuit(tn)
    P.calculate_mps_amplitudes(tn, nodes)
    D.execute_dsl_file(dtype, tl, store)
    psi = np.asarray(be.load_tensor_data("result")).reshape(-1, order="F")
    ref = np.zeros(2 ** n, dtype=np.complex128)
    ref[[0, -1]] = 1 / math.sqrt(2)
    assert np.abs(psi - ref).max() < TOL[dtype]
