"""Batches of independent output bitstrings in one planned run
(compile_plan(bitstrings=...), pq_plan_desc.n_out_bits): every closed output
qubit's leaf holds one column per bitstring, the batch bonds lie above the
slice group, and the FP64 sum keeps one result per bitstring.

  * config 2 (5x5 depth 12): the 8 committed brute-force amplitudes, sliced
    and replayed as a graph, ComplexF64 ........................... 1e-12
    and the bench's precision split ................................ 1e-5
  * Plan.set_bitstrings rewrites the batch in place: the same compiled
    plan and captured graph give the new bitstrings' amplitudes
  * a batch of 5 (padded to 8) with open qubits, against the brute-force
    state vector ................................................... 1e-12
  * fused chains that run once per bitstring (tensors depending on the
    batch bonds only) against the state vector, complex64 ............ 2e-6
"""
import json

import numpy as np
import pytest

from oracle import pq_oracle as O
import picoquant_hip as P
from picoquant_hip import _lib, circuits as C, planner as PL
from picoquant_hip.executor import compile_plan

pytestmark = pytest.mark.gpu


def _net(rows, cols, depth, seed, cfg, dtype=np.complex128):
    n = rows * cols
    tn = P.network_from_gates(n, C.grid_circuit_gates(rows, cols, depth, seed),
                              backend=P.HIPBackend(dtype), decompose=True)
    P.add_input(tn, "0" * n)
    P.add_output(tn, cfg)
    return tn


def _amp(sv, bits):
    return sv[sum(int(b) << q for q, b in enumerate(bits))]


def _rel(got, want):
    got, want = np.asarray(got).reshape(-1), np.asarray(want).reshape(-1)
    return float(np.max(np.abs(got - want) / np.abs(want)))


@pytest.mark.parametrize("precision,tol", [("uniform", 1e-12), ("mixed", 1e-5)])
def test_grid5x5_committed_bitstrings(golden_dir, precision, tol):
    g = json.load(open(golden_dir / "grid5x5_d12.json"))
    bits = g["bitstrings"]
    want = np.array([complex(a, b) for a, b in g["amplitudes"]])
    tn = _net(5, 5, 12, g["seed"], bits[0])
    plan = PL.sweep_plan(tn, PL.grid_site_order(5, 5, "col"))
    sliced = PL.choose_slices(tn, plan, 4)[0]
    p = compile_plan(tn, plan, sliced, group_bits=2, precision=precision, bitstrings=bits)
    assert p.n_out_bits == 3
    for use_graph in (False, True, True):
        p.execute(use_graph)
        got = p.result()
        assert got.shape == (8,)
        assert _rel(got, want) < tol
    # a new batch without recompiling: the captured graph reads the rewritten leaves
    order = [5, 2, 7, 0, 1, 6, 3, 4]
    p.set_bitstrings([bits[i] for i in order])
    for _ in range(2):
        p.execute(True)
        assert _rel(p.result(), want[order]) < tol
    with pytest.raises(ValueError):
        p.set_bitstrings(bits[:3])
    p.close()


def test_plan_run_and_cost_through_the_c_abi(golden_dir):
    """pq_plan_run with raw 64-bit bitstring words (the boundary's one-call
    batched path: set the batch, replay, FP64 result) gives the committed
    config-2 amplitudes, also for a batch shorter than the compiled one
    (padded on the device); pq_plan_cost counts the plan's exact MACs."""
    import ctypes as ct
    g = json.load(open(golden_dir / "grid5x5_d12.json"))
    bits = g["bitstrings"]
    want = np.array([complex(a, b) for a, b in g["amplitudes"]])
    tn = _net(5, 5, 12, g["seed"], bits[0])
    plan = PL.sweep_plan(tn, PL.grid_site_order(5, 5, "col"))
    sliced = PL.choose_slices(tn, plan, 4)[0]
    p = compile_plan(tn, plan, sliced, group_bits=2, precision="uniform", bitstrings=bits)
    ctx, lib = p.ctx, p.ctx.lib
    words = (ct.c_uint64 * 8)(*[sum(int(c) << i for i, c in enumerate(b)) for b in bits])
    n_out = int(np.prod(p.out_shape)) if p.out_shape else 1
    assert n_out == 1
    out = np.empty(2 * 8, dtype=np.float64)
    order = [3, 1, 4, 1, 5, 0, 2, 6]
    for sel in (list(range(8)), order):
        w = (ct.c_uint64 * 8)(*[words[i] for i in sel])
        ctx.check(lib.pq_plan_run(ctx.ptr, p.ptr, 8, w, out.ctypes.data_as(ct.POINTER(ct.c_double))))
        assert _rel(out[0::2] + 1j * out[1::2], want[sel]) < 1e-12
    # three bitstrings: the other five results repeat the first
    w = (ct.c_uint64 * 3)(words[6], words[2], words[5])
    ctx.check(lib.pq_plan_run(ctx.ptr, p.ptr, 3, w, out.ctypes.data_as(ct.POINTER(ct.c_double))))
    assert _rel(out[0::2] + 1j * out[1::2], want[[6, 2, 5, 6, 6, 6, 6, 6]]) < 1e-12
    # a bad count is an error, not a run
    w9 = (ct.c_uint64 * 9)()
    assert lib.pq_plan_run(ctx.ptr, p.ptr, 9, w9, None) != 0
    # cost: the steps' MACs, every step's MACs per slice times the slices it runs for
    c = p.cost()
    assert c["steps"] == len(p.steps) and c["macs"] == pytest.approx(float(np.sum(c["step_macs"])))
    assert c["macs"] == pytest.approx(p.stats()["alg_flops"] / 8, rel=1e-12)
    assert c["launches"] == p.stats()["launches"]
    p.close()


def test_bitstring_batch_with_open_qubits():
    """Five bitstrings (padded to 8 on the device) over a 3x3 grid whose
    qubits 2, 4 and 7 stay open: 5 x 8 amplitudes."""
    cfg = "1x0x11x01"
    sv = O.statevector(9, O.grid_circuit_gates(3, 3, 6, 21))
    rng = np.random.default_rng(4)
    batch = []
    for _ in range(5):
        b = "".join(str(int(v)) for v in rng.integers(0, 2, 9))
        batch.append("".join("x" if c == "x" else b[i] for i, c in enumerate(cfg)))
    tn = _net(3, 3, 6, 21, batch[0])
    open_q = [i + 1 for i, ch in enumerate(cfg) if ch == "x"]
    plan = PL.sweep_plan(tn, PL.grid_site_order(3, 3, open_last=open_q))
    sliced = PL.choose_slices(tn, plan, 2)[0]
    p = compile_plan(tn, plan, sliced, group_bits=1, bitstrings=batch)
    p.execute(True)
    got = p.result()
    assert got.shape == (5, 2, 2, 2)
    for b, bitsb in enumerate(batch):
        for j in range(8):
            full = list(bitsb)
            for t, q in enumerate(open_q):
                full[q - 1] = str((j >> t) & 1)
            want = _amp(sv, full)
            g = got[b].reshape(-1, order="F")[j]
            assert abs(g - want) <= 1e-12 * abs(want)
    p.close()


def test_fused_chains_run_once_per_bitstring():
    cfg0 = "0110" + "x" * 4 + "1001" + "01" + "x" * 2
    sv = O.statevector(16, O.grid_circuit_gates(4, 4, 8, 7))
    rng = np.random.default_rng(9)
    batch = ["".join("x" if c == "x" else str(int(rng.integers(0, 2))) for c in cfg0) for _ in range(4)]
    tn = _net(4, 4, 8, 7, batch[0])
    plan = PL.reconfigure(tn, PL.sweep_plan(tn, PL.grid_site_order(4, 4, "col")), windows=(8,))
    p = compile_plan(tn, plan, precision="mixed", small_elems=16, bitstrings=batch,
                     fuse={"tile_bits": 10, "min_elems": 64, "op_elems": 256})
    chains = [i for i, st in enumerate(p.steps) if st[0] == 2]
    assert chains and any(p.step_info(i)["dep"] for i in chains)
    ctx = tn.backend.ctx
    ctx.profile_enable(True)
    p.execute(False)
    launches = ctx.profile_read(_lib.KC_CHAIN)["launches"]
    ctx.profile_enable(False)
    # a chain that depends on the batch bonds launches once per bitstring
    assert launches == sum(4 if p.step_info(i)["dep"] else 1 for i in chains)
    p.execute(True)
    p.execute(True)
    got = p.result()
    open_q = [i for i, ch in enumerate(cfg0) if ch == "x"]
    for b, bits in enumerate(batch):
        want = []
        for j in range(2 ** len(open_q)):
            full = list(bits)
            for t, q in enumerate(open_q):
                full[q] = str((j >> t) & 1)
            want.append(_amp(sv, full))
        want = np.array(want)
        g = got[b].reshape(-1, order="F")
        assert np.linalg.norm(g - want) / np.linalg.norm(want) < 2e-6
    p.close()


def test_bad_bitstring_batches_are_rejected():
    tn = _net(2, 2, 3, 1, "0110")
    plan = PL.sweep_plan(tn, PL.grid_site_order(2, 2))
    with pytest.raises(ValueError):
        compile_plan(tn, plan, bitstrings=["01x0"])          # an open character at a closed qubit
    with pytest.raises(ValueError):
        compile_plan(tn, plan, bitstrings=["011"])           # too short
    with pytest.raises(ValueError):
        compile_plan(tn, plan, bitstrings=[])
