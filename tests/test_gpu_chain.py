"""Fused-chain steps (PQ_KC_CHAIN) on the GPU: a plan compiled with and
without fusion gives the same amplitudes (complex64 tolerance) and the
oracle's; the chain kernel really ran; the larger operator stages (16 -> 32,
32 x 32, 4 -> 32, 2 -> 4; the complex products of up to 16 x 16
take three real MFMAs, the larger ones four) against FP64 plans and the
committed 5x5 depth-12 amplitudes; bad programs are rejected on the host
before anything is launched."""
import json

import numpy as np
import pytest

from oracle import pq_oracle as O
import picoquant_hip as P
from picoquant_hip import _lib, circuits as C, planner as PL
from picoquant_hip.executor import chain_stages, compile_plan

pytestmark = pytest.mark.gpu

CFG = "0110" + "x" * 4 + "1001" + "01" + "x" * 2


def _net():
    n = 16
    tn = P.network_from_gates(n, C.grid_circuit_gates(4, 4, 8, 7), backend=P.HIPBackend(np.complex128),
                              decompose=True)
    P.add_input(tn, "0" * n)
    P.add_output(tn, CFG)
    return tn


def _oracle():
    to = O.network_from_gates(16, O.grid_circuit_gates(4, 4, 8, 7), decompose=True)
    O.add_input(to, "0" * 16)
    O.add_output(to, CFG)
    return to


@pytest.mark.parametrize("tile_bits", [6, 10, 12])
def test_fused_plan_matches_unfused_and_oracle(tile_bits):
    tn = _net()
    plan = PL.reconfigure(tn, PL.sweep_plan(tn, PL.grid_site_order(4, 4, "col")), windows=(8,))
    limits = {"tile_bits": tile_bits, "min_elems": 64, "op_elems": 256}
    base = compile_plan(tn, plan, precision="mixed", small_elems=16, fuse=False)
    fused = compile_plan(tn, plan, precision="mixed", small_elems=16, fuse=limits)
    kinds = [s[0] for s in fused.steps]
    assert kinds.count(2) >= 1
    ctx = tn.backend.ctx
    ctx.profile_enable(True)
    for use_graph in (False, True, True):
        fused.execute(use_graph)
    got = fused.result().reshape(-1, order="F")
    launches = ctx.profile_read(_lib.KC_CHAIN)["launches"]
    ctx.profile_enable(False)
    assert launches >= kinds.count(2)
    base.execute(True)
    want = base.result().reshape(-1, order="F")
    to = _oracle()
    O.contract_network(to, plan)
    ref = np.asarray(to.backend.load_tensor_data("result")).reshape(-1, order="F")
    scale = np.linalg.norm(ref)
    assert np.linalg.norm(got - want) / scale < 2e-6
    assert np.linalg.norm(got - ref) / scale < 2e-6
    for i, k in enumerate(kinds):
        if k == 2:
            assert fused.step_info(i)["kc"] == _lib.KC_CHAIN
    base.close()
    fused.close()


def _shapes(p):
    out = set()
    for i, st in enumerate(p.steps):
        if st[0] == 2:
            assert p.step_info(i)["kc"] == _lib.KC_CHAIN
            out |= set(chain_stages(p.chain_programs[i]))
    return out


def test_large_operator_stages_against_fp64():
    """5x5 depth 12 (seed 33, last qubit open), reconfigured column sweep,
    2^13-element tiles and operators of up to 32 x 32: chains with 2 -> 4,
    4 -> 32 and 16 -> 32 stages, against the same plan in FP64."""
    tn = P.network_from_gates(25, C.grid_circuit_gates(5, 5, 12, 33), backend=P.HIPBackend(np.complex128),
                              decompose=True)
    P.add_input(tn, "0" * 25)
    P.add_output(tn, "0" * 24 + "x")
    plan = PL.reconfigure(tn, PL.sweep_plan(tn, PL.grid_site_order(5, 5, "col")), windows=(8,))
    limits = {"tile_bits": 13, "min_elems": 64, "op_elems": 1024}
    fused = compile_plan(tn, plan, precision="mixed", small_elems=16, fuse=limits)
    assert {(1, 2), (2, 5), (4, 5)} <= _shapes(fused)
    ref = compile_plan(tn, plan)
    ref.execute(True)
    want = ref.result().reshape(-1, order="F")
    ref.close()
    for use_graph in (False, True, True):
        fused.execute(use_graph)
        got = fused.result().reshape(-1, order="F")
        assert np.linalg.norm(got - want) / np.linalg.norm(want) < 2e-6
    fused.close()


def test_32x32_stage_against_golden(golden_dir):
    """Config 2 (5x5 depth 12, its committed batch of 16 amplitudes) on the
    reconfigured column sweep with the executor's default fusion (2^13-
    element tiles, operators up to 32 x 32, tensors from 2^16 elements):
    chains with a 32 x 32 stage, within 1e-5 of the committed amplitudes."""
    g = json.load(open(golden_dir / "grid5x5_d12.json"))
    cfg = list(g["batch_base"])
    for q in g["open_qubits_1based"]:
        cfg[q - 1] = "x"
    tn = P.network_from_gates(25, C.grid_circuit_gates(5, 5, 12, g["seed"]), backend=P.HIPBackend(np.complex128),
                              decompose=True)
    P.add_input(tn, "0" * 25)
    P.add_output(tn, "".join(cfg))
    plan = PL.reconfigure(tn, PL.sweep_plan(tn, PL.grid_site_order(5, 5, "col", open_last=g["open_qubits_1based"])))
    p = compile_plan(tn, plan, precision="mixed", small_elems=1 << 16)
    assert (5, 5) in _shapes(p)
    want = np.array([complex(a, b) for a, b in g["batch_amplitudes"]])
    for use_graph in (False, True):
        p.execute(use_graph)
        got = p.result().reshape(-1, order="F")
        assert float(np.max(np.abs(got - want) / np.abs(want))) < 1e-5
    p.close()


def test_pair_moves_do_not_change_results():
    """Gather / scatter element pairs as one 16-byte access or as two 8-byte
    accesses (pq_set_chain_pairs 0): bit-identical amplitudes, at tile sizes
    where chains take both forms (a scatter pair of stride 4 at 2^6)."""
    tn = _net()
    ctx = tn.backend.ctx
    plan = PL.reconfigure(tn, PL.sweep_plan(tn, PL.grid_site_order(4, 4, "col")), windows=(8,))
    for tile_bits in (6, 12):
        limits = {"tile_bits": tile_bits, "min_elems": 64, "op_elems": 256}
        fused = compile_plan(tn, plan, precision="mixed", small_elems=16, fuse=limits)
        out = []
        for mode in (-1, 0, -1):
            ctx.set_chain_pairs(mode)
            try:
                fused.execute(False)   # eager: a captured graph keeps its first launch arguments
                out.append(fused.result().copy())
            finally:
                ctx.set_chain_pairs(-1)
        fused.close()
        np.testing.assert_array_equal(out[0], out[1])
        np.testing.assert_array_equal(out[0], out[2])


def test_bad_chain_programs_are_rejected():
    from picoquant_hip import executor as E
    tn = _net()
    plan = PL.reconfigure(tn, PL.sweep_plan(tn, PL.grid_site_order(4, 4, "col")), windows=(8,))
    limits = {"tile_bits": 10, "min_elems": 64, "op_elems": 256}
    real = E.fuse_chains

    def corrupt(how):
        def f(slots, steps, dts, **kw):
            steps, progs, nf = real(slots, steps, dts, **kw)
            # the last program, so a truncated one runs off the end of the words
            i = max(i for i, p in enumerate(progs) if p is not None)
            progs = list(progs)
            progs[i] = how(list(progs[i]))
            return steps, progs, nf
        return f

    cases = [
        lambda w: [w[0], 14] + w[2:],                       # tile beyond the limit
        lambda w: w[:5] + [w[5], 1 << 30] + w[7:],          # gathered stride out of range
        lambda w: w[:-1],                                   # truncated program
    ]
    try:
        for how in cases:
            E.fuse_chains = corrupt(how)
            with pytest.raises(_lib.PQError):
                compile_plan(tn, plan, precision="mixed", small_elems=16, fuse=limits)
    finally:
        E.fuse_chains = real
