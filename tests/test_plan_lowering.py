"""The plan lowering (picoquant_hip.executor.lower_plan) checked on the CPU:
a small numpy interpreter executes the lowered slots/steps exactly as the
device executor addresses memory (per-slice strides, slice offsets into the
leaves, chosen layouts) and must reproduce the oracle's amplitudes."""
import itertools
import math

import numpy as np
import pytest

from oracle import pq_oracle as O
import picoquant_hip as P
from picoquant_hip import circuits as C
from picoquant_hip import planner as PL
from picoquant_hip.executor import lower_plan
from test_host import RecordingBackend


def offsets(modes, which):
    """All offsets of a mode group in its linear order (first mode fastest)."""
    offs = np.zeros(1, dtype=np.int64)
    for m in modes:
        ext, st = m[0], m[1 + which]
        offs = (offs[None, :] + np.arange(ext, dtype=np.int64)[:, None] * st).T.reshape(-1)
    return offs


def interpret(network, leaf_data, plan, sliced, carry_elems=None):
    kw = {} if carry_elems is None else {"carry_elems": carry_elems}
    slots, steps, out_slot = lower_plan(network, plan, sliced, **kw)
    nslices = 2 ** len(sliced)
    total = None
    for s in range(nslices):
        bits = [(s >> j) & 1 for j in range(len(sliced))]
        mem = {i: leaf_data[i] for i in range(len(network.nodes))}
        for sa, sb, sc, M, N, K in steps:
            # a step batched over slices reads slots that keep sliced bonds
            # (leaves, carried intermediates) at this slice's offset; a
            # carrying step has those bonds among its modes instead
            def off(x):
                if not slots[sc].dep:
                    return 0
                return sum(b * st for b, st in zip(bits, slots[x].slice_stride))
            Am = mem[sa][off(sa) + offsets(M, 0)[:, None] + offsets(K, 0)[None, :]]
            Bm = mem[sb][off(sb) + offsets(K, 1)[:, None] + offsets(N, 0)[None, :]]
            out = np.zeros(slots[sc].elems, dtype=np.complex128)
            out[offsets(M, 1)[:, None] + offsets(N, 1)[None, :]] = Am @ Bm
            mem[sc] = out
        res = mem[out_slot]
        total = res if total is None else total + res
    return total, slots[out_slot]


def build(rows, cols, depth, seed, cfg):
    gates = C.grid_circuit_gates(rows, cols, depth, seed)
    n = rows * cols
    tn = P.network_from_gates(n, gates, backend=RecordingBackend(), decompose=True)
    P.add_input(tn, "0" * n)
    P.add_output(tn, cfg)
    to = O.network_from_gates(n, O.grid_circuit_gates(rows, cols, depth, seed), decompose=True)
    O.add_input(to, "0" * n)
    O.add_output(to, cfg)
    data = [np.asarray(to.backend.tensors[k]).reshape(-1, order="F") for k in tn.nodes]
    return tn, data, gates


@pytest.mark.parametrize("carry", [None, 0, 8], ids=["carry", "no-carry", "carry-tiny"])
@pytest.mark.parametrize("nslice", [0, 3, 5])
def test_lowered_plan_reproduces_amplitudes(nslice, carry):
    rows, cols, depth, seed = 3, 3, 5, 4
    cfg = "10x1x0110"
    tn, data, _ = build(rows, cols, depth, seed, cfg)
    sv = O.statevector(9, O.grid_circuit_gates(rows, cols, depth, seed))
    open_q = [i + 1 for i, ch in enumerate(cfg) if ch == "x"]
    plan = PL.sweep_plan(tn, PL.grid_site_order(rows, cols, open_last=open_q))
    sliced = PL.choose_slices(tn, plan, nslice)[0] if nslice else []
    res, out = interpret(tn, data, plan, sliced, carry)
    assert len(out.idx) == 2
    for j in range(4):
        bits = [int(ch) if ch != "x" else 0 for ch in cfg]
        bits[open_q[0] - 1] = j & 1
        bits[open_q[1] - 1] = (j >> 1) & 1
        want = sv[sum(b << q for q, b in enumerate(bits))]
        assert abs(res[j] - want) < 1e-13


def test_lowered_greedy_plan_and_layouts():
    tn, data, gates = build(2, 3, 4, 8, "011010")
    sv = O.statevector(6, O.grid_circuit_gates(2, 3, 4, 8))
    plan, _, _ = PL.bgreedy_single(tn, 1.0, 0.0)
    sliced = PL.choose_slices(tn, plan, 2)[0]
    res, out = interpret(tn, data, plan, sliced)
    bits = [0, 1, 1, 0, 1, 0]
    assert abs(res[0] - sv[sum(b << q for q, b in enumerate(bits))]) < 1e-13
    # intermediates put the indices their consumer sums first
    slots, steps, _ = lower_plan(tn, plan, sliced)
    consumer = {}
    for sa, sb, sc, *_ in steps:
        consumer[sa], consumer[sb] = sb, sa
    for sa, sb, sc, *_ in steps:
        if sc in consumer:
            partner = set(slots[consumer[sc]].idx)
            k = [x in partner for x in slots[sc].idx]
            assert k == sorted(k, reverse=True)


def test_slicer_cost_matches_recount():
    tn, _, _ = build(3, 3, 6, 1, "0" * 9)
    plan = PL.sweep_plan(tn, PL.grid_site_order(3, 3))
    bonds, tot, peak = PL.choose_slices(tn, plan, 4)
    assert len(set(bonds)) == 4
    # recount from the lowered plan: a step runs once per value of the
    # sliced bonds its result depends on (without carrying small steps, as
    # the slicer's model)
    slots, steps, _ = lower_plan(tn, plan, bonds, carry_elems=0)
    recount = 0
    for sa, sb, sc, M, N, K in steps:
        m = math.prod(x[0] for x in M) * math.prod(x[0] for x in N) * math.prod(x[0] for x in K)
        recount += m * 2 ** len(slots[sc].deps)
    assert recount == tot
    unsliced = PL.plan_ledger(tn, plan)["macs"]
    assert tot >= unsliced


def test_carried_bonds():
    """Small world-line intermediates keep their sliced bonds (run once);
    only large steps are batched over the slices; results agree with and
    without carrying on a greedy plan too."""
    tn, data, _ = build(3, 3, 6, 2, "0x" * 4 + "1")
    plan = PL.sweep_plan(tn, PL.grid_site_order(3, 3))
    sliced = PL.choose_slices(tn, plan, 4)[0]
    slots, steps, out_slot = lower_plan(tn, plan, sliced)
    carried = [s for s in slots if s.carry]
    assert carried
    for s in carried:
        assert not s.dep and sum(1 for v in s.slice_stride if v) == len(s.carry)
    assert slots[out_slot].dep and not slots[out_slot].carry
    n_dep = sum(slots[sc].dep for _, _, sc, *_ in steps)
    n_dep0 = sum(x[0] for x in [(lower_plan(tn, plan, sliced, carry_elems=0)[0][sc].dep,)
                                for _, _, sc, *_ in steps])
    assert n_dep < n_dep0
    a, _ = interpret(tn, data, plan, sliced)
    b, _ = interpret(tn, data, plan, sliced, 0)
    np.testing.assert_allclose(a, b, rtol=1e-12, atol=1e-15)
    g, _, _ = PL.bgreedy_single(tn, 1.0, 0.0)
    sl = PL.choose_slices(tn, g, 3)[0]
    a, _ = interpret(tn, data, g, sl)
    b, _ = interpret(tn, data, g, sl, 0)
    np.testing.assert_allclose(a, b, rtol=1e-12, atol=1e-15)


def test_slicing_candidates():
    tn, _, _ = build(3, 3, 6, 1, "0" * 9)
    plan = PL.sweep_plan(tn, PL.grid_site_order(3, 3))
    greedy = PL.choose_slices(tn, plan, 3)[0]
    cands = PL.slicing_candidates(tn, plan, 3, max_elems=2 ** 10, n_random=10, seed=1)
    assert sorted(cands[0][0]) == sorted(greedy)
    keys = [frozenset(b) for b, _, _ in cands]
    assert len(set(keys)) == len(keys)
    steps, ids, ext = PL._plan_masks(tn, plan)
    for bonds, tot, peak in cands:
        assert len(set(bonds)) == 3 and peak <= 2 ** 10
        smask = sum(1 << ids[b] for b in bonds)
        assert (tot, peak) == PL.sliced_cost(steps, ext, smask, 8, "macs")


@pytest.mark.parametrize("nslice,carry", [(0, None), (3, None), (3, 0)])
def test_lowered_bitstring_batch(nslice, carry):
    """A batch of output bitstrings lowered as extra bonds kept by the closed
    output leaves (executor.output_batch / lower_plan extra_bonds): summing
    each bitstring's slices reproduces its brute-force amplitudes."""
    from picoquant_hip.executor import output_batch
    rows, cols, depth, seed = 3, 3, 5, 4
    cfg = "10x1x0110"
    sv = O.statevector(9, O.grid_circuit_gates(rows, cols, depth, seed))
    rng = np.random.default_rng(11)
    batch = ["".join("x" if c == "x" else str(int(rng.integers(0, 2))) for c in cfg) for _ in range(3)]
    tn, data, _ = build(rows, cols, depth, seed, batch[0])
    cols_data, n_bits = output_batch(tn, batch)
    assert n_bits == 2 and all(v.shape == (2, 4) for v in cols_data.values())
    labels = list(tn.nodes)
    extra = {lab: [2 << k for k in range(n_bits)] for lab in cols_data}
    leaf_data = [cols_data[lab].reshape(-1, order="F") if lab in cols_data else data[i]
                 for i, lab in enumerate(labels)]
    strides = {lab: ([1] if lab in cols_data else [int(np.prod(tn.nodes[lab].dims[:i]))
                                                    for i in range(len(tn.nodes[lab].dims))]) for lab in labels}
    open_q = [i + 1 for i, ch in enumerate(cfg) if ch == "x"]
    plan = PL.sweep_plan(tn, PL.grid_site_order(rows, cols, open_last=open_q))
    sliced = PL.choose_slices(tn, plan, nslice)[0] if nslice else []
    kw = {} if carry is None else {"carry_elems": carry}
    slots, steps, out_slot = lower_plan(tn, plan, sliced, leaf_strides=strides, extra_bonds=extra, **kw)
    nb = len(sliced) + n_bits
    assert sorted(slots[out_slot].deps) == list(range(nb))
    for b, bits_b in enumerate(batch):
        total = 0
        for s in range(2 ** len(sliced)):
            v = [(s >> j) & 1 for j in range(len(sliced))] + [(b >> k) & 1 for k in range(n_bits)]
            mem = {i: leaf_data[i] for i in range(len(labels))}
            for sa, sb, sc, M, N, K in steps:
                def off(x):
                    if not slots[sc].dep:
                        return 0
                    return sum(bit * st for bit, st in zip(v, slots[x].slice_stride))
                Am = mem[sa][off(sa) + offsets(M, 0)[:, None] + offsets(K, 0)[None, :]]
                Bm = mem[sb][off(sb) + offsets(K, 1)[:, None] + offsets(N, 0)[None, :]]
                out = np.zeros(slots[sc].elems, dtype=np.complex128)
                out[offsets(M, 1)[:, None] + offsets(N, 1)[None, :]] = Am @ Bm
                mem[sc] = out
            total = total + mem[out_slot]
        for j in range(4):
            full = list(bits_b)
            full[open_q[0] - 1] = str(j & 1)
            full[open_q[1] - 1] = str((j >> 1) & 1)
            want = sv[sum(int(c) << q for q, c in enumerate(full))]
            assert abs(total[j] - want) < 1e-13


def test_bitstring_words_for_the_c_abi():
    """The uint64 words pq_plan_set_bitstrings reads: bit i = the value of
    closed output qubit i (add_output's order), open qubits ignored; the
    closed qubits are exactly those add_output closed."""
    from picoquant_hip.executor import bitstring_words, closed_output_qubits
    cfg = "1x0x11x01"
    tn, _, _ = build(3, 3, 4, 5, cfg)
    closed = closed_output_qubits(tn)
    assert [i for _, i in closed] == [i for i, c in enumerate(cfg) if c != "x"]
    assert len({node for node, _ in closed}) == len(closed)
    batch = ["1x0x11x01", "0x1x00x10", "1x1x10x11"]
    words = bitstring_words(tn, batch)
    for b, bits in enumerate(batch):
        want = sum(int(c) << i for i, c in enumerate(bits) if c != "x")
        assert words[b] == want
    with pytest.raises(ValueError):
        bitstring_words(tn, ["1x0x11x0x"])
