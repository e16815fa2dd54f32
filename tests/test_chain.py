"""Fused chains (executor.fuse_chains / chain_program) on the CPU.

A numpy interpreter runs the lowered steps exactly as the device executor
addresses memory -- contractions through their (extent, stride, stride)
modes, fused-chain steps through their program words (gather into a tile
by bit positions, the stages in place, scatter) -- and the fused step list
must reproduce the unfused one and the oracle's amplitudes."""
import itertools

import numpy as np
import pytest

from oracle import pq_oracle as O
import picoquant_hip as P
from picoquant_hip import _lib, circuits as C, planner as PL
from picoquant_hip.executor import assign_precision, chain_program, fuse_chains, lower_plan
from test_host import RecordingBackend
from test_plan_lowering import offsets


def run_chain(mem, words, x, y_elems):
    L, tb, nin, nout, nspec = words[:5]
    q = 5
    inb = [words[q + 2 * i: q + 2 * i + 2] for i in range(nin)]
    q += 2 * nin
    outb = [words[q + 2 * i: q + 2 * i + 2] for i in range(nout)]
    q += 2 * nout
    spec = [words[q + 3 * i: q + 3 * i + 3] for i in range(nspec)]
    q += 3 * nspec
    stages = []
    for _ in range(L):
        op, nk, nm, nr = words[q: q + 4]
        q += 4
        kb = [words[q + 2 * i: q + 2 * i + 2] for i in range(nk)]
        q += 2 * nk
        mb = [words[q + 2 * i: q + 2 * i + 2] for i in range(nm)]
        q += 2 * nm
        rb = words[q: q + nr]
        q += nr
        stages.append((op, kb, mb, rb))
    assert q == len(words)

    def addr(bits, field):
        n = len(bits)
        v = np.arange(1 << n)
        out = np.zeros(1 << n, dtype=np.int64)
        for b, t in enumerate(bits):
            val = (1 << t) if field == "pos" else t
            out += ((v >> b) & 1) * val
        return out

    cfg = np.zeros(1, dtype=np.int64), np.zeros(1, dtype=np.int64)
    ox, oy = cfg
    for ext, si, so in spec:
        ox = (ox[:, None] + np.arange(ext)[None, :] * si).reshape(-1)
        oy = (oy[:, None] + np.arange(ext)[None, :] * so).reshape(-1)
    T = np.zeros((len(ox), 1 << tb), dtype=np.complex128)
    T[:, addr([p for p, _ in inb], "pos")] = mem[x][ox[:, None] + addr([s for _, s in inb], "str")[None, :]]
    for op, kb, mb, rb in stages:
        S = mem[op]
        ka, ks = addr([p for p, _ in kb], "pos"), addr([s for _, s in kb], "str")
        ma, ms = addr([p for p, _ in mb], "pos"), addr([s for _, s in mb], "str")
        ra = addr(rb, "pos")
        Smat = S[ms[:, None] + ks[None, :]]                      # [m, k]
        Tin = T[:, ra[:, None] + ka[None, :]]                    # [cfg, r, k]
        T[:, ra[:, None] + ma[None, :]] = Tin @ Smat.T           # [cfg, r, m]
    Y = np.full(y_elems, np.nan, dtype=np.complex128)
    Y[oy[:, None] + addr([s for _, s in outb], "str")[None, :]] = T[:, addr([p for p, _ in outb], "pos")]
    return Y


def interpret(slots, steps, progs, leaf_data, out_slot):
    mem = {i: d for i, d in enumerate(leaf_data)}
    for (kind, sa, sb, sc, M, N, K), prog in zip(steps, progs):
        if kind == 1:
            mem[sc] = mem[sa].copy()
        elif kind == 2:
            mem[sc] = run_chain(mem, prog, sa, slots[sc].elems)
        else:
            Am = mem[sa][offsets(M, 0)[:, None] + offsets(K, 0)[None, :]]
            Bm = mem[sb][offsets(K, 1)[:, None] + offsets(N, 0)[None, :]]
            out = np.zeros(slots[sc].elems, dtype=np.complex128)
            out[offsets(M, 1)[:, None] + offsets(N, 1)[None, :]] = Am @ Bm
            mem[sc] = out
    return mem[out_slot]


def _grid(rows, cols, depth, seed, cfg):
    n = rows * cols
    tn = P.network_from_gates(n, C.grid_circuit_gates(rows, cols, depth, seed), backend=RecordingBackend(),
                              decompose=True)
    P.add_input(tn, "0" * n)
    P.add_output(tn, cfg)
    to = O.network_from_gates(n, O.grid_circuit_gates(rows, cols, depth, seed), decompose=True)
    O.add_input(to, "0" * n)
    O.add_output(to, cfg)
    data = [np.asarray(to.backend.tensors[k]).reshape(-1, order="F") for k in tn.nodes]
    return tn, to, data


@pytest.mark.parametrize("tile_bits", [6, 8, 10])
def test_fused_chains_reproduce_the_unfused_steps(tile_bits):
    cfg = "0110" + "x" * 4 + "1001" + "01" + "x" * 2
    tn, to, data = _grid(4, 4, 8, 7, cfg)
    plan = PL.reconfigure(tn, PL.sweep_plan(tn, PL.grid_site_order(4, 4, "col")), windows=(8,))
    slots, steps, out = lower_plan(tn, plan)
    # complex64 above 16 elements, so most steps qualify
    dts, steps, slots = assign_precision(slots, steps, _lib.PQ_C128, "mixed", small_elems=16)
    fused, progs, members = fuse_chains(slots, steps, dts, tile_bits=tile_bits, min_elems=64, op_elems=256)
    n_chain = sum(1 for p in progs if p is not None)
    assert n_chain >= 1 and len(fused) < len(steps)
    contract = [i for i, s in enumerate(steps) if s[0] == 0]
    assert sorted(j for m in members for j in m) == contract
    for p in progs:
        if p is not None:
            assert p[1] <= tile_bits and 2 <= p[0] <= _lib.PQ_CHAIN_MAX_STAGES
    want = interpret(slots, steps, [None] * len(steps), data, out)
    got = interpret(slots, fused, progs, data, out)
    assert not np.isnan(got).any()
    np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-14)
    # and the oracle's contraction of the same network
    O.contract_network(to, plan)
    ref = np.asarray(to.backend.load_tensor_data("result")).reshape(-1, order="F")
    np.testing.assert_allclose(got, ref, rtol=1e-10, atol=1e-13)


def test_chain_program_limits():
    cfg = "0110" + "x" * 4 + "1001" + "01" + "x" * 2
    tn, _, _ = _grid(4, 4, 8, 7, cfg)
    plan = PL.reconfigure(tn, PL.sweep_plan(tn, PL.grid_site_order(4, 4, "col")), windows=(8,))
    slots, steps, out = lower_plan(tn, plan)
    dts, steps, slots = assign_precision(slots, steps, _lib.PQ_C128, "mixed", small_elems=16)
    fused, progs, _ = fuse_chains(slots, steps, dts, tile_bits=10, min_elems=64, op_elems=256)
    # a chain's steps re-split with a smaller tile give no tile above it
    for p in progs:
        if p is not None:
            L, tb, nin, nout, nspec = p[:5]
            assert tb <= 10 and nin <= tb and nout <= tb
    # the whole network as one chain exceeds any tile
    idx = [i for i, s in enumerate(steps) if s[0] == 0]
    assert chain_program(slots, steps, idx[-3:], tile_bits=2) is None


def test_step_choices_follow_the_members_of_each_fused_step():
    """Per-step kernel choices after fusion come from the plan step each
    fused step stands for, also when an unrelated contraction step falls
    between two members of one chain."""
    from picoquant_hip.executor import step_choices
    # before fusion: contraction steps c0 c1 c2 c3 (plan steps 0..3) and a
    # conversion between c1 and c2; c0 and c2 form one chain
    pre = [(0,), (0,), (1,), (0,), (0,)]
    members = [[1], [], [0, 3], [4]]          # c1, the conversion, chain(c0, c2), c3
    kcs, tls = step_choices(pre, members, kc=[10, 11, 12, 13], tiles=[20, 21, 22, 23])
    assert kcs == [11, -1, -1, 13]
    assert tls == [21, -1, -1, 23]
    assert step_choices(pre, members) == ([-1] * 4, [-1] * 4)
    # and on a real fused plan: every unfused contraction step keeps its own
    cfg = "0110" + "x" * 4 + "1001" + "01" + "x" * 2
    tn, _, _ = _grid(4, 4, 8, 7, cfg)
    plan = PL.reconfigure(tn, PL.sweep_plan(tn, PL.grid_site_order(4, 4, "col")), windows=(8,))
    slots, steps, out = lower_plan(tn, plan)
    dts, steps, slots = assign_precision(slots, steps, _lib.PQ_C128, "mixed", small_elems=16)
    fused, progs, members = fuse_chains(slots, steps, dts, tile_bits=10, min_elems=64, op_elems=256)
    kcs, _ = step_choices(steps, members, kc=list(range(len(plan))))
    contract = [i for i, s in enumerate(steps) if s[0] == 0]
    for st, mem, k in zip(fused, members, kcs):
        if st[0] == 0:
            assert k == contract.index(mem[0])
        else:
            assert k == -1
