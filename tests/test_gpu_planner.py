"""The learned planner against FLOP counting (the reference ranks plans by
their cost alone, src/layer2/bgreedy_contraction.jl:53-66), on held-out
grid circuits (seeds 31-36, not in the cost model's training circuits).

The cost model (trained here on this GPU's own timings) ranks 64 candidate
plans of each circuit; the three best predicted and the FLOP pick are
compiled and timed, and the fastest measured one is the pick.  Measured on
these circuits, the plan with the fewest MACs runs up to 5-10 % slower than
another candidate (launch-bound runs of 0.1-0.4 ms, where a plan with more
MACs but fewer or cheaper launches wins; the gap shrinks as the executor
cuts launch costs: 10-26 % with a grouped launch per dependency level, ~5 %
once small steps of up to 2^17 outputs shared grouped launches, 1-3 % on
5x5 depth 12 with the world lines' intermediates kept in LDS and wider
grouped launches -- within timing noise, so no win is asserted any more).
Asserted: the pick is never slower than the FLOP pick beyond 3 % timing
noise.  The gains are printed.  The model's own first choice is
reported beside.  The greedy fallback: no candidate meets an impossible
size cap, so the pick is bgreedy (tau -> 0) and it contracts correctly."""
import time

import numpy as np
import pytest

from oracle import pq_oracle as O
import picoquant_hip as P
from picoquant_hip import _lib, circuits as C, costmodel as CM, learned
from picoquant_hip.executor import compile_plan

pytestmark = pytest.mark.gpu

CIRCUITS = [(4, 5, 10, 31), (5, 5, 12, 33), (6, 6, 10, 36)]


def _net(r, c, d, s):
    n = r * c
    tn = P.network_from_gates(n, C.grid_circuit_gates(r, c, d, s), backend=P.HIPBackend(np.complex128),
                              decompose=True)
    P.add_input(tn, "0" * n)
    oq = [rr * c + c for rr in range(r)]
    P.add_output(tn, "".join("x" if (q + 1) in oq else str(b) for q, b in enumerate(C.seeded_bits(n, s + 1))))
    return tn, oq


def _time(tn, plan, sliced, runs=20):
    p = compile_plan(tn, plan, sliced, precision="mixed")
    ctx = tn.backend.ctx
    try:
        for _ in range(3):
            p.execute(True)
        ctx.synchronize()
        ts = []
        for _ in range(runs):
            t = time.perf_counter()
            p.execute(True)
            ctx.synchronize()
            ts.append(time.perf_counter() - t)
        return 1e3 * float(np.median(ts))
    finally:
        p.close()


@pytest.fixture(scope="module")
def model():
    ctx = _lib.Context(np.complex128)
    plans = CM.grid_training_plans(specs=((4, 5, 10, 11), (5, 5, 12, 12), (5, 6, 12, 13)), per_circuit=12,
                                   slicings=(0,))
    X, us, _ = CM.collect_dataset(ctx, plans, n_samples=6000, seed=0)
    m = CM.MLPCostModel(ctx, hidden=64, seed=0)
    m.fit(X, us, epochs=150)
    yield m
    m.close()
    ctx.close()


def test_learned_pick_never_loses_to_the_flop_pick(model):
    gains = {}
    for spec in CIRCUITS:
        tn, oq = _net(*spec)
        sel = learned.select_plan(tn, spec[0], spec[1], oq, 0, model=model, n_candidates=24,
                                  measure=lambda plan, sl: _time(tn, plan, sl, runs=10))
        assert not sel["summary"]["fallback_greedy"]
        # re-time the two picks back to back, more runs
        t_pick = _time(tn, sel["plan"], sel["sliced"])
        t_flop = _time(tn, sel["flop_plan"], sel["flop_sliced"])
        t_model = _time(tn, sel["model_plan"], sel["model_sliced"])
        gains[spec] = t_flop / t_pick
        print(f"{spec}: pick {sel['summary']['choice']} {t_pick:.3f} ms, FLOP pick "
              f"{sel['summary']['flop_pick']} {t_flop:.3f} ms, model's first choice "
              f"{sel['summary']['model_pick']} {t_model:.3f} ms")
        assert t_pick <= 1.03 * t_flop, (spec, t_pick, t_flop)
    print("gains over the FLOP pick:", gains)


def test_greedy_fallback_when_no_candidate_fits(model):
    spec = (3, 3, 6, 21)
    tn, oq = _net(*spec)
    sel = learned.select_plan(tn, 3, 3, oq, 0, model=model, n_candidates=8, max_elems=2)
    assert sel["summary"]["fallback_greedy"] and sel["summary"]["choice"] == "bgreedy-tau0"
    p = compile_plan(tn, sel["plan"], sel["sliced"])
    p.execute(True)
    got = p.result().reshape(-1, order="F")
    p.close()
    cfg = "".join("x" if (q + 1) in oq else str(b) for q, b in enumerate(C.seeded_bits(9, 22)))
    to = O.network_from_gates(9, O.grid_circuit_gates(3, 3, 6, 21), decompose=True)
    O.add_input(to, "0" * 9)
    O.add_output(to, cfg)
    O.contract_network(to, O.greedy_plan(to)[0])
    want = np.asarray(to.backend.load_tensor_data("result")).reshape(-1, order="F")
    np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-15)
