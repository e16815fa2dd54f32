"""Plan selection with the learned cost model.

For a grid-circuit network: generate candidate contraction orders (world-line
+ sweep plans over column / row snakes and seeded perturbations of them, plus
Boltzmann-greedy samples of the reference's planner), rank all of them by the
MLP's predicted run time, slice a shortlist (planner.choose_slices) and rank
the sliced plans again.  The reference's greedy plan (bgreedy, tau -> 0) is
always a candidate and is the fallback when no candidate meets the slice
size cap.  The pick the FLOP count alone would make is kept for comparison.
"""
from __future__ import annotations

import math

import numpy as np

from . import planner as PL


def _perturbed(order, rng, swaps, keep_last):
    o = list(order)
    for _ in range(swaps):
        i = int(rng.integers(0, len(o) - 1))
        o[i], o[i + 1] = o[i + 1], o[i]
    ks = set(keep_last)
    return [q for q in o if q not in ks] + [q for q in o if q in ks]


def candidate_plans(tn, rows, cols, open_q, n_candidates=64, seed=0, n_greedy=2, n_reconf=2, n_partition=0):
    """[(name, plan)]: n_candidates candidate pair plans for a grid network.
    The column and row sweeps also enter after subtree reconfiguration
    (planner.reconfigure, "-rc"), and so do the first n_reconf seeded
    perturbations of them ("-rcp<k>", one window size).  n_partition > 0
    adds that many top-down partition plans (planner.partition_plan, one
    seed each, "partition-<k>"); off by default: on grid circuits their cost
    is far above the sweeps' (balanced cuts cross every qubit line)."""
    rng = np.random.default_rng(seed)
    base = {"col": PL.grid_site_order(rows, cols, "col", open_last=open_q),
            "row": PL.grid_site_order(rows, cols, "row", open_last=open_q)}
    cands = [(f"sweep-{k}", PL.sweep_plan(tn, v)) for k, v in base.items()]
    for name, plan in list(cands):
        cands.append((f"{name}-rc", PL.reconfigure(tn, plan)))
    prng = np.random.default_rng(seed + 1)
    for k in range(n_reconf):
        name = "col" if k % 2 == 0 else "row"
        plan = PL.sweep_plan(tn, _perturbed(base[name], prng, 2 + k, open_q))
        cands.append((f"sweep-{name}-rcp{k}", PL.reconfigure(tn, plan, windows=(12,))))
    for k, v in base.items():
        rev = list(reversed([q for q in v if q not in set(open_q)])) + [q for q in v if q in set(open_q)]
        cands.append((f"sweep-{k}-rev", PL.sweep_plan(tn, rev)))
    for k in range(n_partition):
        cands.append((f"partition-{k}", PL.partition_plan(tn, imbalance=0.25, leaf=24, seed=seed + k)))
    plan0, _, _ = PL.bgreedy_single(tn, 1.0, 0.0)
    cands.append(("bgreedy-tau0", plan0))
    for g in range(n_greedy):
        alpha = [0.5, 1.5, 1.0, 2.0][g % 4]
        plan, _, _ = PL.bgreedy_single(tn, alpha, 1e6 * (g + 1), rng)
        cands.append((f"bgreedy-a{alpha}-t{g}", plan))
    k = 0
    while len(cands) < n_candidates:
        name = "col" if k % 2 == 0 else "row"
        swaps = 1 + (k // 2) % 6
        cands.append((f"sweep-{name}-p{k}", PL.sweep_plan(tn, _perturbed(base[name], rng, swaps, open_q))))
        k += 1
    out = cands[:n_candidates]
    if not any(name == "bgreedy-tau0" for name, _ in out):
        # the greedy plan is the selector's fallback: it is always a candidate
        out[-1] = ("bgreedy-tau0", plan0)
    return out


def select_plan(tn, rows, cols, open_q, n_sliced, model=None, n_candidates=64, shortlist=8,
                max_elems=2 ** 28, seed=0, measure=None, n_measure=3):
    """Pick a plan + slicing.  model: a trained costmodel.MLPCostModel, or
    None to rank by the roofline-time slicing objective alone.  measure:
    callable (plan, sliced) -> measured ms; when given, the n_measure best
    predicted candidates and the FLOP pick are run and the fastest measured
    one is chosen (the model prunes 64 candidates to a few runs; measured
    time has the last word)."""
    cands = candidate_plans(tn, rows, cols, open_q, n_candidates, seed)
    rows_out = []
    for name, plan in cands:
        led = PL.plan_ledger(tn, plan)
        pred = model.plan_time(tn, plan) if model is not None and led["max_size"] <= 2 ** 34 else math.inf
        rows_out.append({"name": name, "plan": plan, "macs": led["macs"], "max": led["max_size"], "pred_us": pred})
    # shortlist: the best predicted (or the cheapest by MACs without a model)
    key = (lambda r: r["pred_us"]) if model is not None else (lambda r: r["macs"])
    order = sorted(rows_out, key=key)
    short = order[:shortlist]
    by_macs = min(rows_out, key=lambda r: r["macs"])
    if by_macs not in short:
        short.append(by_macs)
    greedy = next(r for r in rows_out if r["name"] == "bgreedy-tau0")
    for r in short:
        try:
            sl, tot, peak = PL.choose_slices(tn, r["plan"], n_sliced, max_elems=max_elems)
        except ValueError:
            r["sliced"], r["sliced_macs"], r["peak"], r["pred_sliced_us"] = None, math.inf, math.inf, math.inf
            continue
        r["sliced"], r["sliced_macs"], r["peak"] = sl, tot, peak
        r["pred_sliced_us"] = model.plan_time(tn, r["plan"], sl) if model is not None else tot
    ok = [r for r in short if r.get("sliced") is not None and r["peak"] <= max_elems]
    measured = {}
    model_pick = None
    if ok:
        # predictions within 2% of the best count as a tie, broken by the
        # MAC count and then by candidate order (a model retrained on fresh
        # timings then keeps picking the same plan)
        top = min(r["pred_sliced_us"] for r in ok)
        tied = [r for r in ok if r["pred_sliced_us"] <= top * 1.02]
        best = model_pick = min(tied, key=lambda r: r["sliced_macs"])
        flop_pick = min(ok, key=lambda r: r["sliced_macs"])
        fallback = False
        if measure is not None:
            runs = [best] + [r for r in sorted(ok, key=lambda r: r["pred_sliced_us"]) if r is not best]
            runs = runs[:n_measure]
            if flop_pick not in runs:
                runs.append(flop_pick)
            for r in runs:
                r["measured_ms"] = float(measure(r["plan"], r["sliced"]))
                measured[r["name"]] = round(r["measured_ms"], 4)
            best = min(runs, key=lambda r: r["measured_ms"])
    else:
        sl, tot, peak = PL.choose_slices(tn, greedy["plan"], n_sliced, max_elems=max_elems)
        greedy.update(sliced=sl, sliced_macs=tot, peak=peak, pred_sliced_us=math.nan)
        best = flop_pick = greedy
        fallback = True
    summary = {
        "choice": best["name"], "ranked_by": "mlp" if model is not None else "roofline slicing objective",
        "fallback_greedy": fallback, "candidates": len(rows_out),
        "unsliced_macs_log2": round(math.log2(best["macs"]), 3),
        "sliced_macs_log2": round(math.log2(best["sliced_macs"]), 3),
        "peak_slice_elems_log2": round(math.log2(best["peak"]), 2),
        "predicted_ms": None if model is None else round(best["pred_sliced_us"] / 1e3, 3),
        "flop_pick": flop_pick["name"],
        "model_pick": model_pick["name"] if model_pick is not None else None,
        "measured_ms": measured,
        "greedy_unsliced_macs_log2": round(math.log2(greedy["macs"]), 2),
        "shortlist": [(r["name"], round(math.log2(r["macs"]), 2),
                       None if model is None else round(r.get("pred_sliced_us", math.inf) / 1e3, 3))
                      for r in short],
    }
    return {"plan": best["plan"], "sliced": best["sliced"], "summary": summary,
            "flop_plan": flop_pick["plan"], "flop_sliced": flop_pick["sliced"], "rows": rows_out,
            "model_plan": (model_pick or best)["plan"], "model_sliced": (model_pick or best)["sliced"],
            "flop_ms": flop_pick.get("measured_ms"), "model_ms": (model_pick or best).get("measured_ms"),
            "chosen_ms": best.get("measured_ms")}
