"""Plan selection: candidate contraction orders + slicings, ranked by a cost
model; the reference's greedy planner (bgreedy, tau -> 0) is always a
candidate and the fallback.

Candidates: world-line + sweep plans over several site orders (column/row
snakes and seeded perturbations of them) and Boltzmann-greedy samples.  Each
candidate gets the greedy slicing of planner.choose_slices.  The ranking uses
the trained MLP cost model when one is loaded (predicted kernel time per
step, summed), otherwise the exact sliced MAC count.
"""
from __future__ import annotations

import math

import numpy as np

from . import planner as PL


def _perturbed_order(order, rng, swaps):
    o = list(order)
    for _ in range(swaps):
        i = int(rng.integers(0, len(o) - 1))
        o[i], o[i + 1] = o[i + 1], o[i]
    return o


def candidate_plans(tn, rows, cols, open_q, n_candidates=8, seed=0, with_greedy=True):
    """[(name, plan)] candidate pair plans for a grid network."""
    rng = np.random.default_rng(seed)
    cands = []
    base = {"col": PL.grid_site_order(rows, cols, "col", open_last=open_q),
            "row": PL.grid_site_order(rows, cols, "row", open_last=open_q)}
    for name, order in base.items():
        cands.append((f"sweep-{name}", PL.sweep_plan(tn, order)))
    if with_greedy:
        plan, _, _ = PL.bgreedy_single(tn, 1.0, 0.0)
        cands.append(("bgreedy-tau0", plan))
    k = 0
    while len(cands) < n_candidates:
        name = "col" if k % 2 == 0 else "row"
        order = _perturbed_order(base[name], rng, 1 + k // 2)
        # keep the open sites last
        order = [q for q in order if q not in set(open_q)] + [q for q in order if q in set(open_q)]
        cands.append((f"sweep-{name}-perturb{k}", PL.sweep_plan(tn, order)))
        k += 1
    return cands


def select_plan(tn, rows, cols, open_q, n_sliced, n_candidates=6, model=None, max_elems=2 ** 28):
    """Pick a plan + slicing.  Returns a dict with plan, sliced bonds and a
    summary of the ranking."""
    cands = candidate_plans(tn, rows, cols, open_q, n_candidates)
    scored = []
    for name, plan in cands:
        led = PL.plan_ledger(tn, plan)
        scored.append((led["macs"], name, plan, led))
    scored.sort(key=lambda x: x[0])
    # slice the cheapest few (slicing search is the expensive part)
    best = None
    for macs, name, plan, led in scored[:3]:
        sliced, tot, peak = PL.choose_slices(tn, plan, n_sliced, max_elems=max_elems)
        score = tot if model is None else model.plan_time(tn, plan, sliced)
        if best is None or score < best[0]:
            best = (score, name, plan, sliced, tot, peak, macs)
    score, name, plan, sliced, tot, peak, macs = best
    summary = {"choice": name, "ranked_by": "mlp" if model is not None else "sliced MACs",
               "unsliced_macs_log2": round(math.log2(macs), 3), "sliced_macs_log2": round(math.log2(tot), 3),
               "peak_slice_elems_log2": round(math.log2(peak), 2),
               "candidates": [(n, round(math.log2(m), 2)) for m, n, _, _ in scored]}
    return {"plan": plan, "sliced": sliced, "summary": summary}
