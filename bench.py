"""Benchmark: amplitudes/s on the 36-qubit (6x6 grid) depth-14 random circuit.

Workload (one step = one full run):
  * circuit: 6x6 grid, 14 cycles, seed 2026 (sqrtX/sqrtY/sqrtW + CZ patterns
    ACBD), input |0...0>;
  * 64 output bitstrings: the last grid column (qubits 6, 12, ..., 36) left
    open, the other 30 bits seeded (PCG64, seed 2027);
  * plan: the planner's pick among candidate orders (world-line + sweep
    plans, Boltzmann-greedy variants) with 10 sliced bonds -> 1024 slices;
  * batched path: every slice and all 64 amplitudes in one hipGraph launch,
    complex64 on FP32 MFMA, FP64 cross-slice sum.
Also reported: relative error against an FP64 run of the same plan, the plain
FP64 GPU path (the reference's drivers on HIPBackend{ComplexF64}, one slice
timed and scaled to 1024), a roofline for the dominant kernel class, and a
CPU baseline (the numpy restatement of InteractiveBackend{ComplexF64} on a
bounded sample).

    python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import json
import math
import os
import pathlib
import sys
import time

ROOT = pathlib.Path(__file__).resolve().parent


def _host_threads():
    """Host threads for the CPU baseline: the cores this process may run on,
    capped by OMP_NUM_THREADS when the host sets it (a shared machine's
    share for one GPU)."""
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    cap = os.environ.get("OMP_NUM_THREADS", "")
    return max(1, min(n, int(cap))) if cap.isdigit() and int(cap) > 0 else max(1, n)


CPU_THREADS = _host_threads()
os.environ.setdefault("OPENBLAS_NUM_THREADS", str(CPU_THREADS))
os.environ.setdefault("OMP_NUM_THREADS", str(CPU_THREADS))
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "picoquant.jl_amd"))

import numpy as np  # noqa: E402

ROWS, COLS, DEPTH, SEED = 6, 6, 14, 2026
N_SLICED = 10
METRIC = "amplitudes/s, 36-qubit d14 grid, batched+learned vs plain GPU at ≤1e-5 rel err"
REL_ERR_BOUND = 1e-5          # worst single amplitude against the FP64 run of the same plan
BINDING_CAP = 2 ** 22         # per-slice size cap of the binding-slicing variant
PEAK = {"mfma_fp32": 157.3e12, "mfma_fp64": 78.6e12, "hbm": 8.0e12}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def build_network(dtype, backend=None):
    import picoquant_hip as P
    from picoquant_hip import circuits as C
    n = ROWS * COLS
    gates = C.grid_circuit_gates(ROWS, COLS, DEPTH, SEED)
    tn = P.network_from_gates(n, gates, backend=backend or P.HIPBackend(dtype), decompose=True)
    P.add_input(tn, "0" * n)
    bits = C.seeded_bits(n, SEED + 1)
    open_q = [r * COLS + COLS for r in range(ROWS)]
    cfg = ["x" if (q + 1) in open_q else str(bits[q]) for q in range(n)]
    P.add_output(tn, "".join(cfg))
    return tn, open_q


def train_cost_model(ctx, n_samples=20000):
    """The learned cost model, trained here on this GPU: ~20k contraction
    steps drawn from candidate plans of 20-36-qubit grid circuits (seeds
    11-14, not the benchmark circuit), timed with pq_time_step, fitted with
    the HIP MLP kernels.  Returns (model, stats)."""
    from picoquant_hip import costmodel as CM
    t0 = time.perf_counter()
    plans = CM.grid_training_plans()
    X, us, keys = CM.collect_dataset(ctx, plans, n_samples=n_samples, seed=0)
    t1 = time.perf_counter()
    uniq = np.unique(keys)
    rng = np.random.default_rng(0)
    held = set(rng.choice(uniq, size=len(uniq) // 5, replace=False).tolist())
    te = np.array([k in held for k in keys])
    model = CM.MLPCostModel(ctx, hidden=64, seed=0)
    model.fit(X[~te], us[~te], epochs=150)
    sel = te & (us > 12.0)
    pred = model.predict_log_us(X[sel])
    rp = np.argsort(np.argsort(pred))
    rt = np.argsort(np.argsort(np.log(us[sel])))
    stats = {"samples": int(len(X)), "unique_shapes": int(len(uniq)),
             "holdout_spearman": float(np.corrcoef(rp, rt)[0, 1]),
             "holdout_median_abs_ratio_err": float(np.median(np.abs(np.exp(pred - np.log(us[sel])) - 1))),
             "collect_s": round(t1 - t0, 1)}
    final = CM.MLPCostModel(ctx, hidden=64, seed=0)       # deployed model: all samples
    stats["train_loss"] = float(final.fit(X, us, epochs=150))
    stats["fit_s"] = round(time.perf_counter() - t1, 1)
    return final, stats


def run_variant(tn, plan, sliced, precision, group_bits, runs=3, tune=True):
    """Autotune (for this precision), compile and time one plan variant:
    (ms per run, result vector)."""
    from picoquant_hip import autotune
    from picoquant_hip.executor import choose_group_bits, compile_plan, lower_plan
    ctx = tn.backend.ctx
    kc = tiles = None
    if group_bits is None:
        # the executor's own choice (largest slice group whose batched
        # intermediates fit its memory budget), fixed here so the autotuner
        # times the same batch sizes
        slots, _, _ = lower_plan(tn, plan, sliced)
        group_bits = choose_group_bits(slots, len(sliced), 64 << 30, 8 if precision != "uniform" else 16)
    if tune and precision != "uniform":
        kc, tiles, _ = autotune.autotune_plan(ctx, tn, plan, sliced, group_bits=group_bits, precision=precision)
    p = compile_plan(tn, plan, sliced, group_bits=group_bits, kc=kc, tiles=tiles, precision=precision)
    t = time_plan(p, ctx, runs)
    res = p.result().reshape(-1)
    classes = {}
    ms = p.profile()
    from picoquant_hip import _lib
    for i in range(len(p.steps)):
        name = _lib.KC_NAMES[p.step_info(i)["kc"]]
        classes[name] = round(classes.get(name, 0.0) + ms[i], 3)
    p.close()
    return 1e3 * t, res, classes


def grid7x7_config():
    """The 49-qubit config: 7x7 grid, 12 cycles (seed 2027), one amplitude
    of a seeded bitstring, the reconfigured column sweep (the planner's
    lowest-MAC order) with 8 bonds sliced under a per-slice cap of 2^22
    elements -- the unsliced plan peaks at 2^24, so the cap binds and the
    slice-batched steps carry most of the MACs; FP64, complex64 on FP32
    MFMA and FP16 hi/lo split, errors against the FP64 run."""
    import math
    import picoquant_hip as P
    from picoquant_hip import circuits as C, planner as PL
    from picoquant_hip.executor import lower_plan
    r = c = 7
    n = r * c
    seed, cap, nb = 2027, 2 ** 22, 8
    tn = P.network_from_gates(n, C.grid_circuit_gates(r, c, 12, seed), backend=P.HIPBackend(np.complex128),
                              decompose=True)
    P.add_input(tn, "0" * n)
    P.add_output(tn, "".join(map(str, C.seeded_bits(n, seed + 1))))
    plan = PL.reconfigure(tn, PL.sweep_plan(tn, PL.grid_site_order(r, c, "col")))
    unsliced_peak = PL.plan_ledger(tn, plan)["max_size"]
    sliced, macs, peak = PL.choose_slices(tn, plan, nb, max_elems=cap)
    out = {"workload": f"7x7 grid depth-12 random circuit (seed 2027), single amplitude, reconfigured column "
                       f"sweep, {nb} bonds sliced under a per-slice cap of 2^22 = {2 ** nb} slices",
           "sliced_macs_log2": round(math.log2(macs), 2), "peak_slice_elems_log2": round(math.log2(peak), 1),
           "unsliced_peak_elems_log2": round(math.log2(unsliced_peak), 1),
           "slice_dependent_mac_fraction": round(dependent_mac_fraction(tn, plan, sliced), 3)}
    ref = None
    for prec in ("uniform", "mixed", "mixed16"):
        ms, res, cls = run_variant(tn, plan, sliced, prec, None)
        if ref is None:
            ref = res
        err = float(np.max(np.abs(res - ref) / np.abs(ref)))
        out[{"uniform": "fp64", "mixed": "fp32_mfma", "mixed16": "fp16_split"}[prec]] = {
            "ms": round(ms, 3), "rel_err_vs_fp64": err, "within_bound": err <= REL_ERR_BOUND, "classes_ms": cls}
    out["amplitude"] = [float(ref[0].real), float(ref[0].imag)]
    return out


def independent_bitstrings(n_bits=64, runs=3):
    """64 independent bitstrings (every qubit closed; PCG64, seed 2028) of
    the benchmark circuit in one run: the reconfigured column sweep of the
    closed network, 10 sliced bonds batched inside each bitstring's group,
    one group per bitstring, complex64 / FP64 split as the headline, one
    hipGraph per run.  Checked against the FP64 run of the same batch and,
    for two of the bitstrings, against single-amplitude FP64 runs."""
    import picoquant_hip as P
    from picoquant_hip import autotune, circuits as C, planner as PL
    from picoquant_hip.executor import compile_plan
    n = ROWS * COLS
    rng = np.random.Generator(np.random.PCG64(SEED + 2))
    batch = ["".join(str(int(b)) for b in rng.integers(0, 2, n)) for _ in range(n_bits)]
    tn = P.network_from_gates(n, C.grid_circuit_gates(ROWS, COLS, DEPTH, SEED),
                              backend=P.HIPBackend(np.complex128), decompose=True)
    P.add_input(tn, "0" * n)
    P.add_output(tn, batch[0])
    ctx = tn.backend.ctx
    plan = PL.reconfigure(tn, PL.sweep_plan(tn, PL.grid_site_order(ROWS, COLS, "col")))
    sliced, macs, peak = PL.choose_slices(tn, plan, N_SLICED)
    kc, tiles, _ = autotune.autotune_plan(ctx, tn, plan, sliced, group_bits=N_SLICED)
    p = compile_plan(tn, plan, sliced, group_bits=N_SLICED, kc=kc, tiles=tiles, precision="mixed",
                     bitstrings=batch)
    t = time_plan(p, ctx, runs)
    got = p.result().reshape(-1)
    st = p.stats()
    p.close()
    p64 = compile_plan(tn, plan, sliced, group_bits=N_SLICED, bitstrings=batch)
    p64.execute(True)
    ref = p64.result().reshape(-1)
    p64.close()
    single = []
    for b in (0, n_bits - 1):
        t2 = P.network_from_gates(n, C.grid_circuit_gates(ROWS, COLS, DEPTH, SEED),
                                  backend=P.HIPBackend(np.complex128), decompose=True)
        P.add_input(t2, "0" * n)
        P.add_output(t2, batch[b])
        q = compile_plan(t2, plan, sliced, group_bits=N_SLICED)
        q.execute(True)
        single.append(abs(complex(q.result().reshape(-1)[0]) - ref[b]) / abs(ref[b]))
        q.close()
    return {"workload": f"{n_bits} independent bitstrings (PCG64 seed {SEED + 2}), every qubit closed, reconfigured "
                        f"column sweep, {N_SLICED} sliced bonds, one group per bitstring, one hipGraph per run",
            "ms_per_run": round(1e3 * t, 3), "amps_per_s": n_bits / t, "launches_per_run": st["launches"],
            "macs_per_amplitude_log2": round(math.log2(macs), 3), "peak_slice_elems_log2": round(math.log2(peak), 1),
            "rel_err_vs_fp64": float(np.max(np.abs(got - ref) / np.abs(ref))),
            "fp64_batch_vs_single_runs": float(max(single))}


def dependent_mac_fraction(tn, plan, sliced):
    """Fraction of a sliced plan's MACs (all slices) in steps that depend on
    a sliced bond, i.e. that run in the slice-batched part of the program."""
    from picoquant_hip.executor import lower_plan
    slots, steps, _ = lower_plan(tn, plan, sliced)
    dep = tot = 0
    for sa, sb, sc, M, N, K in steps:
        m = math.prod(e for e, _, _ in M + N + K) << len(slots[sc].deps)
        tot += m
        if slots[sc].deps:
            dep += m
    return dep / tot


def mps_chain_config():
    """MPS contraction (contract_mps_tensor_network_circuit!, src/layer2.jl:570-625,
    the reference's own GPU benchmark target): a 1 x 20 chain random circuit
    (the grid construction on one row, so every coupler joins neighbours),
    depth 64, seed 2026, bond dimension capped at 256, ComplexF64.  GPU: the
    drivers with pq_contract / pq_decompose (rocSOLVER), second of two runs;
    CPU: the numpy oracle's restatement, one run.  Both time the driver only
    (network construction excluded).  The GPU amplitudes are checked against
    the oracle's run with the same truncation."""
    import picoquant_hip as P
    from picoquant_hip import circuits as C
    from oracle import pq_oracle as O
    n, depth, seed, chi = 20, 64, 2026, 256
    t_gpu, psi, bond = None, None, 0
    for _ in range(2):
        tn = P.network_from_gates(n, C.grid_circuit_gates(1, n, depth, seed), backend=P.HIPBackend(np.complex128),
                                  decompose=True)
        P.add_input(tn, "0" * n)
        ctx = tn.backend.ctx
        ctx.synchronize()
        t0 = time.perf_counter()
        nodes = P.contract_mps_tensor_network_circuit(tn, max_rank=chi)
        ctx.synchronize()
        t_gpu = time.perf_counter() - t0
        bond = max(d for v in nodes for x, d in zip(tn.nodes[v].indices, tn.nodes[v].dims) if tn.edges[x].virtual)
        P.calculate_mps_amplitudes(tn, nodes)
        psi = np.asarray(tn.backend.load_tensor_data("result")).reshape(-1, order="F")
    to = O.network_from_gates(n, O.grid_circuit_gates(1, n, depth, seed), decompose=True)
    O.add_input(to, "0" * n)
    t0 = time.perf_counter()
    onodes = O.contract_mps_tensor_network_circuit(to, max_rank=chi)
    t_cpu = time.perf_counter() - t0
    O.calculate_mps_amplitudes(to, onodes)
    ref = np.asarray(to.backend.load_tensor_data("result")).reshape(-1, order="F")
    return {"workload": f"MPS of a 1x{n} chain random circuit, depth {depth} (seed {seed}), bond cap {chi}, "
                        "complex128", "gpu_s": round(t_gpu, 4), "cpu_oracle_s": round(t_cpu, 3),
            "speedup": round(t_cpu / t_gpu, 2), "max_bond": int(bond),
            "rel_err_vs_oracle": float(np.linalg.norm(psi - ref) / np.linalg.norm(ref))}


def fp64_fidelity(psi, ref):
    """|<ref|psi>|^2 / (<ref|ref><psi|psi>) of two amplitude batches, in FP64."""
    psi = np.asarray(psi, dtype=np.complex128).reshape(-1)
    ref = np.asarray(ref, dtype=np.complex128).reshape(-1)
    return float(abs(np.vdot(ref, psi)) ** 2 / (np.vdot(ref, ref).real * np.vdot(psi, psi).real))


def slicing_variants(tn, plan, ref):
    """The benchmark plan unsliced, and sliced over 10 bonds under a per-slice
    cap of 2^22 elements (the plan peaks at 2^26, so this cap binds and the
    slice-batched part of the program carries most of the MACs); each in
    the benchmark's precision, autotuned, against the FP64 result ref."""
    import picoquant_hip.planner as PL
    out = {}
    for name, nb, cap in (("unsliced", 0, 2 ** 28), ("binding_cap_2^22", N_SLICED, BINDING_CAP)):
        if nb:
            sl, macs, peak = PL.choose_slices(tn, plan, nb, max_elems=cap)
        else:
            led = PL.plan_ledger(tn, plan)
            sl, macs, peak = [], led["macs"], led["max_size"]
        ms, res, cls = run_variant(tn, plan, sl, "mixed", None)
        v = {"slices": 2 ** nb, "ms_per_run": round(ms, 3), "amps_per_s": 64e3 / ms,
             "macs_log2": round(math.log2(macs), 3), "peak_slice_elems_log2": round(math.log2(peak), 1),
             "slice_dependent_mac_fraction": round(dependent_mac_fraction(tn, plan, sl), 3) if nb else 0.0,
             "classes_ms": cls}
        if ref is not None:
            v["rel_err_vs_fp64"] = float(np.max(np.abs(res - ref) / np.abs(ref)))
        out[name] = v
    return out


def choose_plan(tn, open_q, model=None):
    """The learned planner's pick: the cost model ranks 64 candidates and
    their slicings, then the three best predicted and the FLOP pick are
    compiled (benchmark precision, default kernel choices) and timed; the
    fastest measured wins."""
    from picoquant_hip import learned
    from picoquant_hip.executor import compile_plan
    ctx = tn.backend.ctx

    def measure(plan, sliced):
        p = compile_plan(tn, plan, sliced, group_bits=N_SLICED, precision="mixed")
        try:
            return 1e3 * time_plan(p, ctx, 5)
        finally:
            p.close()

    return learned.select_plan(tn, ROWS, COLS, open_q, N_SLICED, model=model,
                               measure=None if model is None else measure)


def time_plan(p, ctx, runs):
    p.execute(True)
    ctx.synchronize()
    t = time.perf_counter()
    for _ in range(runs):
        p.execute(True)
    ctx.synchronize()
    return (time.perf_counter() - t) / runs


def plain_fp64_slice_time(plan, sliced, open_q):
    """The reference's drivers on HIPBackend{ComplexF64}: rebuild the network,
    slice_tensor_network at slice 0, contract_network! with the same plan."""
    import picoquant_hip as P
    from picoquant_hip.slicing import slice_with_map, translate_plan
    tn, _ = build_network(np.complex128)
    old = tn.counters["node"]
    t0 = time.perf_counter()
    m = slice_with_map(tn, sliced, [1] * len(sliced))
    plan2 = translate_plan(plan, m, old, tn.counters["node"])
    P.contract_network(tn, plan2, "vector")
    res = tn.backend.load_tensor_data("result")
    return time.perf_counter() - t0, res


def cpu_baseline(plan, sliced, cap_s=20.0):
    """oracle/pq_oracle.py (numpy: permute + BLAS zgemm, InteractiveBackend's
    algorithm) on slice 0 of the same sliced network and plan, stopped after
    cap_s seconds; the time is scaled by the MACs done to the whole slice."""
    from oracle import pq_oracle as O
    from picoquant_hip import circuits as C
    from picoquant_hip import planner as PL
    from picoquant_hip.slicing import translate_plan
    n = ROWS * COLS
    tn = O.network_from_gates(n, O.grid_circuit_gates(ROWS, COLS, DEPTH, SEED), decompose=True)
    O.add_input(tn, "0" * n)
    bits = C.seeded_bits(n, SEED + 1)
    open_q = [r * COLS + COLS for r in range(ROWS)]
    O.add_output(tn, "".join("x" if (q + 1) in open_q else str(bits[q]) for q in range(n)))
    old = tn.counters["node"]
    label_map = {}
    for b in sliced:
        for nd in (tn.edges[b].src, tn.edges[b].dst):
            view = f"node_{tn.counters['node'] + 1}"
            O.replace_with_view(tn, nd, b, (1, 1))
            label_map[nd] = view
    def cur(x):
        while x in label_map:
            x = label_map[x]
        return x
    plan2 = translate_plan(plan, {k: cur(k) for k in label_map}, old, tn.counters["node"])
    led = [c * k for _, _, c, k in PL.plan_ledger(tn, plan2)["steps"]]
    total = sum(led)
    done = 0
    t0 = time.perf_counter()
    steps = 0
    for (a, b), macs in zip(plan2, led):
        O.contract_pair(tn, a, b)
        done += macs
        steps += 1
        if time.perf_counter() - t0 > cap_s:
            break
    el = time.perf_counter() - t0
    t_slice = el * total / max(done, 1)
    result = None
    if steps == len(plan2):
        # the whole slice was contracted: finish as contract_network! does
        # (permute to the output qubits' order; untimed)
        O.contract_network(tn, [], "vector")        # nothing left to contract: the final permute
        result = np.asarray(tn.backend.load_tensor_data("result"), dtype=np.complex128).reshape(-1, order="F")
    return {"t_slice_s": t_slice, "steps_done": steps, "steps": len(plan2), "mac_fraction": done / total,
            "elapsed_s": el, "result": result}


PROFILES = ROOT / "profiles"


def measured_traffic(kernel_class, d):
    """(HBM bytes per launch, detail) of the dominant kernel class from the
    committed rocprofv3 FETCH_SIZE / WRITE_SIZE runs (scaled by the
    copy-kernel calibration); None with the reason in the detail when no
    summary for this class and plan is committed."""
    traffic_file = PROFILES / f"{kernel_class}_grid6x6d14_traffic.json"
    if not traffic_file.exists():
        return None, {"missing": "no profiles/" + traffic_file.name, "alg_bytes_per_run": d["bytes"]}
    t = json.loads(traffic_file.read_text())
    if t.get("kernel_class") != kernel_class:
        return None, None
    # the summary holds for the plan it was measured on: same class bytes
    prof_alg = t.get("class_alg_bytes_per_run")
    if not prof_alg or abs(prof_alg - d["bytes"]) > 0.01 * d["bytes"]:
        return None, {"stale": "profiles/" + traffic_file.name + " was measured on another plan",
                      "alg_bytes_per_run": d["bytes"], "profiled_alg_bytes_per_run": prof_alg}
    per_launch = t["hbm_bytes_per_run"] / t["launches_per_run"]
    alg_per_launch = d["bytes"] / d["steps"]
    return per_launch, {"hbm_bytes_per_run": t["hbm_bytes_per_run"], "alg_bytes_per_run": d["bytes"],
                        "hbm_over_alg": t["hbm_bytes_per_run"] / d["bytes"],
                        "alg_bytes_per_launch": alg_per_launch, "source": "profiles/" + traffic_file.name}


def calibrate_copy(mib):
    """Strided-copy kernel over a known number of bytes: reads and writes
    mib MiB each per launch, 10 launches."""
    from picoquant_hip import _lib
    ctx = _lib.Context(np.complex64)
    n = (mib << 20) // 8
    h = ctx.upload(np.zeros(n, dtype=np.complex64))
    for _ in range(10):
        out = ctx.materialize(h)
        ctx.free(out)
    ctx.synchronize()
    print(json.dumps({"calibrate_copy_bytes_read": n * 8, "bytes_written": n * 8, "launches": 10}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-model", action="store_true",
                    help="rank candidates by the roofline slicing objective instead of training the "
                         "cost model (for profiler runs; the same plan wins on this circuit)")
    ap.add_argument("--tuned-cache", default="", metavar="JSON",
                    help="reuse the autotuner's per-step choices from this file (written when missing), "
                         "so profiler runs trace only the plan's own kernels")
    ap.add_argument("--no-grid7", action="store_true", help="skip the 7x7 depth-12 precision comparison")
    ap.add_argument("--no-mps", action="store_true", help="skip the MPS chain contraction timing")
    ap.add_argument("--no-fp64-check", action="store_true",
                    help="skip the FP64 reference run (and the FP16-split variant that is checked against it); "
                         "for profiler runs that should hold only the complex64 plan's kernels")
    ap.add_argument("--no-bitstrings", action="store_true",
                    help="skip the batch of 64 independent bitstrings")
    ap.add_argument("--no-slicing-variants", action="store_true",
                    help="skip the unsliced and binding-cap variants of the benchmark plan")
    ap.add_argument("--no-split16", action="store_true",
                    help="skip timing the FP16 hi/lo split variant of the benchmark plan")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run the strided-copy kernel over a MIB-MiB complex64 tensor 10 times "
                         "(scale factor for the HBM byte counters)")
    args = ap.parse_args()
    if args.calibrate_copy:
        calibrate_copy(args.calibrate_copy)
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}); the slices would shard "
              "across GPUs with one FP64 reduction at the end, which this build does not run.")
        sys.exit(2)
    import picoquant_hip as P
    from picoquant_hip import _lib
    from picoquant_hip.executor import compile_plan

    from picoquant_hip import autotune
    t0 = time.perf_counter()
    tn, open_q = build_network(np.complex128)
    ctx = tn.backend.ctx
    if args.no_model:
        model, mstats = None, {"skipped": "--no-model"}
    else:
        model, mstats = train_cost_model(ctx)
    log(f"[bench] cost model {time.perf_counter() - t0:.1f}s: {mstats}")
    sel = choose_plan(tn, open_q, model)
    plan, sliced = sel["plan"], sel["sliced"]
    log(f"[bench] plan {time.perf_counter() - t0:.1f}s: {sel['summary']}")
    cache = pathlib.Path(args.tuned_cache) if args.tuned_cache else None
    if cache is not None and cache.exists():
        kt = json.loads(cache.read_text())
        kcs, tiles, tuned = kt["kc"], kt["tiles"], kt["report"]
        log(f"[bench] autotune choices read from {cache}")
    else:
        kcs, tiles, tuned = autotune.autotune_plan(ctx, tn, plan, sliced, group_bits=N_SLICED)
        if cache is not None:
            cache.write_text(json.dumps({"kc": kcs, "tiles": tiles, "report": tuned}))
    log(f"[bench] autotuned {len(tuned)} steps {time.perf_counter() - t0:.1f}s")
    p = compile_plan(tn, plan, sliced, group_bits=N_SLICED, kc=kcs, tiles=tiles, precision="mixed")
    st = p.stats()
    log(f"[bench] compiled: arena {st['arena_bytes'] / 2**30:.2f} GiB, {st['launches']} launches/run, "
        f"group_bits {p.group_bits}")
    for _ in range(args.warmup):
        p.execute(True)
    ctx.synchronize()
    t1 = time.perf_counter()
    for _ in range(args.steps):
        p.execute(True)
    ctx.synchronize()
    dt = time.perf_counter() - t1
    amps = 64 * args.steps / dt
    res64 = p.result().reshape(-1)
    log(f"[bench] timed {args.steps} runs in {dt:.3f}s -> {amps:.1f} amplitudes/s")

    # Roofline of the dominant kernel class from an event-timed run.
    ms = p.profile()
    per = {}
    for i, (kind, sa, sb, sc, M, N, K) in enumerate(p.steps):
        inf = p.step_info(i)
        e = per.setdefault(inf["kc"], {"ms": 0.0, "flops": 0.0, "bytes": 0.0, "steps": 0,
                                       "flop_s": 0.0, "byte_s": 0.0})
        e["ms"] += ms[i]
        e["flops"] += inf["flops"]
        e["bytes"] += inf["bytes"]
        e["steps"] += 1
        peak = PEAK["mfma_fp64"] if p.dtypes[sc] == _lib.PQ_C128 else PEAK["mfma_fp32"]
        e["flop_s"] += inf["flops"] / peak
        e["byte_s"] += inf["bytes"] / PEAK["hbm"]
    dom = max(per, key=lambda k: per[k]["ms"])
    d = per[dom]
    # the class is MFMA-bound when its steps' flop times outweigh their byte
    # times at the peaks, else HBM-bound
    if d["flop_s"] >= d["byte_s"]:
        achieved = d["flops"] / (d["ms"] * 1e-3)
        roof = {"bound": "mfma", "achieved": achieved / 1e12, "peak": PEAK["mfma_fp32"] / 1e12,
                "unit": "TFLOP/s", "frac": achieved / PEAK["mfma_fp32"]}
    else:
        achieved = d["bytes"] / (d["ms"] * 1e-3)
        roof = {"bound": "hbm", "achieved": achieved / 1e9, "peak": PEAK["hbm"] / 1e9,
                "unit": "GB/s", "frac": achieved / PEAK["hbm"]}
    roof["traffic"], roof["traffic_detail"] = measured_traffic(_lib.KC_NAMES[dom], d)
    roof["kernel_class"] = _lib.KC_NAMES[dom]
    roof["launches_per_run"] = d["steps"]
    roof["class_ms_per_run"] = d["ms"]
    roof["classes"] = {_lib.KC_NAMES[k]: round(v["ms"], 3) for k, v in per.items()}
    roof["class_bound_ms"] = {"flops": d["flop_s"] * 1e3, "bytes": d["byte_s"] * 1e3}
    # Whole-plan roofline: each step's lower bound max(flops / MFMA peak of
    # its element type, bytes / HBM peak), summed, against the measured run.
    lb = 0.0
    for i, (kind, sa, sb, sc, M, N, K) in enumerate(p.steps):
        inf = p.step_info(i)
        peak = PEAK["mfma_fp64"] if p.dtypes[sc] == _lib.PQ_C128 else PEAK["mfma_fp32"]
        lb += max(inf["flops"] / peak, inf["bytes"] / PEAK["hbm"])
    plan_roof = {"roofline_ms": lb * 1e3, "events_ms": float(sum(ms)), "wall_ms": 1e3 * dt / args.steps,
                 "frac_of_wall": lb * 1e3 / (1e3 * dt / args.steps)}
    p.close()

    # What the learned planner buys against FLOP counting alone (the
    # reference's ranking), on the selection's own measurements (same
    # compilation, default kernel choices): the model's top prediction alone
    # and the measured pick among the model's shortlist.
    planner_info = {"cost_model": mstats, "selection": sel["summary"], "learned_ms": 1e3 * dt / args.steps}
    if sel.get("chosen_ms") is not None:
        planner_info.update(
            flop_pick_untuned_ms=sel["flop_ms"], model_pick_untuned_ms=sel["model_ms"],
            chosen_untuned_ms=sel["chosen_ms"],
            learned_gain=sel["flop_ms"] / sel["chosen_ms"],
            model_only_gain=sel["flop_ms"] / sel["model_ms"])
    log(f"[bench] planner: {planner_info}")

    # FP64 run of the same plan: the accuracy reference for the complex64 path.
    rel_err = rel_err_norm = t_fp64_batched = fidelity = None
    if not args.no_fp64_check:
        p64 = compile_plan(tn, plan, sliced, group_bits=N_SLICED)
        p64.execute(True)
        ref = p64.result().reshape(-1)
        ctx.synchronize()
        t2 = time.perf_counter()
        p64.execute(True)
        ctx.synchronize()
        t_fp64_batched = time.perf_counter() - t2
        p64.close()
        rel_err = float(np.max(np.abs(res64 - ref) / np.abs(ref)))
        rel_err_norm = float(np.linalg.norm(res64 - ref) / np.linalg.norm(ref))
        fidelity = fp64_fidelity(res64, ref)
        log(f"[bench] max rel err vs FP64 {rel_err:.2e} (norm-wise {rel_err_norm:.2e}, fidelity "
            f"1 - {1 - fidelity:.2e}); batched FP64 run {t_fp64_batched:.3f}s")

    # Slicing that binds, and no slicing at all, on the same plan.
    slicing = None
    if not args.no_slicing_variants:
        slicing = slicing_variants(tn, plan, ref if not args.no_fp64_check else None)
        log(f"[bench] slicing variants: {slicing}")

    # Independent bitstrings: one contraction per bitstring, all in one run.
    indep = None
    if not args.no_bitstrings:
        indep = independent_bitstrings()
        log(f"[bench] independent bitstrings: {indep}")

    # The same plan with the complex64 GEMMs on FP16 hi/lo matrix cores.
    split16 = None
    if not args.no_split16 and not args.no_fp64_check:
        ms16, res16, cls16 = run_variant(tn, plan, sliced, "mixed16", N_SLICED)
        e16 = float(np.max(np.abs(res16 - ref) / np.abs(ref)))
        split16 = {"ms_per_run": round(ms16, 3), "amps_per_s": 64e3 / ms16,
                   "rel_err_vs_fp64": e16, "within_bound": e16 <= REL_ERR_BOUND,
                   "rel_err_vs_fp64_normwise": float(np.linalg.norm(res16 - ref) / np.linalg.norm(ref)),
                   "classes_ms": cls16}
        log(f"[bench] FP16 hi/lo split variant: {split16}")
    grid7 = None
    if not args.no_grid7:
        grid7 = grid7x7_config()
        log(f"[bench] 7x7 d12: {grid7}")
    mps = None
    if not args.no_mps:
        mps = mps_chain_config()
        log(f"[bench] MPS chain: {mps}")
    plain = None
    if not args.no_plain:
        t_slice, res_slice = plain_fp64_slice_time(plan, sliced, open_q)
        res_slice = np.asarray(res_slice).reshape(-1, order="F")
        plain = {"amps_per_s": 64 / (t_slice * 2 ** N_SLICED), "t_one_slice_s": t_slice,
                 "sample": "slice 0 of 1024 through contract_network! on HIPBackend{ComplexF64}, x1024"}
        log(f"[bench] plain FP64 path: one slice {t_slice:.2f}s -> {plain['amps_per_s']:.4f} amps/s")

    cpu = None
    if not args.no_cpu:
        cb = cpu_baseline(plan, sliced)
        cpu = {"value": 64 / (cb["t_slice_s"] * 2 ** N_SLICED), "unit": "amplitudes/s", "cores": CPU_THREADS,
               "kind": "port",
               "sample": (f"numpy oracle (permute + OpenBLAS zgemm, complex128) on slice 0: first "
                          f"{cb['steps_done']}/{cb['steps']} steps = {100 * cb['mac_fraction']:.1f}% of the "
                          f"slice's MACs in {cb['elapsed_s']:.1f}s, scaled to the slice and x1024 slices")}
        log(f"[bench] cpu baseline {cpu['value']:.3g} amps/s")
        if plain is not None and cb["result"] is not None:
            # the oracle finished slice 0: the plain FP64 GPU path's slice 0 against it
            # worst relative error over the nonzero amplitudes (a slice can
            # close off amplitudes exactly: those must stay zero)
            ref0 = cb["result"]
            nz = ref0 != 0
            plain["slice0_rel_err_vs_oracle"] = float(np.max(np.abs(res_slice[nz] - ref0[nz]) / np.abs(ref0[nz])))
            plain["slice0_zero_amplitudes"] = int((~nz).sum())
            plain["slice0_zeros_max_abs"] = float(np.max(np.abs(res_slice[~nz]))) if (~nz).any() else 0.0
            log(f"[bench] plain FP64 slice 0 vs oracle: {plain['slice0_rel_err_vs_oracle']:.2e} "
                f"({plain['slice0_zero_amplitudes']} exact zeros, largest there {plain['slice0_zeros_max_abs']:.1e})")

    ok = rel_err is None or rel_err <= REL_ERR_BOUND
    out = {
        "metric": METRIC, "value": amps, "unit": "amplitudes/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": 1e3 * dt / args.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None,
        "dtype": "c64",
        "precision": "complex64 on FP32 MFMA for the large steps; FP64 leaves, small steps and slice sum",
        "data": "synthetic",
        "config": {"workload": "6x6 grid depth-14 random circuit (seed 2026), 10 sliced bonds = 1024 slices, "
                               "64 bitstrings (last column open), one hipGraph per run",
                   "qubits": 36, "depth": DEPTH, "slices": 2 ** N_SLICED, "bitstrings": 64,
                   "parallelism": "single GPU"},
        "roofline": roof,
        "plan_roofline": plan_roof,
        "cpu_baseline": cpu,
        "ok": ok,
        "rel_err_bound": REL_ERR_BOUND,
        "rel_err_vs_fp64": rel_err,
        "rel_err_vs_fp64_normwise": rel_err_norm,
        "fidelity_vs_fp64": fidelity,
        "plain_fp64_gpu": plain,
        "speedup_vs_plain": (amps / plain["amps_per_s"]) if plain else None,
        "batched_fp64_amps_per_s": 64 / t_fp64_batched if t_fp64_batched else None,
        "planner": planner_info,
        "fp16_split_variant": split16,
        "slicing_variants": slicing,
        "independent_bitstrings": indep,
        "grid7x7_d12": grid7,
        "mps_chain20": mps,
    }
    print(json.dumps(out))
    if not ok:
        log(f"[bench] FAILED: worst amplitude error {rel_err:.3g} exceeds the {REL_ERR_BOUND:g} bound")
        sys.exit(1)


if __name__ == "__main__":
    main()
