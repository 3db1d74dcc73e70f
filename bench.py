#!/usr/bin/env python3
"""bench.py -- throughput of the QuIDS rule-application step on one MI355X.

Workload (default): plain QCGD, 16-node cycle graph with seeded particles
(srand(42), rand() & 1 in left[0], right[0], left[1], ... order), one
iteration = step -> split_merge(theta=pi/4, phi=0) -> erase_create(theta=pi/3,
phi=pi/5), max_num_object = 2^26, FP64 magnitudes, tolerance 1e-30.

A bench step is one simulate() rule application (the step modifier that
precedes a split_merge is timed with it): steps 1..4 are iterations 1-2,
which grow the state from one graph to 2^26 objects; step 5 is iteration
3's split_merge.  The first `--warmup` steps run untimed, the next `--steps`
are timed; the value is candidates (sy_it.num_object summed over the timed
simulate() calls) per second of timed wall time, with the state resident in
HBM.

Prints one JSON line on stdout; progress goes to stderr.
"""
import argparse
import json
import math
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12  # MI355X HBM3E, bytes/s
METRIC = "children/s per simulate() step; batched+adjoint vs plain GPU at 1e-10 rel err"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def iteration_rules():
    from quids_amd.rules import qcgd
    return [qcgd.step, qcgd.split_merge(math.pi / 4, 0.0), qcgd.erase_create(math.pi / 3, math.pi / 5)]


def alg_bytes(st):
    """Algorithmic HBM bytes of one step: parents read once, per candidate a
    24-byte {hash, mag} write + read + table update, per unique 32 bytes, per
    survivor its bytes + 28 (magnitude, offset, size)."""
    P, C, U, K = st["num_parent"], st["num_candidates"], st["num_unique"], st["num_survivors"]
    gen = st["in_bytes"] + 24 * P + 24 * C
    dedup = 48 * C + 32 * U
    fin = st["output_bytes"] + 28 * K
    return {"generate": gen, "dedup": dedup, "finalize": fin, "total": gen + dedup + fin}


def bench_ops():
    """(modifier or None, rule) of the two bench steps of an iteration."""
    rules = iteration_rules()
    return [(rules[0], rules[1]), (None, rules[2])]


def step_name(k):
    """Bench step k (1-based) as 'iteration i split_merge / erase_create'."""
    return f"iteration {(k + 1) // 2} {'split_merge' if k % 2 else 'erase_create'}"


def host_threads():
    """Host threads the CPU baseline uses: the box's CPU share (OMP_NUM_THREADS
    when the environment sets it), else the CPUs this process may run on."""
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        return int(env)
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def cpu_baseline(nodes, seed, step, log2_cap, budget_s):
    """The compiled reference (oracle/_ref/ref_driver) on the same workload up
    to bench step `step` (the first timed one), on this host's cores; the
    value is that step's candidates over its wall time.  The run uses cap
    2^log2_cap: at the GPU's cap the reference needs about 60 bytes of host
    memory per candidate (sy_it_t, src/quids.hpp:44), some 10^12 bytes."""
    from oracle import oracle as orc
    from quids_amd.state import HostState
    from quids_amd.rules import qcgd
    if not os.path.exists(orc.REF_DRIVER):
        return None
    threads = host_threads()
    hs = HostState.from_objects([qcgd.seeded_graph(nodes, seed)])
    cap = 1 << log2_cap
    ops, op_of_step = [], {}
    for k in range(1, step + 1):
        mod, rule = bench_ops()[(k - 1) % 2]
        if mod is not None:
            ops.append(orc.mod_op(mod))
        op_of_step[len(ops)] = k
        ops.append(orc.rule_op(rule, cap, dump=False))
    try:
        _, out = orc.run_reference(1, hs, ops, threads=threads, timeout=budget_s)
    except Exception as e:  # noqa: BLE001
        log(f"cpu baseline failed: {e}")
        return None
    for line in out.splitlines():
        w = line.split()
        f = dict(kv.split("=") for kv in w[2:])
        if op_of_step.get(int(w[1])) == step:
            C, secs = float(f["C"]), float(f["seconds"])
            return {"value": C / secs if secs else None, "unit": "children/s", "cores": threads, "kind": "reference",
                    "sample": f"reference simulate() on bench step {step} ({step_name(step)}) of the same workload "
                              f"at cap 2^{log2_cap} (steps 1..{step} run at that cap), {int(C)} candidates in "
                              f"{secs:.2f} s, OMP_NUM_THREADS={threads} (os.cpu_count() = {os.cpu_count()})"}
    return None


# Kernels of the three main phases (name prefixes in rocprofv3 traces): the
# generate phase is k_generate_scatter (generate fused with the first radix
# pass), k_generate_fast (split_generate=1) or, for steps beyond the pass
# budget, k_range_count (the counting pass) and one k_range_walkwrite per
# hash-range pass (k_range_bincount + k_range_binwrite when a pass takes two
# walks).
KERNELS_OF_PHASE = {"generate": ("k_generate_scatter", "k_generate_fast", "k_range_count", "k_range_walkwrite",
                                 "k_range_bincount", "k_range_binwrite"),
                    "dedup": ("k_dedup<", "k_dedup_sort", "k_dedup_big"),
                    "finalize": ("k_populate",)}
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic_qcgd16.json")
MARKER_BYTES = 4096  # the plain copy bench.py runs just before its timed region (a marker in traces)


def dominant_kernel_name(phase, ctx, stats):
    if phase == "generate":
        names = []
        if any(not (s["reserved"] & 4) for s in stats):
            names.append("k_generate_fast" if ctx.get_tuning()["split_generate"] else "k_generate_scatter")
        if any(s["reserved"] & 4 for s in stats):
            names += ["k_range_count", "k_range_walkwrite"]
            if any((s["reserved"] & 12) == 12 for s in stats):
                names += ["k_range_bincount", "k_range_binwrite"]
        return "+".join(sorted(names))
    return {"dedup": "k_dedup", "finalize": "k_populate"}[phase]


def traffic_per_launch(phase, kname, candidates, launches, key):
    """HBM bytes per launch of the phase's kernels from the rocprofv3 counter
    runs of this workload (profiles/traffic_qcgd16.json: FETCH_SIZE +
    WRITE_SIZE over the timed dispatches, scaled by the plain-copy
    calibration, per candidate of the timed steps), or None."""
    try:
        with open(TRAFFIC_FILE) as f:
            tr = json.load(f)
    except (OSError, ValueError):
        return None
    kt = tr.get("kernels", {}).get(phase, {})
    if tr.get("workload") != key or kt.get("kernel") != kname:
        return None
    return kt["hbm_bytes_per_candidate"] * candidates / launches


def _csv(path_glob):
    import csv
    import glob
    paths = sorted(glob.glob(path_glob, recursive=True))
    if not paths:
        raise SystemExit(f"no file matches {path_glob}")
    with open(paths[0]) as f:
        return list(csv.DictReader(f))


def _dispatches(rows, prefixes, counter=None):
    """{dispatch id: (kernel name, value)}: the counter summed over the rows of
    each dispatch of a kernel whose name starts with one of the prefixes
    (after 'void qd::'), or the trace row's duration in ns."""
    out = {}
    for r in rows:
        name = r["Kernel_Name"].replace("void ", "").replace("qd::", "")
        if not any(name.startswith(p) for p in prefixes):
            continue
        d = int(r["Dispatch_Id"])
        if counter is None:
            out[d] = (name, float(int(r["End_Timestamp"]) - int(r["Start_Timestamp"])))
        elif r["Counter_Name"] == counter:
            out[d] = (name, out.get(d, (name, 0.0))[1] + float(r["Counter_Value"]))
    return out


def _after_marker(rows, values):
    """The dispatches after the last k_copy_probe (bench.py's marker)."""
    marks = [int(r["Dispatch_Id"]) for r in rows if "k_copy_probe" in r["Kernel_Name"]]
    m = max(marks) if marks else -1
    return {d: v for d, v in values.items() if d > m}


def counter_summary(folder):
    """Per kernel family (name without namespace or template arguments),
    the per-dispatch average of every counter a rocprofv3 --pmc run of this
    script collected, over the timed dispatches (after the marker copy).
    Prints a markdown table, kernels by SQ_INSTS_VALU."""
    rows = _csv(os.path.join(folder, "**", "*counter_collection.csv"))
    after = _after_marker(rows, {int(r["Dispatch_Id"]): r for r in rows})
    sums, disp, names = {}, {}, []
    for r in rows:
        if int(r["Dispatch_Id"]) not in after:
            continue
        fam = r["Kernel_Name"].replace("void ", "").replace("qd::", "").split("(")[0].split("<")[0]
        c = r["Counter_Name"]
        if c not in names:
            names.append(c)
        sums.setdefault(fam, {}).setdefault(c, 0.0)
        sums[fam][c] += float(r["Counter_Value"])
        disp.setdefault(fam, set()).add(int(r["Dispatch_Id"]))
    print("| kernel | dispatches | " + " | ".join(names) + " |")
    print("|---|---:|" + "---:|" * len(names))
    for fam, c in sorted(sums.items(), key=lambda x: -x[1].get("SQ_INSTS_VALU", 0)):
        n = len(disp[fam])
        print(f"| `{fam}` | {n} | " + " | ".join(f"{c.get(k, 0) / n:.4g}" for k in names) + " |")


def rocprof_summary(args):
    """Formats the rocprofv3 output of this script's own runs of one command:
    kernel-trace stats, then per phase kernel family its timed dispatches
    (those after the marker copy that precedes the timed region) with their
    average duration, and FETCH_SIZE / WRITE_SIZE per dispatch scaled by
    what the same counters read for the library's plain copy
    (qd_copy_probe) of a known size.  The traffic is also written per
    candidate of the timed steps (--bench-json: the JSON line of the run) to
    --traffic-out, where bench.py finds it.  Prints markdown."""
    kt, fe, wr, cf, cw = args.rocprof_summary
    stats = _csv(os.path.join(kt, "**", "*kernel_stats.csv"))
    trace = _csv(os.path.join(kt, "**", "*kernel_trace.csv"))
    with open(args.bench_json) as f:
        line = json.loads([x for x in f.read().splitlines() if x.startswith("{")][-1])
    C = line["config"]["candidates"]
    lines = ["| kernel | calls | total ms | avg us | % |", "|---|---:|---:|---:|---:|"]
    for r in sorted(stats, key=lambda r: -float(r["TotalDurationNs"]))[:20]:
        name = r["Name"].split("(")[0].replace("void ", "")
        lines.append(f"| `{name}` | {r['Calls']} | {float(r['TotalDurationNs']) / 1e6:.2f} | "
                     f"{float(r['AverageNs']) / 1e3:.1f} | {float(r['Percentage']):.1f} |")
    with open(args.calib_json) as f:
        calib_bytes = json.loads([x for x in f.read().splitlines() if x.startswith("{")][-1])["copy_probe_bytes"]
    crow_f = _csv(os.path.join(cf, "**", "*counter_collection.csv"))
    crow_w = _csv(os.path.join(cw, "**", "*counter_collection.csv"))
    cfetch = sorted(_dispatches(crow_f, ("k_copy_probe",), "FETCH_SIZE").items())
    cwrite = sorted(_dispatches(crow_w, ("k_copy_probe",), "WRITE_SIZE").items())
    # the copy reads and writes calib_bytes each; counters are in KiB
    f_scale = calib_bytes / (1024.0 * cfetch[-1][1][1])
    w_scale = calib_bytes / (1024.0 * cwrite[-1][1][1])
    frows = _csv(os.path.join(fe, "**", "*counter_collection.csv"))
    wrows = _csv(os.path.join(wr, "**", "*counter_collection.csv"))
    traffic = {"workload": args.workload_key, "candidates": C, "fetch_scale": f_scale, "write_scale": w_scale,
               "kernels": {}}
    lines += ["", f"Timed dispatches (after the marker copy; {C} candidates in the timed steps):", "",
              "| phase | kernels | dispatches | avg us (trace) | FETCH_SIZE KiB/dispatch | WRITE_SIZE KiB/dispatch | "
              "HBM bytes/dispatch (scaled) | HBM bytes/candidate |", "|---|---|---:|---:|---:|---:|---:|---:|"]
    for phase, prefixes in KERNELS_OF_PHASE.items():
        durs = _after_marker(trace, _dispatches(trace, prefixes))
        fv = _after_marker(frows, _dispatches(frows, prefixes, "FETCH_SIZE"))
        wv = _after_marker(wrows, _dispatches(wrows, prefixes, "WRITE_SIZE"))
        if not durs or not fv or not wv:
            continue
        fams = sorted({n.split("(")[0].split("<")[0] for n, _ in durs.values()})
        n = len(durs)
        avg_us = sum(v for _, v in durs.values()) / n / 1e3
        fk = sum(v for _, v in fv.values()) / len(fv)
        wk = sum(v for _, v in wv.values()) / len(wv)
        hbm = 1024.0 * (fk * f_scale + wk * w_scale)
        per_cand = hbm * len(fv) / C
        kname = "+".join(f for f in fams if f != "k_dedup_sort" and f != "k_dedup_big")
        traffic["kernels"][phase] = {"kernel": kname, "dispatches": n, "avg_us": avg_us, "fetch_kib": fk,
                                     "write_kib": wk, "hbm_bytes_per_dispatch": hbm,
                                     "hbm_bytes_per_candidate": per_cand}
        lines.append(f"| {phase} | `{kname}` | {n} | {avg_us:.1f} | {fk:.0f} | {wk:.0f} | {hbm:.4g} | "
                     f"{per_cand:.1f} |")
    lines += ["", f"Counter scale (copy of {calib_bytes} B read + written): FETCH_SIZE x{f_scale:.3f}, "
                  f"WRITE_SIZE x{w_scale:.3f}"]
    with open(args.traffic_out, "w") as f:
        json.dump(traffic, f, indent=1)
    print("\n".join(lines))


def adjoint_bench(nodes=10, seed=42, n_iter=6, log2_cap=18, checkpoint_every=2, gd_steps=0, lr=0.05, reps=3):
    """QCGD adjoint: loss = expected particle count after n_iter iterations
    of step -> split_merge(t1, p1) -> erase_create(t2, p2) from a seeded
    graph, gradient w.r.t. (t1, p1, t2, p2) with a checkpoint every
    `checkpoint_every` iterations.  Untruncated runs of 10 nodes do not fit
    past iteration 2, so every step keeps the top 2^log2_cap objects (the
    adjoint holds that mask fixed).  Times the forward pass alone and the
    gradient (forward with checkpoints + recompute + adjoint steps)."""
    import numpy as np
    import torch
    import quids_amd as q
    from quids_amd import adjoint as adj
    from quids_amd.rules import qcgd
    from quids_amd.state import HostState
    it = q.it_t()
    it.upload(HostState.from_objects([qcgd.seeded_graph(nodes, seed)]))
    prog = adj.qcgd_program(n_iter, 1 << log2_cap)
    p = np.array([math.pi / 4, 0.0, math.pi / 3, math.pi / 5])
    adj.gradient(prog, p, it, adj.OBS_PARTICLES, 0, checkpoint_every)  # warm-up (scratch allocation)
    fwd = []
    for _ in range(reps):
        torch.cuda.synchronize()
        t = time.perf_counter()
        adj.loss(prog, p, it, adj.OBS_PARTICLES)
        torch.cuda.synchronize()
        fwd.append(time.perf_counter() - t)
    grad_t, bwd_t = [], []
    for _ in range(reps):
        tm = {}
        value, g = adj.gradient(prog, p, it, adj.OBS_PARTICLES, 0, checkpoint_every, timings=tm)
        grad_t.append(tm["forward"] + tm["backward"])
        bwd_t.append(tm["backward"])
    f = min(fwd)
    out = {"workload": f"QCGD adjoint, {nodes} nodes seed {seed}, {n_iter} iterations, cap 2^{log2_cap}, "
                       f"checkpoint every {checkpoint_every}, loss = expected particle count",
           "forward_ms": f * 1e3, "gradient_ms": min(grad_t) * 1e3, "backward_ms": min(bwd_t) * 1e3,
           "gradient_over_forward": min(grad_t) / f, "backward_over_forward": min(bwd_t) / f,
           "loss": value, "grad": [float(x) for x in g]}
    if gd_steps:
        losses = []
        t = time.perf_counter()
        for _ in range(gd_steps):
            value, g = adj.gradient(prog, p, it, adj.OBS_PARTICLES, 0, checkpoint_every)
            losses.append(value)
            p = p - lr * g
        out.update({"gd_steps": gd_steps, "lr": lr, "gd_seconds": time.perf_counter() - t,
                    "loss_first": losses[0], "loss_last": losses[-1], "params_last": [float(x) for x in p]})
    it.close()
    return out


def _dev_free_gb():
    import torch
    return torch.cuda.mem_get_info()[0] / 1e9


def batched_bench(n_inst=256, n_iter=8, log2_cap=18, seed=4242, lo_nodes=6, hi_nodes=16, sequential=True):
    """QCGD batch: n_inst instances with node counts uniform in
    [lo_nodes, hi_nodes], particle bits from per-instance seeds and (theta,
    phi) uniform in [0, pi/2] x [0, pi] per instance (the same angles drive
    split_merge and erase_create of an instance), n_iter iterations of
    step -> split_merge -> erase_create with a per-instance cap 2^log2_cap.
    Times the batched steps and, when `sequential`, the same work as n_inst
    separate plain simulate() calls per rule; checks that every instance's
    counters and total probability agree."""
    import numpy as np
    import quids_amd as q
    from quids_amd.batch import apply_modifier, batch_t, simulate_batched
    from quids_amd.rules import qcgd
    from quids_amd.state import HostState
    rng = np.random.default_rng(seed)
    nodes = rng.integers(lo_nodes, hi_nodes + 1, n_inst)
    seeds = rng.integers(0, 2 ** 31 - 1, n_inst)
    theta = rng.uniform(0, math.pi / 2, n_inst)
    phi = rng.uniform(0, math.pi, n_inst)
    states = [HostState.from_objects([qcgd.seeded_graph(int(nodes[i]), int(seeds[i]))]) for i in range(n_inst)]
    cap = 1 << log2_cap
    ctx = q.default_context()

    def rules(which):
        cls = qcgd.split_merge if which == 0 else qcgd.erase_create
        return [cls(float(theta[i]), float(phi[i])) for i in range(n_inst)]

    # batched: one untimed pass first, which sizes the context's scratch
    # (allocations of tens of GB take up to seconds; the sequential calls
    # below find the scratch allocated too)
    def run_batched(label, a, b):
        stats = []
        for it in range(n_iter):
            apply_modifier(a, qcgd.step)
            for which in (0, 1):
                stats.append(simulate_batched(a, rules(which), b, cap))
                a, b = b, a
                log(f"{label} batched iteration {it + 1} rule {which}: "
                    f"C={sum(x['num_candidates'] for x in stats[-1])} "
                    f"objects={int(a.seg[-1])} chunked={bool(stats[-1][0]['reserved'] & 2)} "
                    f"scratch={ctx.scratch_bytes() / 1e9:.1f} GB free={_dev_free_gb():.1f} GB")
        return stats, a, b

    for x in run_batched("untimed", batch_t.from_states(states), batch_t())[1:]:
        x.close()
    a, b = batch_t.from_states(states), batch_t()
    ctx.set_profiling(True)
    ctx.reset_phase_times()
    ctx.synchronize()
    t0 = time.perf_counter()
    bstats, a, b = run_batched("timed", a, b)
    ctx.synchronize()
    t_batch = time.perf_counter() - t0
    a.close()
    b.close()
    bphase = {k: round(v[0], 2) for k, v in ctx.phase_times().items()}
    ctx.reset_phase_times()
    C = sum(st["num_candidates"] for step in bstats for st in step)
    out = {"workload": f"batched QCGD, {n_inst} instances, nodes uniform in {lo_nodes}..{hi_nodes}, "
                       f"theta in [0, pi/2], phi in [0, pi] per instance (same angles for both rules), "
                       f"{n_iter} iterations, per-instance cap 2^{log2_cap}, seed {seed}",
           "candidates": C, "batched_seconds": t_batch, "batched_children_per_s": C / t_batch,
           "batched_phase_ms": bphase}
    if sequential:
        its = []
        for hs in states:
            it = q.it_t()
            it.upload(hs)
            its.append(it)
        sstats = []
        ctx.synchronize()
        t0 = time.perf_counter()
        for it in range(n_iter):
            for i in range(n_inst):
                q.simulate(its[i], qcgd.step)
            for which in (0, 1):
                rs = rules(which)
                row = []
                for i in range(n_inst):
                    nxt = q.it_t()
                    row.append(q.simulate(its[i], rs[i], nxt, None, cap))
                    its[i].close()
                    its[i] = nxt
                sstats.append(row)
            log(f"sequential iteration {it + 1}")
        ctx.synchronize()
        t_seq = time.perf_counter() - t0
        out["sequential_phase_ms"] = {k: round(v[0], 2) for k, v in ctx.phase_times().items()}
        mismatch = 0
        for bs, ss in zip(bstats, sstats):
            for x, y in zip(bs, ss):
                if any(x[k] != y[k] for k in ("num_candidates", "num_unique", "num_survivors", "total_proba")):
                    mismatch += 1
        out.update({"sequential_seconds": t_seq, "sequential_children_per_s": C / t_seq,
                    "speedup_vs_sequential": t_seq / t_batch, "instance_steps_mismatched": mismatch})
        for it in its:
            it.close()
    ctx.set_profiling(False)
    return out


def workload_key(args):
    return f"qcgd nodes={args.nodes} seed={args.seed} cap=2^{args.log2_cap}"


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=6)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--nodes", type=int, default=16)
    ap.add_argument("--seed", type=int, default=42)
    ap.add_argument("--log2-cap", type=int, default=26)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=120.0)
    ap.add_argument("--cpu-log2-cap", type=int, default=18,
                    help="cap of the CPU baseline's run (the reference needs about 60 bytes of host memory per "
                         "candidate, so the GPU's cap does not fit a host)")
    ap.add_argument("--copy-probe", type=float, default=0,
                    help="only run the library's plain copy of this many GiB (counter calibration) and exit")
    ap.add_argument("--rocprof-summary", nargs=5, metavar=("KT", "FETCH", "WRITE", "CAL_FETCH", "CAL_WRITE"),
                    help="summarize rocprofv3 output folders of this script's runs and exit")
    ap.add_argument("--counter-summary", metavar="PMC",
                    help="per-dispatch averages per kernel of a rocprofv3 --pmc output folder of this script's run, "
                         "then exit")
    ap.add_argument("--calib-json", help="the --copy-probe JSON line of the calibration run")
    ap.add_argument("--bench-json", help="the JSON line of the profiled bench run (--rocprof-summary)")
    ap.add_argument("--traffic-out", default=TRAFFIC_FILE)
    ap.add_argument("--workload", choices=["plain", "adjoint", "batched"], default="plain",
                    help="adjoint: the QCGD adjoint workload (with --gd-steps of gradient descent); "
                         "batched: the 256-instance batch against sequential plain calls")
    ap.add_argument("--batch-log2-cap", type=int, default=12)
    ap.add_argument("--batch-iters", type=int, default=5)
    ap.add_argument("--batch-instances", type=int, default=256)
    ap.add_argument("--no-sequential", action="store_true", help="batched workload without the sequential run")
    ap.add_argument("--gd-steps", type=int, default=50)
    ap.add_argument("--no-extras", action="store_true", help="skip the adjoint and batched summaries of the plain line")
    ap.add_argument("--autotune", action="store_true",
                    help="before the timed iterations, let quids_amd.tuning pick the step's launch and bucket "
                         "parameters on the warmed-up state")
    ap.add_argument("--tune", action="append", default=[], metavar="KEY=VALUE",
                    help="set one qd_tuning field before the run (repeatable), e.g. --tune split_generate=1")
    args = ap.parse_args()
    if args.counter_summary:
        counter_summary(args.counter_summary)
        return
    if args.rocprof_summary:
        args.workload_key = workload_key(args)
        rocprof_summary(args)
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}); the library has no multi-GPU path",
              file=sys.stderr)
        sys.exit(2)

    import quids_amd as q
    from quids_amd.rules import qcgd
    from quids_amd.state import HostState

    if args.tune:
        q.default_context().set_tuning(**{k: int(v) for k, v in (kv.split("=", 1) for kv in args.tune)})

    if args.workload == "adjoint":
        res = adjoint_bench(gd_steps=args.gd_steps)
        res.update({"metric": "adjoint gradient cost over one forward pass", "value": res["gradient_over_forward"],
                    "unit": "x", "higher_is_better": False, "n_gpus": 1, "dtype": "f64", "data": "synthetic"})
        print(json.dumps(res), flush=True)
        return

    if args.workload == "batched":
        res = batched_bench(n_inst=args.batch_instances, n_iter=args.batch_iters, log2_cap=args.batch_log2_cap,
                            sequential=not args.no_sequential)
        res.update({"metric": "batched children/s over sequential plain simulate() calls",
                    "value": res.get("speedup_vs_sequential"), "unit": "x", "higher_is_better": True, "n_gpus": 1,
                    "dtype": "f64", "data": "synthetic"})
        print(json.dumps(res), flush=True)
        return

    if args.copy_probe:
        nbytes = int(args.copy_probe * (1 << 30))
        g = q.default_context().copy_probe(nbytes, reps=3)
        print(json.dumps({"copy_probe_bytes": nbytes, "gbps_read_plus_write": g}), flush=True)
        return

    cap = 1 << args.log2_cap
    ctx = q.default_context()
    a, b = q.it_t(ctx=ctx), q.it_t(ctx=ctx)
    sy = q.sy_it_t(ctx=ctx)
    a.upload(HostState.from_objects([qcgd.seeded_graph(args.nodes, args.seed)]))
    ops = bench_ops()

    def run_step(k, timed_stats=None):
        """Bench step k (1-based): the modifier of its iteration if it opens
        one, then one simulate() call."""
        nonlocal a, b
        mod, rule = ops[(k - 1) % 2]
        if mod is not None:
            q.simulate(a, mod)
        in_bytes = a.info()[1]
        t = time.perf_counter()
        st = q.simulate(a, rule, b, sy, cap)
        st["in_bytes"] = in_bytes
        st["seconds"] = time.perf_counter() - t  # host wall time of the call (it synchronises)
        a, b = b, a
        if timed_stats is not None:
            timed_stats.append(st)
        return st

    for k in range(1, args.warmup + 1):
        st = run_step(k)
        log(f"warmup step {k} ({step_name(k)}): C={st['num_candidates']} U={st['num_unique']} "
            f"objects={a.num_object} ({st['seconds']:.2f} s)")

    tuned = None
    if args.autotune:
        from quids_amd import tuning
        tuned = tuning.autotune(ctx, a, [ops[args.warmup % 2][1]], cap, reps=1, verbose=False)
        log(f"autotune: {tuned}")
    ctx.copy_probe(MARKER_BYTES, reps=1)  # marks the start of the timed region in kernel traces
    ctx.set_profiling(True)
    ctx.reset_phase_times()
    stats = []
    ctx.synchronize()
    t0 = time.perf_counter()
    for k in range(args.warmup + 1, args.warmup + args.steps + 1):
        st = run_step(k, stats)
        log(f"timed step {k} ({step_name(k)}): C={st['num_candidates']} U={st['num_unique']} "
            f"objects={a.num_object} passes={'range' if st['reserved'] & 4 else 'single'} ({st['seconds']:.2f} s)")
    ctx.synchronize()
    t1 = time.perf_counter()
    elapsed = t1 - t0
    phases = ctx.phase_times()
    ctx.set_profiling(False)
    a.close()
    b.close()

    C = sum(s["num_candidates"] for s in stats)
    value = C / elapsed
    ab = {k: sum(alg_bytes(s)[k] for s in stats) for k in ("generate", "dedup", "finalize", "total")}
    phase_ms = {k: v[0] for k, v in phases.items()}
    phase_calls = {k: v[1] for k, v in phases.items()}
    kernel_ms = {k: phase_ms[k] for k in ("generate", "dedup")}  # one kernel per phase interval
    dom = max(kernel_ms, key=kernel_ms.get)
    # per launch: the phase's algorithmic bytes of the timed steps over the
    # launches of its kernels in them (one per simulate() call, or one per
    # hash-range pass plus the counting pass)
    launches = max(int(phase_calls[dom]), 1)
    dom_ms = kernel_ms[dom] / launches
    per_launch = ab[dom] / launches
    achieved = per_launch / (dom_ms / 1e3) if dom_ms > 0 else None
    kname = dominant_kernel_name(dom, ctx, stats)
    traffic = traffic_per_launch(dom, kname, C, launches, workload_key(args))
    roofline = {"bound": "hbm", "kernel": kname, "achieved": achieved / 1e9 if achieved else None,
                "peak": HBM_PEAK / 1e9, "unit": "GB/s", "frac": achieved / HBM_PEAK if achieved else None,
                "traffic": traffic, "alg_bytes_per_launch": per_launch, "ms_per_launch": dom_ms,
                "launches": launches}
    step_frac = ab["total"] / elapsed / HBM_PEAK

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(args.nodes, args.seed, args.warmup + 1, args.cpu_log2_cap, args.cpu_budget)
    extras = {}
    if not args.no_extras:
        for key, fn in (("adjoint", lambda: adjoint_bench(gd_steps=args.gd_steps)),
                        ("batched", lambda: batched_bench(n_iter=args.batch_iters, log2_cap=args.batch_log2_cap))):
            try:
                extras[key] = fn()
            except Exception as e:  # noqa: BLE001
                extras[key] = {"error": str(e)}

    first, last = args.warmup + 1, args.warmup + args.steps
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "children/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"plain QCGD, {args.nodes}-node cycle graph, seed {args.seed}, "
                               f"step+split_merge(pi/4,0)+erase_create(pi/3,pi/5), max_num_object=2^{args.log2_cap}, "
                               f"simulate() calls {first}..{last} timed ({step_name(first)} .. {step_name(last)})",
                   "candidates": C, "parallelism": "single GPU",
                   "range_pass_steps": sum(bool(s["reserved"] & 4) for s in stats)},
        "roofline": roofline,
        "step_hbm_frac": step_frac,
        "phase_ms": {k: round(v, 3) for k, v in phase_ms.items()},
        "cpu_baseline": cpu,
        "tuning": ctx.get_tuning() if tuned is None else dict(tuned["tuning"], autotuned=True),
        **extras,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
