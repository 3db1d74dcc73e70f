#!/usr/bin/env python3
"""bench.py -- throughput of the QuIDS rule-application step on one MI355X.

Workload: plain QCGD, 12-node cycle graph with seeded particles (srand(seed),
rand() & 1 in left[0], right[0], left[1], ... order), one iteration =
step -> split_merge(theta=pi/4, phi=0) -> erase_create(theta=pi/3, phi=pi/5),
max_num_object = 2^22, FP64 magnitudes, tolerance 1e-30.  The first
`--warmup` iterations grow the state untimed; the next `--steps` iterations
are timed.  A bench step is one iteration (two simulate() calls and the
step modifier); the value is candidates (sy_it.num_object summed over the
timed simulate() calls) per second of timed wall time.

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


def cpu_baseline(nodes, seed, cap, budget_s, iterations=2):
    """The compiled reference (oracle/_ref/ref_driver) on the same workload's
    first `iterations` iterations (the second one is the first large step:
    ~1.7e7 candidates truncated to 2^22), on this host's cores."""
    from oracle import oracle as orc
    from quids_amd.rules import qcgd
    from quids_amd.state import HostState
    if not os.path.exists(orc.REF_DRIVER):
        return None
    threads = min(16, os.cpu_count() or 1)
    hs = HostState.from_objects([qcgd.seeded_graph(nodes, seed)])
    rules = iteration_rules()
    ops = [orc.mod_op(rules[0]), orc.rule_op(rules[1], cap, dump=False), orc.rule_op(rules[2], cap, dump=False)]
    ops = ops * iterations
    try:
        _, out = orc.run_reference(1, hs, ops, threads=threads, timeout=budget_s)
    except Exception as e:  # noqa: BLE001
        log(f"cpu baseline failed: {e}")
        return None
    C = secs = 0.0
    for line in out.splitlines():
        f = dict(kv.split("=") for kv in line.split()[2:])
        C += float(f["C"])
        secs += float(f["seconds"])
    return {"value": C / secs if secs else None, "unit": "children/s", "cores": threads, "kind": "reference",
            "sample": f"reference simulate() on iterations 1..{iterations} of the same workload "
                      f"({nodes} nodes, cap 2^{cap.bit_length() - 1}), "
                      f"{int(C)} candidates in {secs:.2f} s, OMP_NUM_THREADS={threads}"}


# The generate phase's kernel: k_generate_scatter (generate fused with the
# first radix pass, the default) or k_generate_fast (split_generate=1).
KERNEL_OF_PHASE = {"generate": "k_generate_scatter", "dedup": "k_dedup<", "finalize": "k_populate"}
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic_qcgd12.json")


def _csv(path_glob):
    import csv
    import glob
    paths = sorted(glob.glob(path_glob, recursive=True))
    if not paths:
        raise SystemExit(f"no file matches {path_glob}")
    with open(paths[0]) as f:
        return list(csv.DictReader(f))


def _counter_per_dispatch(rows, kernel, counter):
    """{dispatch id: counter value (summed over the rows of the dispatch)}"""
    out = {}
    for r in rows:
        if kernel in r["Kernel_Name"] and r["Counter_Name"] == counter:
            d = int(r["Dispatch_Id"])
            out[d] = out.get(d, 0.0) + float(r["Counter_Value"])
    return [out[d] for d in sorted(out)]


def rocprof_summary(args):
    """Formats the rocprofv3 output of this script's own runs: kernel-trace
    stats, the timed launches of the dominant kernel (the last 2*steps
    launches: one per simulate() call), and FETCH_SIZE / WRITE_SIZE per
    launch scaled by what the same counters read for the library's plain
    copy (qd_copy_probe) of a known size.  Prints markdown and writes the
    traffic figures to --traffic-out."""
    kt, fe, wr, cf, cw = args.rocprof_summary
    stats = _csv(os.path.join(kt, "**", "*kernel_stats.csv"))
    trace = _csv(os.path.join(kt, "**", "*kernel_trace.csv"))
    nt = 2 * args.steps
    lines = ["| kernel | calls | total ms | avg us | % |", "|---|---:|---:|---:|---:|"]
    for r in sorted(stats, key=lambda r: -float(r["TotalDurationNs"]))[:16]:
        name = r["Name"].split("(")[0].replace("void ", "")
        lines.append(f"| `{name}` | {r['Calls']} | {float(r['TotalDurationNs']) / 1e6:.2f} | "
                     f"{float(r['AverageNs']) / 1e3:.1f} | {float(r['Percentage']):.1f} |")
    calib_bytes = None
    with open(args.calib_json) as f:
        calib_bytes = json.load(f)["copy_probe_bytes"]
    cfetch = _counter_per_dispatch(_csv(os.path.join(cf, "**", "*counter_collection.csv")), "k_copy_probe", "FETCH_SIZE")
    cwrite = _counter_per_dispatch(_csv(os.path.join(cw, "**", "*counter_collection.csv")), "k_copy_probe", "WRITE_SIZE")
    # the copy reads and writes calib_bytes each; counters are in KiB
    f_scale = calib_bytes / (1024.0 * cfetch[-1])
    w_scale = calib_bytes / (1024.0 * cwrite[-1])
    fetch = _csv(os.path.join(fe, "**", "*counter_collection.csv"))
    write = _csv(os.path.join(wr, "**", "*counter_collection.csv"))
    traffic = {"workload": args.workload_key, "timed_launches": nt, "fetch_scale": f_scale, "write_scale": w_scale,
               "kernels": {}}
    lines += ["", "Timed launches (the last 2 x steps of each kernel, one per simulate() call):", "",
              "| kernel | launches | avg us (trace) | FETCH_SIZE KiB/launch | WRITE_SIZE KiB/launch | "
              "HBM bytes/launch (scaled) |", "|---|---:|---:|---:|---:|---:|"]
    for phase, kname in KERNEL_OF_PHASE.items():
        durs = [(int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) for r in trace if kname in r["Kernel_Name"]]
        fv = _counter_per_dispatch(fetch, kname, "FETCH_SIZE")
        wv = _counter_per_dispatch(write, kname, "WRITE_SIZE")
        if len(durs) < nt or len(fv) < nt or len(wv) < nt:
            continue
        durs, fv, wv = durs[-nt:], fv[-nt:], wv[-nt:]
        avg_us = sum(durs) / nt / 1e3
        fk, wk = sum(fv) / nt, sum(wv) / nt
        hbm = 1024.0 * (fk * f_scale + wk * w_scale)
        traffic["kernels"][phase] = {"kernel": kname.rstrip("<"), "avg_us": avg_us, "fetch_kib": fk,
                                     "write_kib": wk, "hbm_bytes_per_launch": hbm}
        lines.append(f"| `{kname.rstrip('<')}` | {nt} | {avg_us:.1f} | {fk:.0f} | {wk:.0f} | {hbm:.4g} |")
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
    return f"qcgd nodes={args.nodes} seed={args.seed} cap=2^{args.log2_cap} warmup={args.warmup} steps={args.steps}"


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=4)
    ap.add_argument("--warmup", type=int, default=6)
    ap.add_argument("--nodes", type=int, default=12)
    ap.add_argument("--seed", type=int, default=42)
    ap.add_argument("--log2-cap", type=int, default=22)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=120.0)
    ap.add_argument("--copy-probe", type=float, default=0,
                    help="only run the library's plain copy of this many GiB (counter calibration) and exit")
    ap.add_argument("--rocprof-summary", nargs=5, metavar=("KT", "FETCH", "WRITE", "CAL_FETCH", "CAL_WRITE"),
                    help="summarize rocprofv3 output folders of this script's runs and exit")
    ap.add_argument("--calib-json", help="the --copy-probe JSON line of the calibration run")
    ap.add_argument("--traffic-out", default=TRAFFIC_FILE)
    ap.add_argument("--workload", choices=["plain", "adjoint", "batched"], default="plain",
                    help="adjoint: the QCGD adjoint workload (with --gd-steps of gradient descent); "
                         "batched: the 256-instance batch against sequential plain calls")
    ap.add_argument("--batch-log2-cap", type=int, default=14)
    ap.add_argument("--batch-iters", type=int, default=4)
    ap.add_argument("--batch-instances", type=int, default=256)
    ap.add_argument("--no-sequential", action="store_true", help="batched workload without the sequential run")
    ap.add_argument("--gd-steps", type=int, default=50)
    ap.add_argument("--no-extras", action="store_true", help="skip the adjoint summary of the plain line")
    ap.add_argument("--autotune", action="store_true",
                    help="before the timed iterations, let quids_amd.tuning pick the step's launch and bucket "
                         "parameters on the warmed-up state")
    ap.add_argument("--tune", action="append", default=[], metavar="KEY=VALUE",
                    help="set one qd_tuning field before the run (repeatable), e.g. --tune split_generate=1")
    args = ap.parse_args()
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
    rules = iteration_rules()
    ctx = q.default_context()
    a, b = q.it_t(ctx=ctx), q.it_t(ctx=ctx)
    sy = q.sy_it_t(ctx=ctx)
    a.upload(HostState.from_objects([qcgd.seeded_graph(args.nodes, args.seed)]))

    def run_iteration(timed_stats=None):
        nonlocal a, b
        q.simulate(a, rules[0])
        for r in rules[1:]:
            in_bytes = a.info()[1] if timed_stats is not None else 0
            st = q.simulate(a, r, b, sy, cap)
            if timed_stats is not None:
                st["in_bytes"] = in_bytes
                timed_stats.append(st)
            a, b = b, a

    for i in range(args.warmup):
        t = time.time()
        run_iteration()
        ctx.synchronize()
        log(f"warmup iteration {i + 1}: {a.num_object} objects, last C={sy.num_object} ({time.time() - t:.2f} s)")

    tuned = None
    if args.autotune:
        from quids_amd import tuning
        q.simulate(a, rules[0])
        tuned = tuning.autotune(ctx, a, rules[1:], cap, reps=1, verbose=False)
        q.simulate(a, qcgd.reversed_step)  # back to the warmed-up state
        log(f"autotune: {tuned}")
    ctx.set_profiling(True)
    ctx.reset_phase_times()
    stats = []
    ctx.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        run_iteration(stats)
        log(f"timed iteration {args.warmup + i + 1}: {a.num_object} objects, C={stats[-1]['num_candidates']}")
    ctx.synchronize()
    t1 = time.perf_counter()
    elapsed = t1 - t0
    phases = ctx.phase_times()
    ctx.set_profiling(False)

    C = sum(s["num_candidates"] for s in stats)
    value = C / elapsed
    ab = {k: sum(alg_bytes(s)[k] for s in stats) for k in ("generate", "dedup", "finalize", "total")}
    kernel_ms = {k: phases[k][0] for k in ("generate", "dedup", "finalize")}
    dom = max(kernel_ms, key=kernel_ms.get)
    launches = len(stats)  # each timed simulate() launches the phase's main kernel once
    dom_ms = kernel_ms[dom] / launches
    per_launch = ab[dom] / launches
    achieved = per_launch / (dom_ms / 1e3) if dom_ms > 0 else None
    kname = KERNEL_OF_PHASE[dom].rstrip("<")
    if dom == "generate" and ctx.get_tuning()["split_generate"]:
        kname = "k_generate_fast"
    traffic = None
    try:
        with open(TRAFFIC_FILE) as f:
            tr = json.load(f)
        kt = tr.get("kernels", {}).get(dom, {})
        if tr.get("workload") == workload_key(args) and kt.get("kernel") == kname:
            traffic = kt["hbm_bytes_per_launch"]
    except (OSError, ValueError):
        pass
    roofline = {"bound": "hbm", "kernel": kname, "achieved": achieved / 1e9 if achieved else None,
                "peak": HBM_PEAK / 1e9, "unit": "GB/s", "frac": achieved / HBM_PEAK if achieved else None,
                "traffic": traffic, "alg_bytes_per_launch": per_launch, "ms_per_launch": dom_ms,
                "launches": launches}
    step_frac = ab["total"] / elapsed / HBM_PEAK

    cpu = None if args.no_cpu_baseline else cpu_baseline(args.nodes, args.seed, cap, args.cpu_budget)
    extras = {}
    if not args.no_extras:
        try:
            extras["adjoint"] = adjoint_bench()
        except Exception as e:  # noqa: BLE001
            extras["adjoint"] = {"error": str(e)}

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
                               f"iterations {args.warmup + 1}..{args.warmup + args.steps} timed",
                   "candidates": C, "parallelism": "single GPU"},
        "roofline": roofline,
        "step_hbm_frac": step_frac,
        "phase_ms": {k: round(v[0], 3) for k, v in phases.items()},
        "cpu_baseline": cpu,
        "tuning": ctx.get_tuning() if tuned is None else dict(tuned["tuning"], autotuned=True),
        **extras,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
