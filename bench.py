#!/usr/bin/env python3
"""Benchmark: single-LP PDHG solve to 1e-4 relative KKT on one MI355X.

Workload (default): m = 1,000,000 rows, n = 2,000,000 columns, nnz = 20,000,000,
row lengths Zipf(1.5) clipped to [1, 50000] and rescaled to the exact nnz,
uniform distinct columns, A_ij ~ N(0,1) with 10% replaced by +-1, feasible by
construction (seed 1000), FP64 values, int32 indices, default parameters
(presolve off: there is no presolver), eps 1e-4.

A step is one full solve (scaling, power iteration, restarted Halpern PDHG
loop, result) of that LP from a problem already resident in HBM.  The value is
the mean wall time of a step (seconds, lower is better).  Kernel roofline: the
two fused phase kernels (SpMV + PDHG update + Halpern) are timed live with HIP
events on the solve's stream during the timed steps.

CPU baseline: the OpenMP build of the FP64 oracle on the host cores solving the
same LP completely (median of 3 solves; end-to-end and loop-only times).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "d-pdlp_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

HBM_PEAK_GBPS = 8000.0

WORKLOADS = {
    "zipf-1m": dict(m=1_000_000, n=2_000_000, nnz=20_000_000, row_rule=1, pm1_frac=0.1, seed=1000,
                 desc="m=1e6 n=2e6 nnz=2e7 Zipf(1.5) rows, FP64, eps 1e-4"),
    "small-2k": dict(m=2_000, n=5_000, nnz=50_000, row_rule=0, pm1_frac=0.0, seed=0,
                 desc="m=2e3 n=5e3 nnz=5e4 uniform positions, FP64, eps 1e-4"),
    "zipf-4m": dict(m=4_000_000, n=8_000_000, nnz=120_000_000, row_rule=1, pm1_frac=0.0,
                 seed=2000, desc="m=4e6 n=8e6 nnz=1.2e8 Zipf(1.5) rows"),
    "banded-1m": dict(m=1_000_000, n=2_000_000, per_row=20,
                 desc="m=1e6 n=2e6 nnz=2e7, row i's 20 entries in columns 2i-10..2i+9 "
                      "(local columns), FP64; SpMV and PDHG-phase bandwidth only"),
    "batch-512": dict(count=512, base_seed=3000, m_lo=500, m_hi=200_000,
                 desc="batch of 512 LPs: m log-uniform [500, 2e5], n = m*U[1.5,3], "
                      "Poisson(8)+1 nnz/row, seeds 3000..3511, eps 1e-4"),
    "serve-poisson": dict(rates=(50.0, 200.0, 800.0), large_frac=0.05, large_m=1_000_000,
                 large_nnz=10_000_000, budget_gb=200.0, pool_small=512, pool_large=4,
                 desc="serving: Poisson arrivals, LPs from the batch generator + 5% "
                      "m=1e6 nnz=1e7, 200 GB admission, eps 1e-4"),
}


BATCH_CAP = 4096  # row-block capacity of the batched LPs (nonzeros per workgroup)


def finite(v):
    """JSON-safe number (None for inf / nan)."""
    import math
    if isinstance(v, float) and not math.isfinite(v):
        return None
    return v


def clean(o):
    if isinstance(o, dict):
        return {k: clean(v) for k, v in o.items()}
    if isinstance(o, (list, tuple)):
        return [clean(v) for v in o]
    return finite(o)


def cpu_batch_baseline(probs, budget_s):
    """OpenMP oracle solving batch members in index order for about budget_s."""
    import oracle_py as orc
    L = orc.lib(omp=True)
    threads = L.orc_num_threads()
    prm = orc.default_params(presolve=False)
    t0 = time.perf_counter()
    done = 0
    for prob, _ in probs:
        orc.solve(prob.ptr, prm, omp=True)
        done += 1
        if time.perf_counter() - t0 > budget_s:
            break
    el = time.perf_counter() - t0
    return dict(value=done / el, unit="LPs/s", cores=threads, kind="port",
                sample=f"OpenMP FP64 oracle ({threads} threads) solving batch members 0..{done - 1} "
                       f"to eps 1e-4 one after another: {done} LPs in {el:.1f} s")


def batch_traffic(kernel):
    """Mean HBM bytes per launch of a batched phase kernel, from the committed
    counter profile (FETCH_SIZE/WRITE_SIZE passes over this same leg)."""
    tfile = os.path.join(ROOT, "profiles", "batch512_hbm_traffic.json")
    if not os.path.exists(tfile):
        return None
    with open(tfile) as f:
        return json.load(f).get("bytes_per_launch", {}).get(kernel)


def bench_batch(args, dpdlp):
    w = WORKLOADS["batch-512"]
    count = args.batch_count or w["count"]
    t = time.perf_counter()
    probs = [dpdlp.batch_member(k, w["base_seed"], w["m_lo"], w["m_hi"]) for k in range(count)]
    gen_s = time.perf_counter() - t
    nnz = sum(p.nnz for p, _ in probs)
    print(f"# generated {count} LPs, nnz={nnz} in {gen_s:.1f}s", file=sys.stderr, flush=True)
    params = dpdlp.default_params(presolve=False)
    # per size class, the row-block layout (capacity, rows per block) of A and
    # A' for which the batched phase kernels run fastest, cached by device
    tune = None
    lays = None
    if not args.no_autotune:
        cache = args.tune_cache or os.path.join(os.environ.get("TMPDIR", "/tmp"),
                                                "dpdlp_batch_layouts.txt")
        lays, rep = dpdlp.batch_autotune([p for p, _ in probs], params,
                                         precision=dpdlp.PREC_MIXED, cache_path=cache)
        hist = {}
        for L in lays:
            hist[str(L)] = hist.get(str(L), 0) + 1
        tune = dict(report=rep, layouts=hist, cache=cache)
        print(f"# autotune: {rep}", file=sys.stderr, flush=True)
    t = time.perf_counter()
    plain_res = [dpdlp.Resident(p) for p, _ in probs] if not args.no_plain else []
    cap_at = args.batch_cap_at or BATCH_CAP
    if lays is not None:
        batch_res = [dpdlp.Resident(p, layout=L) for (p, _), L in zip(probs, lays)]
    else:
        batch_res = [dpdlp.Resident(p, spmv_cap=BATCH_CAP, panels=(1, 1), spmv_cap_At=cap_at)
                     for p, _ in probs]
    up_s = time.perf_counter() - t
    for k in range(args.warmup):
        if plain_res:
            plain_res[0].solve(params)
        dpdlp.solve_batch(batch_res[:4], params, spmv_cap=BATCH_CAP, spmv_cap_At=cap_at,
                          precision=dpdlp.PREC_MIXED)
    # plain loop: the single-LP FP64 path, one LP after another
    plain, plain_s = None, None
    if not args.no_plain:
        t = time.perf_counter()
        plain = [r.solve(params)[0] for r in plain_res]
        plain_s = time.perf_counter() - t
        print(f"# plain loop: {plain_s:.2f}s {count / plain_s:.1f} LPs/s", file=sys.stderr,
              flush=True)

    def run_batch(prec, profile):
        st = {}
        t0 = time.perf_counter()
        got = dpdlp.solve_batch(batch_res, params, stats=st, max_lps=count, spmv_cap=BATCH_CAP,
                                spmv_cap_At=cap_at, precision=prec, profile_phases=profile)
        el = time.perf_counter() - t0
        cnt = {k: st[k] for k in ("windows", "syncs", "launches")}
        ph = {"samples": st["phase_samples"], "primal_ms": st["primal_ms"],
              "primal_bytes": st["primal_bytes"], "dual_ms": st["dual_ms"],
              "dual_bytes": st["dual_bytes"]}
        return got, el, cnt, ph

    runs = []
    for k in range(args.steps):
        res, el, cnt, ph = run_batch(dpdlp.PREC_MIXED, 1)
        runs.append(el)
        print(f"# batch mixed step {k}: {el:.2f}s {count / el:.1f} LPs/s {cnt}", file=sys.stderr,
              flush=True)
    mixed, mixed_ph, mixed_cnt = res, ph, cnt
    untuned_s = None
    if lays is not None and not args.no_untuned:
        # the same batch with every LP on the default 4096-nonzero, 256-row blocks
        ures = [dpdlp.Resident(p, spmv_cap=BATCH_CAP, panels=(1, 1)) for p, _ in probs]
        t0 = time.perf_counter()
        dpdlp.solve_batch(ures, params, max_lps=count, spmv_cap=BATCH_CAP,
                          spmv_cap_At=BATCH_CAP, precision=dpdlp.PREC_MIXED)
        untuned_s = time.perf_counter() - t0
        for r in ures:
            r.free()
        print(f"# batch mixed, untuned layout: {untuned_s:.2f}s {count / untuned_s:.1f} LPs/s",
              file=sys.stderr, flush=True)
    fp64, fp64_s, fp64_cnt = None, None, None
    if not args.no_fp64_batch:
        fp64, fp64_s, fp64_cnt, _ = run_batch(dpdlp.PREC_FP64, 0)
        print(f"# batch fp64: {fp64_s:.2f}s {count / fp64_s:.1f} LPs/s", file=sys.stderr,
              flush=True)
    step_s = sum(runs) / len(runs)
    persist = None
    if args.persist_max_nnz > 0:
        persist = persist_legs(dpdlp, batch_res, probs, params, count, cap_at,
                               args.persist_max_nnz)

    def obj_err(a, b):
        if a is None or b is None:
            return None
        return max(abs(x.primal_obj - y.primal_obj) / max(1.0, abs(y.primal_obj))
                   for x, y in zip(a, b))
    statuses = {}
    for r in mixed:
        statuses[r.status] = statuses.get(r.status, 0) + 1
    p1 = mixed_ph["primal_bytes"] / (mixed_ph["primal_ms"] * 1e-3) / 1e9 if mixed_ph["primal_ms"] else None
    p2 = mixed_ph["dual_bytes"] / (mixed_ph["dual_ms"] * 1e-3) / 1e9 if mixed_ph["dual_ms"] else None
    dom = ("batched_primal_phase", p1, mixed_ph["primal_ms"], mixed_ph["primal_bytes"]) \
        if mixed_ph["primal_ms"] >= mixed_ph["dual_ms"] else \
        ("batched_dual_phase", p2, mixed_ph["dual_ms"], mixed_ph["dual_bytes"])
    cpu = None
    if not args.no_cpu:
        print("# cpu baseline sample ...", file=sys.stderr, flush=True)
        cpu = cpu_batch_baseline(probs, args.cpu_budget)
    n_s = max(mixed_ph["samples"], 1)
    out = {
        "metric": "time-to-1e-4 KKT & SpMV HBM GB/s; served LPs/s + p99 latency vs plain loop",
        "value": count / step_s, "unit": "LPs/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": step_s * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32-matrix/f64",
        "data": "synthetic (seeded generator, feasible by construction)",
        "config": {"workload": f"batch-512: {w['desc']}", "lps": count, "nnz_total": nnz,
                   "m_total": sum(p.m for p, _ in probs), "n_total": sum(p.n for p, _ in probs),
                   "parallelism": "one GPU, batched launches"},
        "statuses": statuses,
        "plain_loop": None if plain is None else {
            "s": plain_s, "LPs_per_s": count / plain_s, "dtype": "f64",
            "iterations_mean": sum(r.iterations for r in plain) / count},
        "speedup_vs_plain_loop": None if plain is None else plain_s / step_s,
        "batch_fp64": None if fp64 is None else {
            "s": fp64_s, "LPs_per_s": count / fp64_s,
            "speedup_vs_plain_loop": None if plain is None else plain_s / fp64_s,
            "counters": fp64_cnt},
        "batch_counters": mixed_cnt,
        "autotune": tune,
        "untuned_layout": None if untuned_s is None else {
            "layout": [BATCH_CAP, 256, BATCH_CAP, 256], "s": untuned_s,
            "LPs_per_s": count / untuned_s},
        "objective_rel_diff": {"mixed_vs_plain": obj_err(mixed, plain),
                               "fp64_batch_vs_plain": obj_err(fp64, plain)},
        "phase_ms_per_sample": {"primal": mixed_ph["primal_ms"] / n_s,
                                "dual": mixed_ph["dual_ms"] / n_s, "samples": mixed_ph["samples"]},
        "roofline": {"bound": "hbm", "kernel": dom[0], "achieved": dom[1], "peak": HBM_PEAK_GBPS,
                     "unit": "GB/s", "frac": (dom[1] or 0) / HBM_PEAK_GBPS,
                     "traffic": batch_traffic(dom[0]),
                     "algorithmic_bytes_per_launch": dom[3] / n_s},
        "phase_GBps": {"primal": p1, "dual": p2},
        "persistent_group": persist,
        "cpu_baseline": cpu, "generate_s": gen_s, "upload_s": up_s,
    }
    print(json.dumps(clean(out)))


def persist_legs(dpdlp, res, probs, params, count, cap_at, cut):
    """The mixed batch stepped by hand without and with a persistent group
    (LPs of at most `cut` nonzeros whole on the device): LPs/s, and the
    completion latency of the LPs at or below the cut."""
    out = {"cut_nnz": cut, "lps_at_or_below_cut": sum(1 for p, _ in probs if p.nnz <= cut)}
    for name, c in (("windowed", 0), ("persistent", cut)):
        B = dpdlp.Batch(params, max_lps=count, spmv_cap=BATCH_CAP, spmv_cap_At=cap_at,
                        precision=dpdlp.PREC_MIXED, persist_max_nnz=c)
        t0 = time.perf_counter()
        for k, r in enumerate(res):
            B.add(r, k)
        lat = {}
        while B.in_flight() > 0:
            B.step()
            for tag, r, l in B.take():
                lat[tag] = (l, r.status)
        for tag, r, l in B.take():
            lat[tag] = (l, r.status)
        el = time.perf_counter() - t0
        B.free()
        small = sorted(lat[k][0] for k in range(count) if probs[k][0].nnz <= cut)
        out[name] = {"s": el, "LPs_per_s": count / el,
                     "optimal": sum(1 for v in lat.values() if v[1] == "OPTIMAL"),
                     "small_p50_s": small[len(small) // 2] if small else None,
                     "small_max_s": small[-1] if small else None}
        lat_txt = (f"p50 {small[len(small) // 2]:.3f}s max {small[-1]:.3f}s" if small
                   else "none at or below the cut")
        print(f"# batch {name} (cut {cut}): {el:.2f}s {count / el:.1f} LPs/s, LPs <= cut "
              f"{lat_txt}", file=sys.stderr, flush=True)
    return out


def bench_serve(args, dpdlp):
    import numpy as np
    w = WORKLOADS["serve-poisson"]
    params = dpdlp.default_params(presolve=False)
    rates = [float(x) for x in args.rates.split(",")] if args.rates else list(w["rates"])
    per = []
    for rate in rates:
        row = {"rate": rate}
        for mode, name in ((0, "batched"), (1, "plain_loop")):
            if mode == 1 and args.no_plain:
                continue
            t = time.perf_counter()
            rep, lat, idx, _ = dpdlp.serve(
                params, rate=rate, duration_sec=args.serve_duration, drain_sec=args.serve_drain,
                seed=4000, mode=mode, precision=dpdlp.PREC_MIXED if mode == 0 else dpdlp.PREC_FP64,
                large_frac=w["large_frac"] if args.large_frac < 0 else args.large_frac,
                large_m=w["large_m"], large_nnz=w["large_nnz"],
                pool_small=w["pool_small"], pool_large=w["pool_large"],
                hbm_budget_gb=w["budget_gb"], max_lps=args.serve_max_lps,
                persist_max_nnz=args.persist_max_nnz if mode == 0 else 0)
            rep["wall_s"] = time.perf_counter() - t
            # per request class (pool entries past the small pool are the
            # m = 1e6 requests): arrival -> result latency, inf = unfinished
            small = np.asarray(idx) < w["pool_small"]
            lat = np.asarray(lat)
            for cls, sel in (("small", small), ("large", ~small)):
                lc = lat[sel]
                done = np.isfinite(lc)
                rep[cls] = {"arrived": int(sel.sum()), "completed": int(done.sum()),
                            "p50_sec": float(np.percentile(lc, 50)) if lc.size else None,
                            "p99_sec": float(np.percentile(lc, 99)) if lc.size else None}
            row[name] = rep
            print(f"# rate {rate:.0f}/s {name}: arrived {rep['arrived']} done {rep['completed']} "
                  f"{rep['lps_per_sec']:.1f} LPs/s p50 {rep['p50_sec']:.2f}s p99 {rep['p99_sec']:.2f}s "
                  f"in-flight max {rep['max_in_flight']}", file=sys.stderr, flush=True)
        per.append(row)
    top = per[-1]
    out = {
        "metric": "time-to-1e-4 KKT & SpMV HBM GB/s; served LPs/s + p99 latency vs plain loop",
        "value": top["batched"]["lps_per_sec"], "unit": "LPs/s", "n_gpus": 1, "steps": 1,
        "warmup": 0, "ms_per_step": top["batched"]["makespan_sec"] * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32-matrix/f64", "data": "synthetic (seeded generator, feasible by construction)",
        "config": {"workload": f"serve-poisson: {w['desc']}", "rates": rates,
                   "arrival_window_s": args.serve_duration, "drain_cap_s": args.serve_drain,
                   "max_lps": args.serve_max_lps, "persist_max_nnz": args.persist_max_nnz,
                   "large_frac": w["large_frac"] if args.large_frac < 0 else args.large_frac},
        "per_rate": per,
        "speedup_vs_plain_loop_at_top_rate": top["batched"]["lps_per_sec"] /
        max(top["plain_loop"]["lps_per_sec"], 1e-9) if "plain_loop" in top else None,
        "roofline": None, "cpu_baseline": None,
    }
    print(json.dumps(clean(out)))


def precision_compare(R, dpdlp, opt, eps):
    """Solve the same resident LP to eps with FP64 and with the mixed mode;
    report time-to-1e-4 / time-to-eps and the objective error of each."""
    out = {}
    params = dpdlp.default_params(presolve=False,
                                  termination_criteria__eps_optimal_relative=eps,
                                  termination_criteria__eps_feasible_relative=eps)
    sols = {}
    for name, prec in (("fp64", dpdlp.PREC_FP64), ("mixed", dpdlp.PREC_MIXED)):
        t = time.perf_counter()
        r, o = R.solve(params, precision=prec, profile_phases=1)
        wall = time.perf_counter() - t
        sols[name] = r
        out[name] = dict(status=r.status, iterations=r.iterations, fp32_iterations=o.iterations_fp32,
                         wall_s=wall, loop_s=o.loop_time_sec, time_to_1e4_s=o.time_to_1e4_sec,
                         time_to_1e6_s=o.time_to_1e6_sec, objective=r.primal_obj,
                         rel_err_vs_known_opt=abs(r.primal_obj - opt) / max(1.0, abs(opt)),
                         rel_kkt=[r.rel_primal, r.rel_dual, r.rel_gap],
                         phase_ms=[o.primal_phase_ms, o.dual_phase_ms])
        print(f"# {name}: {r.status} iters={r.iterations} (fp32 {o.iterations_fp32}) "
              f"wall={wall:.2f}s t1e-4={o.time_to_1e4_sec:.2f}s t1e-6={o.time_to_1e6_sec:.2f}s",
              file=sys.stderr, flush=True)
    import numpy as np
    a, b = sols["fp64"], sols["mixed"]
    out["mixed_vs_fp64"] = dict(
        objective_rel=abs(a.primal_obj - b.primal_obj) / max(1.0, abs(a.primal_obj)),
        x_rel=float(np.linalg.norm(a.x - b.x) / max(np.linalg.norm(a.x), 1e-300)),
        y_rel=float(np.linalg.norm(a.y - b.y) / max(np.linalg.norm(a.y), 1e-300)))
    return out


def bench_banded(args, dpdlp):
    """The zipf-1m sizes with local columns: row i holds 20 entries in
    columns 2i-10 .. 2i+9, so neighbouring rows gather the same lines and the
    gathers hit in cache.  The same kernels then run at the bandwidth of the
    matrix stream, which separates the kernels' streaming efficiency from
    the gather rate that bounds random-column matrices."""
    import numpy as np
    w = WORKLOADS["banded-1m"]
    m, n, per = w["m"], w["n"], w["per_row"]
    nnz = m * per
    start = np.clip(2 * np.arange(m, dtype=np.int64) - per // 2, 0, n - per)
    ci = (start[:, None] + np.arange(per)[None, :]).astype(np.int32).ravel()
    rp = np.arange(0, nnz + 1, per, dtype=np.int32)
    v = np.random.default_rng(7).standard_normal(nnz)
    prob = dpdlp.Problem.from_csr(m, n, rp, ci, v, c=-np.ones(n), con_lb=np.full(m, -np.inf),
                                  con_ub=np.ones(m), var_lb=np.zeros(n), var_ub=np.ones(n))
    R = dpdlp.Resident(prob, spmv_cap=4096, panels=(1, 1))
    ta = R.time_spmv(False, reps=20)
    tt = R.time_spmv(True, reps=20)
    ba, bt = spmv_bytes(m, n, nnz), spmv_bytes(n, m, nnz)
    # 300 iterations timed, whatever the KKT error (the tolerances are out of reach)
    params = dpdlp.default_params(presolve=False, termination_criteria__iteration_limit=300,
                                  termination_criteria__eps_optimal_relative=1e-15,
                                  termination_criteria__eps_feasible_relative=1e-15)
    for _ in range(args.warmup):
        R.solve(params, profile_phases=1)
    t1 = t2 = 0.0
    steps = max(args.steps, 1)
    t = time.perf_counter()
    for _ in range(steps):
        r, o = R.solve(params, profile_phases=1)
        t1 += o.primal_phase_ms
        t2 += o.dual_phase_ms
    wall = (time.perf_counter() - t) / steps
    t1, t2 = t1 / steps, t2 / steps
    p1, p2 = phase_bytes(m, n, nnz)
    g = lambda b, ms: b / (ms * 1e-3) / 1e9
    out = {
        "metric": "SpMV and PDHG-phase HBM GB/s on a banded matrix", "value": g(ba, ta),
        "unit": "GB/s", "n_gpus": 1, "steps": steps, "warmup": args.warmup,
        "ms_per_step": wall * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic (banded, seeded values)",
        "config": {"workload": f"banded-1m: {w['desc']}", "m": m, "n": n, "nnz": nnz,
                   "layout": "A and A' untiled, 4096-nonzero row blocks"},
        "spmv": {"A_ms": ta, "At_ms": tt, "A_GBps": g(ba, ta), "At_GBps": g(bt, tt),
                 "A_bytes": ba, "At_bytes": bt},
        "phase_ms": {"phase1_At_y_primal": t1, "phase2_A_x_dual": t2, "iterations": 300},
        "roofline": {"bound": "hbm", "kernel": "phase2_A_x_dual", "achieved": g(p2, t2),
                     "peak": HBM_PEAK_GBPS, "unit": "GB/s", "frac": g(p2, t2) / HBM_PEAK_GBPS,
                     "traffic": None, "algorithmic_bytes_per_launch": p2},
        "phase1_GBps": g(p1, t1), "phase1_frac": g(p1, t1) / HBM_PEAK_GBPS,
        "spmv_A_frac": g(ba, ta) / HBM_PEAK_GBPS, "spmv_At_frac": g(bt, tt) / HBM_PEAK_GBPS,
        "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)
    R.free()


def spmv_bytes(rows, cols, nnz, sv=8, sx=8):
    """nnz*(s_v+4) + 4(r+1) + s_x*c + 8r : values+indices, row pointers, the
    gathered vector once, the output once."""
    return nnz * (sv + 4) + 4 * (rows + 1) + sx * cols + 8 * rows


def phase_bytes(m, n, nnz, sv=8, sx=8):
    """Algorithmic bytes of the two fused phases of one PDHG iteration.
    Phase 1 (A'y, primal step, Halpern on x): A' stream, y gather, reads x_cur,
    x_init, c, l_v, u_v and writes x_cur, x_ref.  Phase 2 (A x_ref, dual step,
    Halpern on y): A stream, x_ref gather, reads y_cur, y_init, l_c, u_c, writes
    y_cur."""
    p1 = nnz * (sv + 4) + 4 * (n + 1) + sx * m + 48 * n + sx * n
    p2 = nnz * (sv + 4) + 4 * (m + 1) + sx * n + 40 * m
    return p1, p2


def host_cores():
    """Physical cores and logical CPUs of this host (lscpu), for the record."""
    import subprocess
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
    except (OSError, subprocess.SubprocessError):
        return None
    kv = {}
    for line in out.splitlines():
        if ":" in line:
            k, v = line.split(":", 1)
            kv[k.strip()] = v.strip()
    try:
        cores = int(kv["Core(s) per socket"]) * int(kv["Socket(s)"])
    except (KeyError, ValueError):
        cores = None
    return {"physical_cores": cores, "logical_cpus": int(kv.get("CPU(s)", "0") or 0),
            "model": kv.get("Model name")}


def cpu_baseline(problem, repeats, gpu_result):
    """Complete solves of the same LP by the OpenMP build of the FP64 oracle
    (scaling, transpose, power iteration, PDHG loop to 1e-4, result), median of
    `repeats`: end-to-end wall time and the loop time (the reference's solve
    time, which starts after setup).  The oracle's status, iterations and
    objective are reported beside the GPU's."""
    import statistics
    import oracle_py as orc
    L = orc.lib(omp=True)
    threads = L.orc_num_threads()
    prm = orc.default_params(presolve=False)
    walls, loops, last = [], [], None
    for k in range(repeats):
        t = time.perf_counter()
        last, _ = orc.solve(problem.ptr, prm, omp=True)
        walls.append(time.perf_counter() - t)
        loops.append(orc.last_timings(omp=True)["loop"])
        print(f"# cpu baseline solve {k}: {last['status']} iters={last['total_count']} "
              f"wall={walls[-1]:.2f}s loop={loops[-1]:.2f}s", file=sys.stderr, flush=True)
    tm = orc.last_timings(omp=True)
    r, o = gpu_result
    obj = last["primal_objective_value"]
    return dict(value=statistics.median(walls), unit="s", cores=threads, kind="port",
                sample=(f"{repeats} complete solves of the same LP by the OpenMP FP64 oracle "
                        f"({threads} threads) to eps 1e-4; value = median end-to-end wall"),
                loop_s=statistics.median(loops), walls_s=walls, loops_s=loops,
                setup_s=tm["rescale"] + tm["transpose"], power_s=tm["power"],
                host=host_cores(),
                oracle=dict(status=last["status"], iterations=last["total_count"], objective=obj,
                            gpu_objective_rel_diff=abs(r.primal_obj - obj) / max(1.0, abs(obj)),
                            gpu_iterations=r.iterations))


def gather_sweep(dpdlp, gathers=20_000_000, reps=10):
    """Random 8-byte gathers with the SpMV's index stream at growing working
    sets: each XCD reading only its own slice (32 KiB .. 256 MiB per XCD), and
    every workgroup reading one shared vector (the two SpMVs' vectors)."""
    out = {"xcd_local": {}, "global": {}}
    for sz in (32 << 10, 256 << 10, 2 << 20, 4 << 20, 16 << 20, 256 << 20):
        out["xcd_local"][str(sz)] = dpdlp.lib.dpdlp_gather_calibration_ex(sz, gathers, reps, 1, 2048)
    for sz in (2 << 20, 8_000_000, 16_000_000, 256 << 20):
        out["global"][str(sz)] = dpdlp.lib.dpdlp_gather_calibration_ex(sz, gathers, reps, 0, 2048)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="zipf-1m", choices=sorted(WORKLOADS))
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-repeats", type=int, default=3,
                    help="complete OpenMP oracle solves for the CPU baseline (median)")
    ap.add_argument("--spmv-cap", type=int, default=0)
    ap.add_argument("--no-tune", action="store_true", help="panel counts from the size rule")
    ap.add_argument("--calib-copy", action="store_true",
                    help="also run a 256 MiB plain copy kernel (counter calibration)")
    ap.add_argument("--mixed", action="store_true", help="FP32 matrix / FP64 state")
    ap.add_argument("--verbose", action="store_true", help="print the solver's iteration log")
    ap.add_argument("--batch-count", type=int, default=0, help="batch: LPs in the batch")
    ap.add_argument("--cpu-budget", type=float, default=20.0, help="batch: CPU sample seconds")
    ap.add_argument("--no-plain", action="store_true", help="batch: skip the plain-loop leg")
    ap.add_argument("--batch-cap-at", type=int, default=0,
                    help="batch: row-block capacity of the batched A' (0: as A)")
    ap.add_argument("--no-fp64-batch", action="store_true", help="batch: skip the FP64 batch leg")
    ap.add_argument("--no-autotune", action="store_true",
                    help="batch: every LP on 4096-nonzero blocks instead of the autotuned layouts")
    ap.add_argument("--no-untuned", action="store_true",
                    help="batch: skip the comparison run on the untuned layout")
    ap.add_argument("--tune-cache", default="", help="batch: layout cache file of the autotuner")
    ap.add_argument("--rates", default="", help="serving: comma-separated arrival rates")
    ap.add_argument("--serve-duration", type=float, default=60.0, help="serving: arrival window s")
    ap.add_argument("--serve-drain", type=float, default=30.0, help="serving: drain cap s")
    ap.add_argument("--serve-max-lps", type=int, default=512, help="serving: LPs in flight")
    ap.add_argument("--large-frac", type=float, default=-1.0,
                    help="serving: share of m = 1e6 requests (default: the workload's 0.05)")
    ap.add_argument("--persist-max-nnz", type=int, default=0,
                    help="batch and serving: LPs with at most this many nonzeros run whole in "
                         "the persistent kernel (0: none)")
    ap.add_argument("--compare-precision", type=float, default=0.0, metavar="EPS",
                    help="solve to EPS in FP64 and in mixed precision and report both")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this solver runs one LP on one GPU; multi-GPU runs are not "
                          "part of this build (bench.py --gpus 1)"}))
        sys.exit(2)

    import dpdlp
    if args.workload == "batch-512":
        return bench_batch(args, dpdlp)
    if args.workload == "serve-poisson":
        return bench_serve(args, dpdlp)
    if args.workload == "banded-1m":
        return bench_banded(args, dpdlp)
    w = WORKLOADS[args.workload]
    t = time.perf_counter()
    prob, opt = dpdlp.generate(w["m"], w["n"], w["nnz"], row_rule=w["row_rule"], seed=w["seed"],
                               pm1_frac=w["pm1_frac"])
    gen_s = time.perf_counter() - t
    print(f"# generated {prob.m}x{prob.n} nnz={prob.nnz} in {gen_s:.1f}s", file=sys.stderr,
          flush=True)
    import numpy as np
    rp = prob.arrays()["row_ptr"] if prob.m <= 5_000_000 else None
    row_stats = None
    if rp is not None:
        lens = np.diff(rp)
        row_stats = dict(max=int(lens.max()), mean=float(lens.mean()),
                         p99=float(np.percentile(lens, 99)))
    t = time.perf_counter()
    R = dpdlp.Resident(prob, spmv_cap=args.spmv_cap, tune=not args.no_tune)
    up_s = time.perf_counter() - t
    print(f"# uploaded in {up_s:.1f}s, row blocks A={R.nblk_A} A'={R.nblk_At} "
          f"panels={R.panels}", file=sys.stderr, flush=True)
    if args.compare_precision > 0:
        res = precision_compare(R, dpdlp, opt, args.compare_precision)
        print(json.dumps({"workload": f"{args.workload}: {w['desc']}", "m": prob.m, "n": prob.n,
                          "nnz": prob.nnz, "eps": args.compare_precision, "known_optimum": opt,
                          "panels": R.panels, **res}))
        return
    params = dpdlp.default_params(presolve=False, verbose=args.verbose)
    prec = dpdlp.PREC_MIXED if args.mixed else dpdlp.PREC_FP64
    for k in range(args.warmup):
        r, o = R.solve(params, precision=prec)
        print(f"# warmup {k}: {r.status} iters={r.iterations} total={o.total_time_sec:.3f}s",
              file=sys.stderr, flush=True)
    times, loops, p1s, p2s, nsamp = [], [], [], [], 0
    last = None
    for k in range(args.steps):
        t = time.perf_counter()
        r, o = R.solve(params, precision=prec, profile_phases=1)
        dt = time.perf_counter() - t
        times.append(dt)
        loops.append(o.loop_time_sec)
        if o.phase_samples:
            p1s.append(o.primal_phase_ms)
            p2s.append(o.dual_phase_ms)
            nsamp += o.phase_samples
        last = (r, o)
        print(f"# step {k}: {r.status} iters={r.iterations} wall={dt:.3f}s loop={o.loop_time_sec:.3f}s "
              f"pi={o.power_iterations} ({o.power_time_sec:.3f}s) setup={o.setup_time_sec:.3f}s "
              f"p1={o.primal_phase_ms:.4f}ms p2={o.dual_phase_ms:.4f}ms", file=sys.stderr, flush=True)
    r, o = last
    step_s = sum(times) / len(times)
    m, n, nnz = prob.m, prob.n, prob.nnz
    sv = 4 if args.mixed else 8
    b1, b2 = phase_bytes(m, n, nnz, sv=sv, sx=sv)
    p1 = sum(p1s) / len(p1s) if p1s else float("nan")
    p2 = sum(p2s) / len(p2s) if p2s else float("nan")
    # dominant kernel: the phase with the larger share of loop time
    dom = ("phase1_At_y_primal", b1, p1) if p1 >= p2 else ("phase2_A_x_dual", b2, p2)
    achieved = dom[1] / (dom[2] * 1e-3) / 1e9
    # standalone SpMV timing (same kernels, plain store epilogue)
    ta = R.time_spmv(False, reps=20)
    tt = R.time_spmv(True, reps=20)
    spmv = dict(A_ms=ta, At_ms=tt,
                A_GBps=spmv_bytes(m, n, nnz) / (ta * 1e-3) / 1e9,
                At_GBps=spmv_bytes(n, m, nnz) / (tt * 1e-3) / 1e9)
    iter_GBps = (b1 + b2) / ((p1 + p2) * 1e-3) / 1e9
    copy_GBps = None
    if args.calib_copy:
        nbytes = 256 << 20
        cms = dpdlp.lib.dpdlp_copy_calibration(nbytes, 5)
        copy_GBps = 2 * nbytes / (cms * 1e-3) / 1e9
    # what bounds the phases: every nonzero gathers one vector entry per
    # launch; the ceiling is the measured rate of random gathers with the
    # same index stream over the same working set (phase 1 gathers y from
    # one m-vector; phase 2 gathers x, each XCD from its panel's n/P slice)
    sweep = gather_sweep(dpdlp)
    pA = R.panels[0]
    ws1 = dict(kind="global", bytes=8 * m)
    ws2 = dict(kind="xcd_local", bytes=8 * n // pA) if pA > 1 else dict(kind="global", bytes=8 * n)

    def ceiling(ws, values):
        return dpdlp.lib.dpdlp_gather_calibration_values(
            ws["bytes"], nnz, 10, 1 if ws["kind"] == "xcd_local" else 0, 2048, values)
    c1, c2 = ceiling(ws1, 0), ceiling(ws2, 0)
    # the same gathers with the FP64 values streamed beside the indices: the
    # ceiling of a pass that reads the matrix as the SpMV does (12 B/nonzero)
    v1, v2 = ceiling(ws1, 1), ceiling(ws2, 1)
    g1, g2 = nnz / (p1 * 1e-3) / 1e9, nnz / (p2 * 1e-3) / 1e9
    gather = dict(gathers_per_launch=nnz,
                  phase1_G_per_s=g1, phase1_working_set=ws1, phase1_ceiling_G_per_s=c1,
                  phase1_frac=g1 / c1 if c1 > 0 else None,
                  phase1_ceiling_with_values_G_per_s=v1,
                  phase1_frac_with_values=g1 / v1 if v1 > 0 else None,
                  phase2_G_per_s=g2, phase2_working_set=ws2, phase2_ceiling_G_per_s=c2,
                  phase2_frac=g2 / c2 if c2 > 0 else None,
                  phase2_ceiling_with_values_G_per_s=v2,
                  phase2_frac_with_values=g2 / v2 if v2 > 0 else None,
                  ceiling_kernel="k_gather_calib: coalesced index stream, 2048 indices per "
                                 "workgroup, 8 gathers in flight per thread; *_with_values: "
                                 "FP64 values streamed beside the indices",
                  sweep_G_per_s=sweep)
    traffic = None
    tfile = os.path.join(ROOT, "profiles", "zipf1m_fp64_hbm_traffic.json")
    if args.workload == "zipf-1m" and not args.mixed and os.path.exists(tfile):
        with open(tfile) as f:
            tj = json.load(f)
        traffic = tj.get("bytes_per_launch", {}).get(dom[0])
    cpu = None
    if not args.no_cpu:
        print("# cpu baseline ...", file=sys.stderr, flush=True)
        cpu = cpu_baseline(prob, args.cpu_repeats, last)
    out = {
        "metric": "time-to-1e-4 KKT & SpMV HBM GB/s; served LPs/s + p99 latency vs plain loop",
        "value": step_s,
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": step_s * 1e3,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32-matrix/f64" if args.mixed else "f64",
        "data": "synthetic (seeded generator, feasible by construction)",
        "config": {"workload": f"{args.workload}: {w['desc']}", "m": m, "n": n, "nnz": nnz,
                   "row_lengths": row_stats, "seed": w["seed"], "known_optimum": opt},
        "status": r.status,
        "iterations": r.iterations,
        "loop_time_s": sum(loops) / len(loops),
        "power_iterations": o.power_iterations,
        "power_time_s": o.power_time_sec,
        "setup_time_s": o.setup_time_sec,
        "time_to_1e4_s": o.time_to_1e4_sec,
        "objective": r.primal_obj,
        "objective_rel_err_vs_known_opt": abs(r.primal_obj - opt) / max(1.0, abs(opt)),
        "rel_kkt": [r.rel_primal, r.rel_dual, r.rel_gap],
        "phase_ms": {"phase1_At_y_primal": p1, "phase2_A_x_dual": p2, "samples": nsamp},
        "iteration_GBps": iter_GBps,
        "spmv": spmv,
        "roofline": {"bound": "hbm", "kernel": dom[0], "achieved": achieved,
                     "peak": HBM_PEAK_GBPS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBPS,
                     "traffic": traffic, "algorithmic_bytes_per_launch": dom[1]},
        "gather_bound": gather,
        "copy_kernel_GBps": copy_GBps,
        "cpu_baseline": cpu,
        "xcd_panels": R.panels,
        "layout": {"panels": R.panels, "row_block_caps": R.caps, "rows_renumbered": R.renumbered},
        "generate_s": gen_s,
        "upload_s": up_s,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
