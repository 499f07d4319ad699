"""Benchmark: simulation steps per second of the single-large-body workload.

Workload (one GPU): a 64 x 64 x 64-cell tet grid (274,625 particles,
1,310,720 corotational tetrahedra, E = 1e5 Pa, nu = 0.3, 1000 kg/m^3,
jitter U(-0.005, 0.005) m, static top face) plus a 512 x 512-vertex cloth
(stretch k = 1e4, shear k = 1e3, pinned corners) draping over a 0.5 m sphere;
dt = 1/60 s, one Newton iteration, PCG tolerance 1e-5 with at most 50
iterations. One step advances both bodies (b3Body::Step on each), as two
handles on their own streams (--batch: one batch of two instances, stepped by
one launch sequence).

Prints one JSON line: steps/s (whole job), the roofline of the dominant
kernel (the PCG's block-sparse SpMV, timed with one HIP event pair around
back-to-back launches on the system of the last timed step) and the CPU
baseline (the reference library compiled from its own sources, oracle/_ref,
timed on this host's cores on a bounded sample; the oracle's restatement
stands in when that build is absent).
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
PROFILE_PMC = os.path.join(ROOT, "profiles", "k_cg_spmv_dq_tet64_pmc.txt")
PROFILE_PMC_BATCH = os.path.join(ROOT, "profiles", "k_cg_spmv_dq_batch4096_pmc.txt")
PROFILE_PMC_CONTACT = os.path.join(ROOT, "profiles", "k_cg_spmv_dq_contact256_pmc.txt")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def traffic_from_profile(path=PROFILE_PMC, workload=None):
    """HBM bytes per SpMV launch measured by rocprofv3 --pmc, from a profile summary under profiles/, if present.
    With `workload`, the summary's `workload:` line must equal it: a figure applies only to the system it was
    measured on."""
    fields = {}
    try:
        for line in open(path):
            key, sep, rest = line.partition(":")
            if sep and key in ("workload", "traffic_bytes_per_launch"):
                fields[key] = rest.strip()
    except OSError:
        return None
    if workload is not None and fields.get("workload") != workload:
        return None
    try:
        return float(fields["traffic_bytes_per_launch"])
    except (KeyError, ValueError):
        return None


def host_cpu():
    """The host's CPU model and the number of CPUs this process may run on."""
    model = "unknown CPU"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return f"{model}, {len(os.sched_getaffinity(0))} CPUs available to the process"


def emit(line):
    """Prints the bench line; a CPU baseline also names the host it ran on."""
    if isinstance(line.get("cpu_baseline"), dict):
        line["cpu_baseline"]["host"] = host_cpu()
    print(json.dumps(line), flush=True)


def cpu_baseline(tet_arr, cloth_arr, dt, sub, sub_tol, sample_steps):
    """The reference build (or the oracle) on both bodies: after set-up, one untimed warm-up step, then
    `sample_steps` steps timed one by one; a step of both bodies costs the sum of the two medians."""
    import oracle

    kind = "reference" if oracle.reference_available() else "port"
    medians, spent = [], 0.0
    for arr in (tet_arr, cloth_arr):
        body = oracle.RefBody(arr, fast=True) if kind == "reference" else oracle.OracleBody(arr)
        spent += body.step_timed(1, dt, 1, sub, 0.0, sub_tol)   # warm-up
        times = [body.step_timed(1, dt, 1, sub, 0.0, sub_tol) for _ in range(sample_steps)]
        spent += sum(times)
        medians.append(float(np.median(times)))
        del body
    return {
        "value": round(1.0 / sum(medians), 6),
        "unit": "steps/s",
        "cores": 1,
        "kind": kind,
        "sample": f"the 64^3 tet body and the 512^2 cloth after set-up, single thread: 1 warm-up step, then the "
                  f"median of {sample_steps} timed steps each (tet {medians[0]:.2f} s, cloth {medians[1]:.2f} s); "
                  f"{spent:.1f} s of CPU time",
    }


def _ref_step(job):
    """One reference step of one body in a worker process: (worker pid, seconds of the step alone)."""
    arr, dt, sub, tol = job
    import oracle
    body = oracle.RefBody(arr, fast=True) if oracle.reference_available() else oracle.OracleBody(arr)
    return os.getpid(), body.step_timed(1, dt, 1, sub, 0.0, tol)


def cpu_body_parallel(sample, dt, sub, tol):
    """The reference stepping independent bodies on all the host CPUs it may use (at most 16): one process per
    CPU, each body one step. Returns (seconds, workers), the seconds being the busiest worker's summed step
    time (set-up excluded). The workers are forked, so this runs before the process makes its first HIP call."""
    import multiprocessing as mp

    workers = max(1, min(16, len(os.sched_getaffinity(0)), len(sample)))
    # largest bodies first, so the pool's dynamic hand-out balances the workers
    order = sorted(sample, key=lambda a: -len(a["p_type"]))
    with mp.get_context("fork").Pool(workers) as pool:
        done = pool.map(_ref_step, [(a, dt, sub, tol) for a in order], chunksize=1)
    per_worker = {}
    for pid, t in done:
        per_worker[pid] = per_worker.get(pid, 0.0) + t
    return max(per_worker.values()), workers


def autotune(bodies, cache):
    """b3g_autotune on each body after its timed steps: the chosen schedule and the search's ms per step (whole
    steps from the body's state, which is restored) with it and with the default."""
    from bounce_softbody_amd import _lib

    out = []
    for b in bodies:
        t0 = time.perf_counter()
        chosen, tuned_ms, default_ms = b.autotune(cache)
        out.append({"chosen": dict(zip(_lib.SCHEDULE_FIELDS, chosen)),
                    "default": dict(zip(_lib.SCHEDULE_FIELDS, _lib.DEFAULT_SCHEDULE)),
                    "search_ms_per_step": {"chosen": round(tuned_ms, 4), "default": round(default_ms, 4)},
                    "search_s": round(time.perf_counter() - t0, 2)})
        log(f"autotune: {out[-1]}")
    return out


def run_precision(args):
    """Precision form on the stiffness-contrast workload: a tet grid (args.tet cells per axis) whose elements
    have Young's moduli log-uniform in [1e4, 1e7] Pa. Every form steps from the same deformed, moving state
    (after args.pre_steps plain steps); its error is the relative max-norm of one step's velocities against
    the FP64 oracle (the reference restated in double, PCG tolerance 1e-10) from that state."""
    from bounce_softbody_amd import _lib, configs
    from bounce_softbody_amd.body import Body

    dt = configs.DT
    n = args.tet
    arr = configs.tet_grid(n, young_per_element=configs.log_uniform_young(n, seed=7)).arrays()
    body = Body(arr)
    for _ in range(args.pre_steps):
        body.Step(dt, 1, 200, 0.0, 1e-6)
    body.synchronize()
    start = body.save_state()
    x0, v0 = body.state()
    v64 = None
    if not args.no_cpu_baseline:
        import oracle
        t0 = time.time()
        o = oracle.OracleBody(arr, precision=64)
        o.set_state(x0, v0)
        o.step(dt, 1, 20000, 0.0, 1e-10)
        v64 = o.state64()[1]
        del o
        log(f"FP64 oracle step in {time.time() - t0:.1f} s")
    forms = [
        ("plain_fp32", _lib.SOLVER_PCG, (), args.sub_iterations, args.sub_tolerance),
        ("plain_fp32_tight", _lib.SOLVER_PCG, (), 1000, 1e-7),
        ("refined_fp32", _lib.SOLVER_REFINED_FP32, (args.inner_iterations, args.inner_tolerance),
         args.max_sweeps * args.inner_iterations, args.refine_tolerance),
        ("refined_bf16", _lib.SOLVER_REFINED_BF16, (args.inner_iterations, args.inner_tolerance),
         args.max_sweeps * args.inner_iterations, args.refine_tolerance),
    ]
    out = {}
    for name, mode, inner, sub, tol in forms:
        body.set_solver(mode, *inner)
        body.load_state(start)
        body.Step(dt, 1, sub, 0.0, tol)
        _, v1 = body.state()
        err = float(np.abs(v1 - v64).max() / np.abs(v64).max()) if v64 is not None else None
        st = body.instance_stats()[0][0]
        sweeps, res = body.refinement()[0]
        body.load_state(start)
        for _ in range(args.warmup):
            body.Step(dt, 1, sub, 0.0, tol)
        body.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            body.Step(dt, 1, sub, 0.0, tol)
        body.synchronize()
        el = time.perf_counter() - t0
        mean_ms, by = body.time_spmv(args.spmv_launches)
        out[name] = {"steps_per_s": round(args.steps / el, 3), "rel_err_vs_fp64": err,
                     "inner_iterations_first_step": int(st[2]), "sweeps": int(sweeps),
                     "fp64_relative_residual": float(res) if mode != _lib.SOLVER_PCG else None,
                     "pcg": {"max_iterations": sub, "tolerance": tol, "inner": list(inner)},
                     "spmv": {"mean_us": round(mean_ms * 1e3, 2), "bytes_per_launch": by,
                              "gbs": round(by / (mean_ms * 1e-3) / 1e9, 1)}}
        log(f"{name}: {out[name]}")
    best_name = max(("refined_fp32", "refined_bf16"), key=lambda k: out[k]["steps_per_s"])
    best = out[best_name]
    line = {
        "metric": "sim steps/s",
        "value": best["steps_per_s"],
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 / best["steps_per_s"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 inner / f64 residual",
        "data": "synthetic",
        "config": {"workload": f"precision form: tet grid {n}^3 cells, E log-uniform in [1e4, 1e7] Pa",
                   "dt": "1/60", "newton_iterations": 1, "form": best_name + " (value)"},
        "forms": out,
        "speedup_vs_plain_fp32_tight": round(best["steps_per_s"] / out["plain_fp32_tight"]["steps_per_s"], 3),
        "note": "plain FP32 cannot reach 1e-5 here: its right-hand side loses ~1e-5 to cancellation in FP32",
        "roofline": {"kernel": "k_cg_spmv_dq, " + ("BF16 off-diagonal" if best_name == "refined_bf16" else "FP32")
                               + " operator (back-to-back pass)", "bound": "hbm",
                     "achieved": best["spmv"]["gbs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(best["spmv"]["gbs"] / HBM_PEAK_GBS, 4), "traffic": None,
                     "bytes_per_launch": best["spmv"]["bytes_per_launch"], "mean_launch_us": best["spmv"]["mean_us"]},
        "cpu_baseline": None,
    }
    emit(line)


def run_batch_precision(args):
    """Precision form on the many-body batch (args.bodies tet boxes, as in the batch workload): every solver form
    steps the whole batch, captured, from the same state (after args.pre_steps steps); its error is the largest,
    over the first args.cpu_sample_bodies bodies, relative max-norm of one step's velocities against the FP64
    oracle (the reference restated in double, PCG tolerance 1e-10) from that body's state. value = steps/s of the
    fastest refined form within 1e-5; config has every form and the ratio to the fastest plain FP32 form that
    meets 1e-5 (or the tightest plain form when none does)."""
    import oracle
    from bounce_softbody_amd import _lib, configs
    from bounce_softbody_amd.body import Body

    dt = configs.DT
    t0 = time.time()
    arrs = configs.batch_boxes(args.bodies)
    batch = Body(arrs)
    batch.set_graph(not args.no_graph)
    log(f"{args.bodies} bodies built in {time.time() - t0:.1f} s; system {batch.system_size()}")
    for _ in range(args.pre_steps):
        batch.Step(dt, 1, 200, 0.0, 1e-6)
    batch.synchronize()
    start = batch.save_state()
    x0, v0 = batch.state()
    xs, vs = batch.split(x0), batch.split(v0)
    sample = list(range(min(args.cpu_sample_bodies, len(arrs))))
    t0 = time.time()
    v64 = {}
    for k in sample:
        o = oracle.OracleBody(arrs[k], precision=64)
        o.set_state(xs[k], vs[k])
        o.step(dt, 1, 20000, 0.0, 1e-10)
        v64[k] = o.state64()[1]
        del o
    log(f"FP64 oracle steps of {len(sample)} bodies in {time.time() - t0:.1f} s")
    forms = [
        ("plain_fp32", _lib.SOLVER_PCG, (), args.sub_iterations, args.sub_tolerance),
        ("plain_fp32_200", _lib.SOLVER_PCG, (), 200, 1e-6),
        ("plain_fp32_tight", _lib.SOLVER_PCG, (), 1000, 1e-7),
        ("refined_fp32", _lib.SOLVER_REFINED_FP32, (args.inner_iterations, args.inner_tolerance),
         args.max_sweeps * args.inner_iterations, args.refine_tolerance),
        ("refined_bf16", _lib.SOLVER_REFINED_BF16, (args.inner_iterations, args.inner_tolerance),
         args.max_sweeps * args.inner_iterations, args.refine_tolerance),
    ]
    out = {}
    for name, mode, inner, sub, tol in forms:
        batch.set_solver(mode, *inner)
        batch.load_state(start)
        batch.Step(dt, 1, sub, 0.0, tol)
        _, v1 = batch.state()
        v1s = batch.split(v1)
        err = max(float(np.abs(v1s[k] - v64[k]).max() / np.abs(v64[k]).max()) for k in sample)
        abs_err = max(float(np.abs(v1s[k] - v64[k]).max()) for k in sample)
        batch.load_state(start)
        for _ in range(args.warmup):
            batch.Step(dt, 1, sub, 0.0, tol)
        batch.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            batch.Step(dt, 1, sub, 0.0, tol)
        batch.synchronize()
        el = time.perf_counter() - t0
        out[name] = {"steps_per_s": round(args.steps / el, 3), "max_rel_err_vs_fp64": err,
                     "max_abs_err_m_per_s": abs_err,
                     "pcg": {"max_iterations": sub, "tolerance": tol, "inner": list(inner)}}
        log(f"{name}: {out[name]}")
    refined = [k for k in ("refined_fp32", "refined_bf16") if out[k]["max_rel_err_vs_fp64"] <= 1e-5]
    best_name = max(refined or ["refined_fp32", "refined_bf16"], key=lambda k: out[k]["steps_per_s"])
    # the same first bodies stepped by plain FP32 as one handle each and as one batch, against the chosen
    # refined form as one batch (all from the same saved states, steps as above)
    bmode, binner, bsub, btol = next((m, i, s_, t) for n_, m, i, s_, t in forms if n_ == best_name)
    sub_arrs = arrs[: args.handles_subset]
    handles = [Body(a) for a in sub_arrs]
    small = Body(sub_arrs)
    small.set_graph(not args.no_graph)
    for k, h in enumerate(handles):
        h.set_state(xs[k], vs[k])
    sx = np.concatenate([xs[k] for k in range(len(sub_arrs))])
    sv = np.concatenate([vs[k] for k in range(len(sub_arrs))])

    def timed(bodies, setup, sub, tol):
        for b in bodies:
            setup(b)
        for b in bodies:
            b.Step(dt, 1, sub, 0.0, tol)
        for b in bodies:
            b.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            for b in bodies:
                b.Step(dt, 1, sub, 0.0, tol)
        for b in bodies:
            b.synchronize()
        return args.steps / (time.perf_counter() - t0)

    plain = (args.sub_iterations, args.sub_tolerance)
    small.set_state(sx, sv)
    hs = [h.save_state() for h in handles]
    ss = small.save_state()
    plain_handles = timed(handles, lambda b: None, *plain)
    for h, blob in zip(handles, hs):
        h.load_state(blob)
    plain_batch = timed([small], lambda b: b.load_state(ss), *plain)
    refined_batch = timed([small], lambda b: (b.load_state(ss), b.set_solver(bmode, *binner)), bsub, btol)
    del handles, small
    subset = {"bodies": len(sub_arrs), "steps_per_s": {"plain_fp32_handles": round(plain_handles, 3),
                                                      "plain_fp32_batch": round(plain_batch, 3),
                                                      best_name + "_batch": round(refined_batch, 3)},
              "speedup_vs_plain_fp32_handles": round(refined_batch / plain_handles, 3),
              "speedup_vs_plain_fp32_batch": round(refined_batch / plain_batch, 3)}
    log(f"first {len(sub_arrs)} bodies: {subset}")
    plain_ok = [k for k in ("plain_fp32", "plain_fp32_200", "plain_fp32_tight") if out[k]["max_rel_err_vs_fp64"] <= 1e-5]
    plain_name = max(plain_ok, key=lambda k: out[k]["steps_per_s"]) if plain_ok else "plain_fp32_tight"
    best = out[best_name]
    line = {
        "metric": "sim steps/s", "value": best["steps_per_s"], "unit": "steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(1e3 / best["steps_per_s"], 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 inner / f64 residual" if best_name == "refined_fp32" else "bf16 off-diagonal inner / f64 residual",
        "data": "synthetic",
        "config": {"workload": f"precision form on the batch of {args.bodies} tet boxes (a, b, c ~ U{{3..24}} cells)",
                   "dt": "1/60", "newton_iterations": 1, "form": f"{best_name} (value), one batch"
                   + ("" if args.no_graph else ", captured steps"),
                   "error_sample_bodies": len(sample), "forms": out,
                   "vs_plain_fp32": {"whole_batch_steps_per_s": {"plain_fp32_batch": out["plain_fp32"]["steps_per_s"],
                                                                 best_name + "_batch": best["steps_per_s"]},
                                     "subset": subset},
                   "plain_reference_form": plain_name,
                   "speedup_vs_plain_at_1e-5": round(best["steps_per_s"] / out[plain_name]["steps_per_s"], 3),
                   "plain_meets_1e-5": bool(plain_ok)},
        "roofline": None, "cpu_baseline": None,
    }
    emit(line)


def run_batch(args):
    """Many-body batch: args.bodies uneven tet boxes in one handle (one launch sequence per step, captured as a
    HIP graph). Also times the same first args.handles_subset bodies as separate handles, and the reference
    CPU path on a sample of bodies scaled to the whole batch by tetrahedron count."""
    from bounce_softbody_amd import configs
    from bounce_softbody_amd.body import Body

    dt = configs.DT
    t0 = time.time()
    arrs = configs.batch_boxes(args.bodies)
    nv = sum(len(a["p_type"]) for a in arrs)
    nt = sum(len(a["tet_param"]) // 3 for a in arrs)
    log(f"{args.bodies} bodies built in {time.time() - t0:.1f} s: {nv} particles, {nt} tets")
    parallel = None
    if not args.no_cpu_baseline:   # before the first HIP call: the workers are forked
        psample = arrs[: 32 * min(16, len(os.sched_getaffinity(0)))]
        pst = sum(len(a["tet_param"]) // 3 for a in psample)
        pt, pw = cpu_body_parallel(psample, dt, args.sub_iterations, args.sub_tolerance)
        parallel = (pt * nt / max(1, pst), pw, len(psample), pst)
        log(f"CPU body-parallel: {len(psample)} bodies on {pw} processes, busiest {pt:.1f} s")
    t0 = time.time()
    batch = Body(arrs)
    batch.set_graph(not args.no_graph)
    if args.persist_tiles:
        from bounce_softbody_amd import _lib
        batch.set_schedule(*_lib.DEFAULT_SCHEDULE[:5], args.persist_tiles, args.persist_groups)
    rows, blocks = batch.system_size()
    log(f"batch built on the device in {time.time() - t0:.1f} s: {rows} rows, {blocks} blocks")

    def steps(bodies, k):
        for _ in range(k):
            for b in bodies:
                b.Step(dt, 1, args.sub_iterations, 0.0, args.sub_tolerance)
        for b in bodies:
            b.synchronize()

    steps([batch], args.warmup)
    start = batch.save_state() if args.autotune else None
    t0 = time.perf_counter()
    steps([batch], args.steps)
    el = time.perf_counter() - t0
    value = args.steps / el
    schedule = None
    if args.autotune:
        # the same steps again from the same state under the chosen schedule (one untimed step first)
        tuned = autotune([batch], args.schedule_cache)[0]
        batch.load_state(start)
        steps([batch], 1)
        batch.load_state(start)
        t0 = time.perf_counter()
        steps([batch], args.steps)
        tuned_value = args.steps / (time.perf_counter() - t0)
        log(f"batch with the chosen schedule: {tuned_value:.2f} steps/s (default schedule {value:.2f})")
        schedule = dict(tuned, steps_per_s={"chosen": round(tuned_value, 3), "default": round(value, 3)})
        value = tuned_value
    its = [int(s[0][2]) for s in batch.instance_stats()]
    mean_ms, by = batch.time_spmv(args.spmv_launches)
    achieved = by / (mean_ms * 1e-3) / 1e9
    log(f"batch: {value:.2f} steps/s ({value * args.bodies:.0f} body-steps/s), PCG iterations min/max "
        f"{min(its)}/{max(its)}; SpMV {mean_ms * 1e3:.1f} us, {achieved:.0f} GB/s")
    # one handle per body on a subset, against the batch of the same subset
    sub = arrs[: args.handles_subset]
    handles = [Body(a) for a in sub]
    small = Body(sub)
    small.set_graph(not args.no_graph)
    steps(handles, 2)
    steps([small], 2)
    t0 = time.perf_counter()
    steps(handles, 5)
    t_handles = (time.perf_counter() - t0) / 5
    t0 = time.perf_counter()
    steps([small], 5)
    t_small = (time.perf_counter() - t0) / 5
    log(f"{len(sub)} bodies: {1e3 * t_handles:.2f} ms/step as handles, {1e3 * t_small:.2f} ms/step as one batch")
    del handles, small
    baseline = None
    if not args.no_cpu_baseline:
        import oracle
        kind = "reference" if oracle.reference_available() else "port"
        sample = arrs[: args.cpu_sample_bodies]
        st = sum(len(a["tet_param"]) // 3 for a in sample)
        total = 0.0
        for a in sample:
            body = oracle.RefBody(a, fast=True) if kind == "reference" else oracle.OracleBody(a)
            total += body.step_timed(1, dt, 1, args.sub_iterations, 0.0, args.sub_tolerance)
            del body
        per_batch_step = total * nt / max(1, st)
        pstep, pw, pn, pst = parallel
        baseline = {"value": round(1.0 / pstep, 6), "unit": "steps/s", "cores": pw, "kind": kind,
                    "sample": f"one step of each of the first {pn} bodies ({pst} tets), one process per CPU, "
                              f"the busiest process's summed step time scaled to the batch's {nt} tets",
                    "single_thread": {"value": round(1.0 / per_batch_step, 6), "cores": 1,
                                      "sample": f"one step of each of the first {len(sample)} bodies ({st} tets, "
                                                f"{total:.1f} s of CPU time), scaled to the batch's {nt} tets"}}
        log(f"CPU baseline: {baseline}")
    workload = (f"batch of {args.bodies} tet boxes (a, b, c ~ U{{3..24}} cells), {nv} particles, {nt} tets, "
                f"{blocks} blocks")
    line = {
        "metric": "sim steps/s",
        "value": round(value, 3),
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 / value, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": workload,
                   "dt": "1/60", "newton_iterations": 1, "pcg_max_iterations": args.sub_iterations,
                   "pcg_tolerance": args.sub_tolerance, "form": "one batch" + ("" if args.no_graph else ", captured steps"),
                   "body_steps_per_s": round(value * args.bodies, 1),
                   "pcg_iterations_min_max": [min(its), max(its)],
                   "subset_ms_per_step": {"bodies": len(sub), "handles": round(1e3 * t_handles, 3),
                                          "batch": round(1e3 * t_small, 3)},
                   "schedule": schedule or ({"persist_tiles": args.persist_tiles, "persist_groups": args.persist_groups}
                                            if args.persist_tiles else "default")},
        "roofline": {"kernel": "k_cg_spmv_dq (all instances, back-to-back pass)", "bound": "hbm",
                     "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic_from_profile(PROFILE_PMC_BATCH, workload), "bytes_per_launch": by,
                     "mean_launch_us": round(mean_ms * 1e3, 2)},
        "cpu_baseline": baseline,
    }
    emit(line)


def run_contact(args):
    """Contact workload: args.bodies bodies of [cloth (args.contact_cloth+1)^2 vertices + tet args.contact_tet^3
    cells] over one 128^3 SDF, every vertex a sphere fixture, in one handle with captured steps. The bodies start
    clear of the terrain; args.pre_steps untimed steps let them settle into contact first."""
    from bounce_softbody_amd import configs
    from bounce_softbody_amd.body import Body

    dt = configs.DT
    t0 = time.time()
    arrs = configs.contact_batch(args.bodies, cloth_cells=args.contact_cloth, tet_cells=args.contact_tet,
                                 sdf_res=args.sdf_res)
    nv = sum(len(a["p_type"]) for a in arrs)
    nt = sum(len(a["tet_param"]) // 3 for a in arrs)
    ntri = sum(len(a["st_param"]) // 12 for a in arrs)
    log(f"{args.bodies} bodies built in {time.time() - t0:.1f} s: {nv} particles, {nt} tets, {ntri} cloth triangles")
    parallel = None
    if not args.no_cpu_baseline:   # before the first HIP call: the workers are forked
        psample = arrs[: 4 * min(16, len(os.sched_getaffinity(0)))]
        pt, pw = cpu_body_parallel(psample, dt, args.sub_iterations, args.sub_tolerance)
        parallel = (pt * len(arrs) / max(1, len(psample)), pw, len(psample))
        log(f"CPU body-parallel: {len(psample)} bodies on {pw} processes, busiest {pt:.1f} s")
    t0 = time.time()
    batch = Body(arrs)
    batch.set_graph(not args.no_graph)
    rows, blocks = batch.system_size()
    log(f"batch built on the device in {time.time() - t0:.1f} s: {rows} rows, {blocks} blocks")
    from bounce_softbody_amd import _lib
    solver = args.contact_solver
    if solver == "pcg":
        sub, tol = args.sub_iterations, args.sub_tolerance
    else:
        mode = _lib.SOLVER_REFINED_FP32 if solver == "refined_fp32" else _lib.SOLVER_REFINED_BF16
        batch.set_solver(mode, args.inner_iterations, args.inner_tolerance)
        sub, tol = args.max_sweeps * args.inner_iterations, args.refine_tolerance

    def steps(k):
        for _ in range(k):
            batch.Step(dt, 1, sub, 0.0, tol)
        batch.synchronize()

    steps(args.pre_steps)
    contacts_start = len(batch.contacts()[0])
    steps(args.warmup)
    t0 = time.perf_counter()
    steps(args.steps)
    el = time.perf_counter() - t0
    value = args.steps / el
    contacts_end = len(batch.contacts()[0])
    its = [int(s[0][2]) for s in batch.instance_stats()]
    mean_ms, by = batch.time_spmv(args.spmv_launches)
    achieved = by / (mean_ms * 1e-3) / 1e9
    log(f"contact batch: {value:.2f} steps/s ({value * args.bodies:.0f} body-steps/s), contacts {contacts_start} -> "
        f"{contacts_end}, PCG iterations min/max {min(its)}/{max(its)}; SpMV {mean_ms * 1e3:.1f} us, {achieved:.0f} GB/s")
    baseline = None
    if not args.no_cpu_baseline:
        import oracle
        kind = "reference" if oracle.reference_available() else "port"
        sample = arrs[: args.cpu_sample_bodies]
        total = 0.0
        for a in sample:
            body = oracle.RefBody(a, fast=True) if kind == "reference" else oracle.OracleBody(a)
            total += body.step_timed(1, dt, 1, args.sub_iterations, 0.0, args.sub_tolerance)
            del body
        per_batch_step = total * len(arrs) / max(1, len(sample))
        pstep, pw, pn = parallel
        baseline = {"value": round(1.0 / pstep, 6), "unit": "steps/s", "cores": pw, "kind": kind,
                    "sample": f"the first step of each of the first {pn} bodies, one process per CPU, the busiest "
                              f"process's summed step time scaled to the batch's {len(arrs)} equal-sized bodies",
                    "single_thread": {"value": round(1.0 / per_batch_step, 6), "cores": 1,
                                      "sample": f"the first step of each of the first {len(sample)} bodies "
                                                f"({total:.1f} s of CPU time), scaled to the batch's {len(arrs)} "
                                                f"equal-sized bodies"}}
        log(f"CPU baseline: {baseline}")
    workload = (f"contact: {args.bodies} x [cloth {args.contact_cloth + 1}^2 vertices + tet {args.contact_tet}^3 cells] "
                f"over one {args.sdf_res}^3 SDF, {nv} particles (all sphere fixtures), {nt} tets, {ntri} cloth "
                f"triangles, {blocks} blocks")
    line = {
        "metric": "sim steps/s",
        "value": round(value, 3),
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 / value, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": workload,
                   "dt": "1/60", "newton_iterations": 1, "pcg_max_iterations": sub,
                   "pcg_tolerance": tol, "pre_steps": args.pre_steps,
                   "solver": solver if solver == "pcg" else
                   f"{solver}: {args.inner_iterations} inner iterations at {args.inner_tolerance} per sweep, "
                   f"<= {args.max_sweeps} sweeps to {args.refine_tolerance}",
                   "form": "one batch" + ("" if args.no_graph else ", captured steps"),
                   "body_steps_per_s": round(value * args.bodies, 1),
                   "contacts_before_after": [contacts_start, contacts_end],
                   "pcg_iterations_min_max": [min(its), max(its)]},
        "roofline": {"kernel": "k_cg_spmv_dq (all instances, back-to-back pass)", "bound": "hbm",
                     "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4),
                     # the summary was measured on the FP32 blocks; the BF16 operator moves other bytes
                     "traffic": None if solver == "refined_bf16" else traffic_from_profile(PROFILE_PMC_CONTACT, workload),
                     "bytes_per_launch": by,
                     "mean_launch_us": round(mean_ms * 1e3, 2)},
        "cpu_baseline": baseline,
    }
    emit(line)


def run_sdf_build(args):
    """b3BuildSDF on the device: a closed icosphere (args.sdf_subdivisions subdivisions, radius 0.5) on an
    args.sdf_res^3 grid (box extension 0.2); the reference's CPU path (oracle/_ref where built, else the oracle)
    on a 24^3 grid of the same mesh, scaled by voxel count."""
    from bounce_softbody_amd import _lib, configs

    v, t = configs.icosphere(args.sdf_subdivisions, 0.5)
    cell = 1.4 / (args.sdf_res - 1)
    for _ in range(args.warmup):
        _lib.build_sdf(v, t, (cell, cell, cell), 0.2)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        lo, hi, dims, vox = _lib.build_sdf(v, t, (cell, cell, cell), 0.2)
    el = (time.perf_counter() - t0) / args.steps
    nvox = dims[0] * dims[1] * dims[2]
    pairs = nvox * len(t)
    log(f"SDF build: grid {dims}, {len(t)} triangles, {el * 1e3:.1f} ms per build, {pairs / el / 1e9:.1f} G pairs/s")
    baseline = None
    if not args.no_cpu_baseline:
        import oracle
        kind = "reference" if oracle.reference_available() else "port"
        c = 1.4 / 23
        t0 = time.perf_counter()
        _, _, d2, _ = oracle.build_sdf(v, t, (c, c, c), 0.2, which=kind if kind == "reference" else "oracle")
        cel = time.perf_counter() - t0
        n2 = d2[0] * d2[1] * d2[2]
        baseline = {"value": round(1.0 / (cel * nvox / n2), 6), "unit": "builds/s", "cores": 1, "kind": kind,
                    "sample": f"a {d2[0]}x{d2[1]}x{d2[2]} grid of the same mesh ({cel:.1f} s), scaled by voxel count"}
        log(f"CPU baseline: {baseline}")
    line = {"metric": "SDF builds/s", "value": round(1.0 / el, 4), "unit": "builds/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": round(el * 1e3, 3), "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "f32", "data": "synthetic",
            "config": {"workload": f"b3BuildSDF of an icosphere ({len(t)} triangles) on a {dims[0]}x{dims[1]}x{dims[2]} grid",
                       "voxel_triangle_pairs_per_s": round(pairs / el, 1)},
            "roofline": None, "cpu_baseline": baseline}
    emit(line)


def run_ray_cast(args):
    """b3Body::RayCastSingle on the device: the 512^2 cloth after 20 steps, args.rays rays per query, rays and
    results resident in device memory (b3g_ray_cast_device); each query after the cloth moved (tree rebuilt) and
    again on the same state (tree reused). The host-array entry point (b3g_ray_cast, PCIe copies included) is
    timed beside it and reported in config."""
    from bounce_softbody_amd import configs
    from bounce_softbody_amd.body import Body

    arr = configs.cloth_on_sphere(args.cloth).arrays()
    b = Body(arr)
    for _ in range(20):
        b.Step(configs.DT, 1, 50, 0.0, 1e-5)
    x, _ = b.state()
    rng = np.random.default_rng(9)
    lo, hi = x.min(0) - 0.1, x.max(0) + 0.1
    p1 = (lo + (hi - lo) * rng.random((args.rays, 3))).astype(np.float32)
    p1[:, 1] = hi[1] + 0.5
    p2 = p1.copy()
    p2[:, 1] = lo[1] - 0.5
    from bounce_softbody_amd._lib import DeviceArray
    d1, d2 = DeviceArray.of(p1), DeviceArray.of(p2)
    di, df, dn = DeviceArray(8 * args.rays), DeviceArray(4 * args.rays), DeviceArray(12 * args.rays)

    def cast():
        b.ray_cast_device(args.rays, d1.address, d2.address, di.address, df.address, dn.address)

    for _ in range(args.warmup):
        b.set_state(x)
        cast()
    rebuild = reuse = host = 0.0
    for _ in range(args.steps):
        b.set_state(x)   # moves the particles' version: the next query rebuilds the tree
        b.synchronize()
        t0 = time.perf_counter()
        cast()
        t1 = time.perf_counter()
        cast()
        rebuild += t1 - t0
        reuse += time.perf_counter() - t1
    hits = int((di.download(np.int64, args.rays) >= 0).sum())
    host_steps = max(1, min(args.steps, 10))
    for _ in range(host_steps):
        t0 = time.perf_counter()
        hidx, _, _ = b.RayCastSingle(p1, p2)
        host += time.perf_counter() - t0
    if int((hidx >= 0).sum()) != hits:
        raise SystemExit("host and device ray casts disagree")
    rebuild /= args.steps
    reuse /= args.steps
    host /= host_steps
    log(f"ray cast: {args.rays} rays in device memory, {1e3 * rebuild:.2f} ms with tree rebuild, "
        f"{1e3 * reuse:.2f} ms reused; host arrays {1e3 * host:.2f} ms; {hits} hits")
    line = {"metric": "rays/s", "value": round(args.rays / rebuild, 1), "unit": "rays/s", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(rebuild * 1e3, 3),
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
            "config": {"workload": f"RayCastSingle x {args.rays} against a {args.cloth + 1}^2 cloth "
                                   f"({len(arr['tf_density'])} triangle fixtures), rays and results in device "
                                   f"memory, tree rebuilt per query",
                       "ms_tree_reused": round(reuse * 1e3, 3),
                       "host_arrays_rays_per_s": round(args.rays / host, 1)},
            "roofline": None, "cpu_baseline": None}
    emit(line)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None, help="timed steps (default 100; batch_precision: 5 per form)")
    ap.add_argument("--warmup", type=int, default=None, help="untimed steps first (default 5; batch_precision: 1)")
    ap.add_argument("--tet", type=int, default=64, help="cells per axis of the tet grid")
    ap.add_argument("--cloth", type=int, default=511, help="cells per side of the cloth (vertices = cells + 1)")
    ap.add_argument("--sub-iterations", type=int, default=50)
    ap.add_argument("--sub-tolerance", type=float, default=1e-5)
    ap.add_argument("--cpu-sample-steps", type=int, default=3, help="timed CPU reference steps (median) after 1 warm-up")
    ap.add_argument("--spmv-launches", type=int, default=200, help="launches of the SpMV roofline pass")
    ap.add_argument("--cold-launches", type=int, default=50,
                    help="launches of the SpMV pass from HBM (each after a cache-evicting write)")
    ap.add_argument("--batch", action="store_true",
                    help="step the two bodies as one batch of two instances instead of two handles")
    ap.add_argument("--graph", action="store_true", help="replay each body's step as a captured HIP graph")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--workload", choices=["two_body", "precision", "batch", "batch_precision", "contact", "sdf_build",
                                           "ray_cast"],
                    default="two_body")
    ap.add_argument("--sdf-subdivisions", type=int, default=4, help="sdf_build workload: icosphere subdivisions")
    ap.add_argument("--rays", type=int, default=1000000, help="ray_cast workload: rays per query")
    ap.add_argument("--bodies", type=int, default=None, help="batch / contact workloads: number of bodies "
                                                               "(default 4096 / 256)")
    ap.add_argument("--handles-subset", type=int, default=64, help="batch workload: bodies also timed as handles")
    ap.add_argument("--cpu-sample-bodies", type=int, default=None,
                    help="batch / contact workloads: bodies in the single-thread CPU sample (default 256 / 16)")
    ap.add_argument("--no-graph", action="store_true", help="batch / contact workloads: uncaptured steps")
    ap.add_argument("--persist-tiles", type=int, default=0,
                    help="batch workload: persistent PCG for the instances of at most this many SpMV tiles (0: off)")
    ap.add_argument("--persist-groups", type=int, default=0,
                    help="batch workload: workgroups of the persistent PCG (0: one per persistent instance)")
    ap.add_argument("--autotune", action="store_true",
                    help="two_body / batch workloads: after the timed steps, autotune each body's PCG schedule and "
                         "time the steps again with it (reported as value; the default schedule's rate beside it)")
    ap.add_argument("--schedule-cache", default=None, help="autotune: schedule cache file (looked up, appended)")
    ap.add_argument("--contact-cloth", type=int, default=127, help="contact workload: cloth cells per side")
    ap.add_argument("--contact-tet", type=int, default=16, help="contact workload: tet block cells per axis")
    ap.add_argument("--contact-solver", choices=["pcg", "refined_fp32", "refined_bf16"], default="pcg",
                    help="contact workload: the reference's PCG, or FP64 refinement around an FP32 / BF16 inner PCG")
    ap.add_argument("--sdf-res", type=int, default=128, help="contact and sdf_build workloads: SDF voxels per axis")
    ap.add_argument("--pre-steps", type=int, default=None,
                    help="precision / contact workloads: untimed steps first (default 10 / 40)")
    ap.add_argument("--inner-iterations", type=int, default=None,
                    help="precision workloads: PCG iterations per sweep (default 30; contact: 50)")
    ap.add_argument("--inner-tolerance", type=float, default=None,
                    help="precision workloads: PCG tolerance per sweep (default 1e-2; contact: 1e-3)")
    ap.add_argument("--max-sweeps", type=int, default=None,
                    help="precision workloads: refinement sweeps at most (default 4; batch_precision: 10)")
    ap.add_argument("--refine-tolerance", type=float, default=None,
                    help="precision workloads: FP64 residual target relative to the first sweep's (default 1e-4; "
                         "batch_precision: 1e-5)")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: this workload is a single body per GPU and the multi-GPU path is not "
              "part of this build (run with --gpus 1)", file=sys.stderr)
        sys.exit(2)

    contact = args.workload == "contact"
    if args.bodies is None:
        args.bodies = 256 if contact else 4096
    bp = args.workload == "batch_precision"
    if args.steps is None:
        args.steps = 5 if bp else 100
    if args.warmup is None:
        args.warmup = 1 if bp else 5
    if args.cpu_sample_bodies is None:
        args.cpu_sample_bodies = 16 if contact else (8 if bp else 256)
    # precision workloads: the fastest refined settings within 1e-5 of the FP64 oracle in sweeps of inner
    # iterations / tolerances and sweep limits on the 64^3 grid and on the 4096-body batch
    if args.inner_iterations is None:
        args.inner_iterations = 50 if contact else 30
    if args.inner_tolerance is None:
        args.inner_tolerance = 1e-3 if contact else 1e-2
    if args.max_sweeps is None:
        args.max_sweeps = 10 if bp else 4
    if args.refine_tolerance is None:
        args.refine_tolerance = 1e-5 if bp else 1e-4
    if args.pre_steps is None:
        args.pre_steps = 40 if contact else 10
    if args.workload == "precision":
        run_precision(args)
        return
    if args.workload == "sdf_build":
        run_sdf_build(args)
        return
    if args.workload == "ray_cast":
        run_ray_cast(args)
        return
    if contact:
        run_contact(args)
        return
    if args.workload == "batch":
        run_batch(args)
        return
    if args.workload == "batch_precision":
        run_batch_precision(args)
        return

    from bounce_softbody_amd import configs
    from bounce_softbody_amd.body import Body

    dt = configs.DT
    t0 = time.time()
    tet_arr = configs.tet_grid(args.tet).arrays()
    cloth_arr = configs.cloth_on_sphere(args.cloth).arrays()
    log(f"scenes built in {time.time() - t0:.1f} s: tet {len(tet_arr['p_type'])} particles, "
        f"cloth {len(cloth_arr['p_type'])} particles")
    t0 = time.time()
    if not args.batch:
        tet, cloth = Body(tet_arr), Body(cloth_arr)
        bodies = [tet, cloth]
        roof = tet
    else:
        both = Body([tet_arr, cloth_arr])   # one batch: one launch sequence steps both bodies
        bodies = [both]
        roof = both
    log(f"device bodies built in {time.time() - t0:.1f} s; systems {[b.system_size() for b in bodies]}")
    if args.graph:
        for b in bodies:
            b.set_graph(True)

    def step():
        for b in bodies:
            b.Step(dt, 1, args.sub_iterations, 0.0, args.sub_tolerance)

    def sync():
        for b in bodies:
            b.synchronize()

    for _ in range(args.warmup):
        step()
    sync()
    log(f"warmup done ({args.warmup} steps)")
    start = [b.save_state() for b in bodies] if args.autotune else None

    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    sync()
    elapsed = time.perf_counter() - t0
    schedule = None
    if args.autotune:
        # the same steps again from the same states under the chosen schedules (one untimed step first)
        tuned = autotune(bodies, args.schedule_cache)
        for b, blob in zip(bodies, start):
            b.load_state(blob)
        step()
        sync()
        for b, blob in zip(bodies, start):
            b.load_state(blob)
        t0 = time.perf_counter()
        for _ in range(args.steps):
            step()
        sync()
        tuned_elapsed = time.perf_counter() - t0
        log(f"chosen schedules: {args.steps / tuned_elapsed:.2f} steps/s (default schedule {args.steps / elapsed:.2f})")
        schedule = {"bodies": tuned, "steps_per_s": {"chosen": round(args.steps / tuned_elapsed, 3),
                                                     "default": round(args.steps / elapsed, 3)}}
        elapsed = tuned_elapsed
    # roofline pass: the PCG SpMV kernel back-to-back on the system of the last step (its 147 MB matrix stays
    # in the 256 MB MALL between launches), then with the caches evicted before every launch (from HBM)
    mean_ms, bytes_per_launch = roof.time_spmv(args.spmv_launches)
    cold_ms, _ = roof.time_spmv_cold(args.cold_launches)
    if not args.batch:
        stats, _ = tet.stats()
        cstats, ccontacts = cloth.stats()
    else:
        (stats, _), (cstats, ccontacts) = both.instance_stats()
    for b in bodies:
        x, v = b.state()
        if not (np.all(np.isfinite(x)) and np.all(np.isfinite(v))):
            raise SystemExit("non-finite state after the timed steps")
    value = args.steps / elapsed
    log(f"timed {args.steps} steps in {elapsed:.3f} s; tet PCG iterations min/max {stats[1]}/{stats[2]}, "
        f"cloth {cstats[1]}/{cstats[2]}, cloth contacts {ccontacts}; SpMV pass {args.spmv_launches} launches, "
        f"mean {mean_ms*1e3:.2f} us; from HBM (caches evicted) {args.cold_launches} launches, mean {cold_ms*1e3:.2f} us")

    achieved = bytes_per_launch / (mean_ms * 1e-3) / 1e9 if mean_ms > 0 else 0.0
    achieved_hbm = bytes_per_launch / (cold_ms * 1e-3) / 1e9 if cold_ms > 0 else 0.0
    baseline = None
    if not args.no_cpu_baseline:
        log("timing the CPU baseline ...")
        baseline = cpu_baseline(tet_arr, cloth_arr, dt, args.sub_iterations, args.sub_tolerance, args.cpu_sample_steps)
        log(f"CPU baseline: {baseline}")
    out = {
        "metric": "sim steps/s",
        "value": round(value, 3),
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * elapsed / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"tet grid {args.tet}^3 cells ({len(tet_arr['p_type'])} particles, "
                        f"{len(tet_arr['tet_param']) // 3} tets) + cloth {args.cloth + 1}^2 vertices on a sphere",
            "dt": "1/60",
            "newton_iterations": 1,
            "pcg_max_iterations": args.sub_iterations,
            "pcg_tolerance": args.sub_tolerance,
            "parallelism": "single GPU",
            "form": ("two handles" if not args.batch else "one batch of two instances")
                    + (", captured steps" if args.graph else ""),
            "schedule": schedule or "default",
        },
        "roofline": {
            "kernel": "k_cg_spmv_dq (" + ("tet body" if not args.batch else "both bodies, one batched launch")
                      + ", passes after the timed steps)",
            "bound": "hbm",
            "achieved": round(achieved_hbm, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved_hbm / HBM_PEAK_GBS, 4),
            "frac_hbm": round(achieved_hbm / HBM_PEAK_GBS, 4),
            "frac_mall": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic_from_profile() if not args.batch else None,
            "bytes_per_launch": bytes_per_launch,
            "hbm_pass": f"{args.cold_launches} launches, each after a 1 GiB streaming read that evicts the L2s and "
                        f"the MALL; mean {cold_ms * 1e3:.2f} us per launch including one launch gap",
            "mall_pass": f"{args.spmv_launches} back-to-back launches, the matrix resident in the 256 MB MALL; "
                         f"mean {mean_ms * 1e3:.2f} us",
            "mean_launch_us": round(cold_ms * 1e3, 2),
            "mean_launch_us_mall": round(mean_ms * 1e3, 2),
        },
        "cpu_baseline": baseline,
    }
    emit(out)


if __name__ == "__main__":
    main()
