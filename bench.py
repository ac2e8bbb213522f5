"""Benchmark of the crash-solver core on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W]
                    [--workload sweep|calib|large|crash|serve|pulse-fit]
                    [--instances N] [--no-cpu] [--no-plain]

Workloads
  sweep  (default) the batched sweep: N (4096) independent head-on/offset
         pairs of the four tuned vehicle TOMLs, speeds U(10, 40) mph each,
         lateral offset U(0, 0.6) m, element size 0.3 or 0.2 m, rigid road,
         0.15 s each, solved together by the batched engine (graph-replayed
         tile kernels).
         One bench step is one batched solve of every instance to its end.
  calib  the batched calibration set: N (1024) rigid-barrier tests at
         35 mph of the four TOMLs at 0.3 m, crush curves perturbed as a
         calibration driver probes them (yield x LogU(0.8, 1.25), knots x
         U(0.9, 1.1), rail box force_fraction U(0.5, 0.8), width
         U(0.4, 0.8) m).  One bench step solves every instance.
  large  the large single instance: the crash presets remeshed at 0.075 m
         (76,662 hexes, 84,416 nodes), 8-knot tabulated curves with lock-up,
         mass_scaling 0.25.  One bench step is one solve.
  crash  plain single solve: `crush crash 30 30` (Silverado vs Neon presets,
         1290 hexes, no road) with the one-model path.  One bench step is
         one solve.
  serve  Poisson arrivals of delta-v requests (head-on/offset, T-bone,
         rear-end) into the continuous-batching server, at several offered
         loads; value is the highest achieved solves/s.
  pulse-fit  the pulse calibration driver (calibrate_pulse) on NHTSA test
         2320: one step is one pattern-search round, one batch of 18
         barrier probes, timed end to end (model build, upload, solve,
         pulse scoring).

Prints one JSON line.  `value` is solves/s of the workload with every model
already resident in device memory; preparation and upload are reported
beside it, not inside it.
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

import numpy as np  # noqa: E402

import crushrs_amd as ca  # noqa: E402
from crushrs_amd import workloads  # noqa: E402

METRIC = "batched crash solves/s + p50/p99 latency vs plain GPU; Gelem-steps/s, %HBM"
HBM_PEAK_GBS = 8000.0

# Algorithmic bytes of one explicit step in the reference's storage model
# (f32 element data, f64 nodal data, geometry constants stored per element):
# force pass 776 B per honeycomb element (connectivity 32, gradients/volume/
# hourglass constants 540, state 100 read + 100 written, part/erosion 4),
# 576 B per elastic element (no state), 36 B per node (gather X and u, write
# f_int); central-difference update 76 B per node; the stable-step pass every
# adaptive_check_steps steps 36 B per element and 24 B per node.  Contact is
# not counted.
B_HC, B_EL, B_NODE_FORCE, B_NODE_UPDATE = 776, 576, 36, 76
# Bytes the honeycomb force kernel must move in this library's layout, where
# the geometry constants are stored once per geometry class and stay in
# cache: per element-step the connectivity 32, the class index 4, the
# erosion flag 1, the state 100 read + 100 written and the per-corner forces
# 96 written (333 B), plus each node of the honeycomb elements' displacement
# once (24 B, f64 xyz).  This is the roofline's algorithmic figure.
LAYOUT_HC_ELEM, LAYOUT_HC_NODE = 333, 24
# HBM bytes the honeycomb force kernel actually moves per element-step,
# from rocprofv3 FETCH_SIZE and WRITE_SIZE counters on a batched sweep,
# each calibrated against a 1 GiB device copy (FETCH_SIZE x 2.0,
# WRITE_SIZE x 1.0): 217.6 B read, 198.7 B written.
TRAFFIC_HC_READ, TRAFFIC_HC_WRITE = 217.6, 198.7
B_DT_ELEM, B_DT_NODE = 36, 24
# Worker threads of one GPU's job on a shared host: the share the GPU box
# gives one GPU (its scheduler runs one job per GPU with this many CPUs).
CPU_SHARE = 16
KERNELS = ("k_ctl", "k_grid", "k_search_tiles", "k_faceforce_tiles", "k_node_tiles",
           "k_force_hc", "k_force_el")


def step_bytes(n_hc: int, n_el: int, n_nodes: int, steps: int, check_every: int = 5) -> float:
    per = B_HC * n_hc + B_EL * n_el + (B_NODE_FORCE + B_NODE_UPDATE) * n_nodes
    dt = (B_DT_ELEM * (n_hc + n_el) + B_DT_NODE * n_nodes) * -(-steps // check_every)
    return float(per) * steps + dt


def cpu_info() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def host_cpus() -> dict:
    """What the host offers this process: logical CPUs of the machine, CPUs
    in the affinity mask, the cgroup CPU quota, physical cores and sockets,
    and the worker count the CPU baselines use (every CPU of the process's
    share: the affinity mask and quota, at most CPU_SHARE)."""
    logical = os.cpu_count() or 1
    try:
        affinity = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        affinity = logical
    quota = None
    try:
        q, per = open("/sys/fs/cgroup/cpu.max").read().split()[:2]
        if q != "max":
            quota = float(q) / float(per)
    except (OSError, ValueError):
        pass
    cores, sockets = set(), set()
    try:
        phys = core = None
        for line in open("/proc/cpuinfo"):
            if line.startswith("physical id"):
                phys = line.split(":", 1)[1].strip()
            elif line.startswith("core id"):
                core = line.split(":", 1)[1].strip()
            elif not line.strip() and phys is not None:
                sockets.add(phys)
                cores.add((phys, core))
                phys = core = None
    except OSError:
        pass
    share = affinity if quota is None else min(affinity, max(1, int(quota)))
    return {"cpu_model": cpu_info(), "logical_cpus": logical, "physical_cores": len(cores) or None,
            "sockets": len(sockets) or None, "affinity_cpus": affinity, "cgroup_quota_cpus": quota,
            "threads": max(1, min(share, CPU_SHARE))}


def cpu_baseline(models, elem_steps_per_solve: float, budget_s: float = 20.0, sample_steps: int = 100,
                 path: str = "lanes"):
    """The reference's CPU path restated in C, timed on the host cores.

    path "lanes": the default element path of the reference (Settings.simd =
    true, src/kernel_simd.rs): honeycomb elements as 8-lane f32 AVX2 blocks,
    other elements scalar f64.  path "f64": the all-scalar f64 oracle.

    Throughput form of a batch on the CPU: one solve per core, in parallel
    (as the reference's calibration drivers do).  Each thread runs
    `sample_steps` explicit steps of one sampled model at a time; the step
    rate, less the set-up time of a zero-step call, is converted to solves/s
    of this workload through its mean element-steps per solve."""
    sys.path.insert(0, str(ROOT / "tests"))
    import ctypes as C
    import oracle_ffi as of
    lib = of.load()
    host = host_cpus()
    cores = host["threads"]
    rng = np.random.default_rng(1)
    order = rng.permutation(len(models))
    t_end = time.perf_counter() + budget_s

    def steps(m, n, done):
        if path == "lanes":
            return lib.oracle_run_steps_lanes(m.ptr, n, 1, None, None, None, None, None, C.byref(done))
        return lib.oracle_run_steps(m.ptr, n, None, None, None, None, None, C.byref(done))

    def worker(w):
        busy, es, k = 0.0, 0.0, 0
        done = C.c_int32()
        for j in order[w::cores]:
            if time.perf_counter() > t_end and k > 0:
                break
            m = models[j]
            t0 = time.perf_counter()
            steps(m, 0, done)
            t1 = time.perf_counter()
            steps(m, sample_steps, done)
            t2 = time.perf_counter()
            busy += max((t2 - t1) - (t1 - t0), 1e-9)
            es += m.n_hexes * done.value
            k += 1
        return (es / busy if busy > 0.0 else 0.0), k

    with ThreadPoolExecutor(max_workers=cores) as ex:
        res = list(ex.map(worker, range(cores)))
    rate = sum(r for r, _ in res)
    n = sum(k for _, k in res)
    per_core = rate / cores
    what = ("AVX2 8-lane f32 restatement of the reference's default element path (kernel_simd.rs; "
            "elastic parts scalar f64), -O3 -mavx2 -mfma" if path == "lanes" else
            "f64 scalar restatement of the reference's solver (kernel.rs path)")
    return {
        "value": rate / elem_steps_per_solve,
        "unit": "solves/s",
        "cores": cores,
        "kind": "port",
        "sample": (f"{what}; one solve per thread, {n} sampled models x {sample_steps} steps on {cores} threads of "
                   f"{host['cpu_model']}; {rate / 1e6:.2f} M element-steps/s converted with this workload's mean "
                   f"{elem_steps_per_solve:.3g} element-steps per solve"),
        "gelem_steps_per_s": rate / 1e9,
        "host": host,
        # one solve per core scales linearly (independent solves, no shared
        # data): the rate on every physical core of the machine, for scale
        "all_cores_extrapolated": (per_core * (host["physical_cores"] or host["logical_cpus"])
                                   / elem_steps_per_solve),
    }


def plain_gpu(models, k: int = 8):
    """The one-model GPU path (crush_run, one solve at a time) on k sampled
    models of the workload."""
    rng = np.random.default_rng(2)
    idx = rng.choice(len(models), size=min(k, len(models)), replace=False)
    ca.run(models[int(idx[0])])  # load kernels
    t, es = [], 0.0
    for i in idx:
        s = time.perf_counter()
        r = ca.run(models[int(i)])
        t.append(time.perf_counter() - s)
        es += models[int(i)].n_hexes * r.steps
    tot = sum(t)
    return {"value": len(t) / tot, "unit": "solves/s", "sample": f"{len(t)} sampled models solved one at a time",
            "gelem_steps_per_s": es / tot / 1e9, "ms_per_solve_p50": float(np.median(t) * 1e3)}


def copy_probe(nbytes: int) -> None:
    """One device-to-device copy of nbytes with the HIP runtime's copy
    kernel: the reference reading for the HBM byte counters."""
    import ctypes as C
    ca._abi.load()  # loads the HIP runtime the library links
    hip = None
    for name in ("libamdhip64.so", "libamdhip64.so.7", "libamdhip64.so.6", "/opt/rocm/lib/libamdhip64.so"):
        try:
            hip = C.CDLL(name)
            break
        except OSError:
            continue
    if hip is None:
        raise RuntimeError("HIP runtime not found")
    a, b = C.c_void_p(), C.c_void_p()
    if hip.hipMalloc(C.byref(a), C.c_size_t(nbytes)) or hip.hipMalloc(C.byref(b), C.c_size_t(nbytes)):
        raise RuntimeError("hipMalloc failed")
    hip.hipMemset(a, 1, C.c_size_t(nbytes))
    hip.hipDeviceSynchronize()
    rc = hip.hipMemcpy(b, a, C.c_size_t(nbytes), 3)  # hipMemcpyDeviceToDevice
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)
    print(json.dumps({"copy_probe_bytes": nbytes, "rc": rc}))


def progress(**kw) -> None:
    """One progress line on stderr (the JSON result line is on stdout)."""
    print(json.dumps({"progress": kw}), file=sys.stderr, flush=True)


def hc_node_count(model) -> int:
    """Distinct nodes of a model's honeycomb elements."""
    v = model.view
    parts = np.ctypeslib.as_array(v.hex_part, shape=(v.n_hexes,))
    kinds = np.array([v.materials[p].model for p in range(v.n_parts)], dtype=np.int32)
    hexes = np.ctypeslib.as_array(v.hexes, shape=(v.n_hexes, 8))
    return int(np.unique(hexes[kinds[parts] == 1]).size)


def timed_runs(batch, steps: int, warmup: int, budget_s: float, label: str):
    """At most one untimed warm-up run, then whole runs of the batch, each
    bracketed by device synchronisation inside crush_batch_run, until
    `steps` runs are timed or the next one would pass `budget_s` of timed
    wall time (at least one is timed).  Returns the per-run stats."""
    for w in range(min(warmup, 1)):
        t0 = time.perf_counter()
        batch.run()
        batch.reset()
        progress(leg=label, warmup=w + 1, s=round(time.perf_counter() - t0, 2))
    out = []
    while len(out) < max(steps, 1):
        if out:
            batch.reset()
        st = batch.run()
        out.append(st)
        spent = sum(x.wall_ms for x in out) / 1e3
        progress(leg=label, timed=len(out), wall_s=round(st.wall_ms / 1e3, 3), timed_s=round(spent, 2))
        if spent + spent / len(out) > budget_s:
            break
    return out


def force_roofline(tk, layout_bytes_per_hc_step: float) -> dict:
    """Roofline of the honeycomb force kernel from a kernel-timing run: the
    kernel's summed duration (HIP events around every launch, on the
    kernels' stream) over its launches, against the bytes this layout must
    move; the reference storage model's 776 B and the counter-measured
    traffic beside it."""
    k_launches = max(1, tk.launches)
    avg_ms = tk.k_force_hc_ms / k_launches
    hc_per_launch = tk.hc_element_steps / k_launches
    alg = layout_bytes_per_hc_step * hc_per_launch
    achieved = alg / (avg_ms * 1e-3) / 1e9
    traffic = (TRAFFIC_HC_READ + TRAFFIC_HC_WRITE) * hc_per_launch
    return {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
            "kernel": "k_force_hc", "launch_ms": avg_ms, "launches": int(tk.launches),
            "element_steps_per_launch": hc_per_launch,
            "algorithmic_bytes_per_launch": alg, "bytes_per_element_step": layout_bytes_per_hc_step,
            "algorithmic_definition": (f"{LAYOUT_HC_ELEM} B per honeycomb element-step (connectivity 32, class 4, "
                                       "erosion 1, state 100 read + 100 written, corner forces 96) + 24 B per "
                                       "distinct node of the honeycomb elements (displacement), this layout's "
                                       "minimum; geometry constants are per class and stay in cache"),
            "traffic_gbs": traffic / (avg_ms * 1e-3) / 1e9,
            "reference_storage_gbs": B_HC * hc_per_launch / (avg_ms * 1e-3) / 1e9,
            "reference_storage_note": ("776 B per element-step, the reference's per-element storage (geometry "
                                       "constants 540 B per element); not moved by this layout")}


def bench_sweep(args):
    """The batched workloads: the head-on/offset sweep and the barrier
    calibration set, solved together by the batched path.  One bench step
    is one run of the whole batch to the end of every instance."""
    t0 = time.perf_counter()
    if args.workload == "calib":
        cases = workloads.calibration_set(args.instances, seed=args.seed)
        desc = (f"batched calibration set: {args.instances} rigid-barrier tests at 35 mph of the four tuned TOMLs, "
                "0.3 m hexes, yield x LogU(0.8,1.25), force-table knots x U(0.9,1.1), rail-box force_fraction "
                "U(0.5,0.8) and width U(0.4,0.8) m, 0.15 s")
        data = "synthetic: seeded rigid-barrier tests of the reference's four vehicle TOMLs with perturbed crush curves"
    else:
        cases = workloads.headon_sweep(args.instances, seed=args.seed, end_time=args.end_time)
        desc = (f"batched sweep: {args.instances} head-on/offset pairs (neon/silverado/expedition/navigator TOMLs), "
                f"speeds U(10,40) mph, offset U(0,0.6) m, h in {{0.3,0.2}} m, rigid road, {args.end_time} s")
        data = "synthetic: seeded head-on/offset pairs of the reference's four vehicle TOMLs, generated meshes"
    models = workloads.build_cases(cases)
    t_build = time.perf_counter() - t0
    n_hc = np.array([workloads.honeycomb_count(m) for m in models])
    n_hc_nodes = np.array([hc_node_count(m) for m in models])
    E = np.array([m.n_hexes for m in models])
    N = np.array([m.n_nodes for m in models])
    progress(leg="build", models=len(models), s=round(t_build, 2))
    opt = ca.batch_options(threads_per_instance=args.threads_per_instance, force_tile=args.force_tile,
                           steps_per_graph=args.steps_per_graph, autotune=int(args.autotune))
    s0 = time.perf_counter()
    batch = ca.Batch(models, opt)
    if args.autotune:
        progress(leg="autotune", s=round(time.perf_counter() - s0, 2),
                 threads_per_instance=batch.options.threads_per_instance, force_tile=batch.options.force_tile,
                 class_scales=batch.class_scales())
    if args.eager:
        batch.time_kernels(True)
    runs = timed_runs(batch, args.steps, args.warmup, args.budget_s, "batch")
    st = runs[-1]
    n_timed = len(runs)
    steps = np.array([batch.result(i, fields=False).steps for i in range(len(models))])
    n = len(models)
    es = float((E * steps).sum())
    t_run = sum(r.wall_ms for r in runs) / 1e3
    value = n * n_timed / t_run
    total_bytes = sum(step_bytes(int(n_hc[i]), int(E[i] - n_hc[i]), int(N[i]), int(steps[i])) for i in range(n))
    # per honeycomb element-step, this layout's bytes (element part plus the
    # honeycomb nodes' displacement, weighted by each instance's steps)
    hc_es = float((n_hc * steps).sum())
    layout = LAYOUT_HC_ELEM + LAYOUT_HC_NODE * float((n_hc_nodes * steps).sum()) / max(hc_es, 1.0)
    # roofline of the dominant kernel, the honeycomb element-force tiles:
    # one more run of the same batch with HIP events recorded around every
    # launch (on the stream the kernels run on) gives the summed duration
    if args.eager:
        tk = st
    else:
        batch.time_kernels(True)
        batch.reset()
        s0 = time.perf_counter()
        tk = batch.run()
        batch.time_kernels(False)
        progress(leg="kernel_timing", s=round(time.perf_counter() - s0, 2))
    el_steps = tk.element_steps - tk.hc_element_steps
    kernels = {
        "ms": {name: tk.k_ms[j] for j, name in enumerate(KERNELS)},
        "launches_each": tk.launches,
        "force_el_gbs_reference_storage": B_EL * el_steps / max(tk.k_force_el_ms, 1e-9) / 1e6,
        "whole_step_reference_storage_gbs": total_bytes / (tk.device_ms * 1e-3) / 1e9,
    }
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": n_timed,
        "warmup": min(args.warmup, 1),
        "ms_per_step": 1e3 * t_run / n_timed,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": data,
        "config": {
            "workload": desc,
            "instances": n, "seed": args.seed,
            "elements": {"min": int(E.min()), "max": int(E.max()), "mean": float(E.mean())},
            "solver_steps": {"min": int(steps.min()), "max": int(steps.max()), "mean": float(steps.mean())},
            "threads_per_instance": batch.options.threads_per_instance, "force_tile": batch.options.force_tile,
            "contact_cell_scale": batch.options.contact_cell_scale, "steps_per_graph": batch.options.steps_per_graph,
            "autotune": ({"class_scales": [{"max_elements": e, "cell_scale": cs} for e, cs in batch.class_scales()]}
                         if args.autotune else None),
        },
        "step_definition": (f"one run of all {n} instances to their end; steps requested {args.steps}, warm-up "
                            f"requested {args.warmup}; runs are timed until {args.steps} are done or the next would "
                            f"pass {args.budget_s:.0f} s of timed wall time, warm-up capped at one run"),
        "gelem_steps_per_s": es * n_timed / t_run / 1e9,
        "latency_ms": {"p50": float(np.mean([r.latency_p50_ms for r in runs])),
                       "p99": float(np.mean([r.latency_p99_ms for r in runs])),
                       "definition": "batch start to instance retirement"},
        "device_ms": float(np.mean([r.device_ms for r in runs])),
        "host": {"build_models_s": t_build, "prepare_ms": st.prepare_ms, "upload_ms": st.upload_ms,
                 "pool_gib": st.pool_bytes / 2**30},
        "roofline": force_roofline(tk, layout),
        "kernels": kernels,
    }
    if not args.no_plain:
        s0 = time.perf_counter()
        pg = plain_gpu(models)
        out["plain_gpu"] = pg
        out["speedup_vs_plain"] = value / pg["value"]
        progress(leg="plain_gpu", s=round(time.perf_counter() - s0, 2))
    if not args.no_cpu:
        s0 = time.perf_counter()
        cb = cpu_baseline(models, es / n, budget_s=15.0)
        out["cpu_baseline"] = cb
        out["speedup_vs_cpu"] = value / cb["value"]
        out["speedup_vs_cpu_all_cores_extrapolated"] = value / cb["all_cores_extrapolated"]
        progress(leg="cpu_lanes", s=round(time.perf_counter() - s0, 2))
        s0 = time.perf_counter()
        f64 = cpu_baseline(models, es / n, budget_s=8.0, path="f64")
        out["cpu_f64_oracle"] = f64
        out["speedup_vs_cpu_f64"] = value / f64["value"]
        progress(leg="cpu_f64", s=round(time.perf_counter() - s0, 2))
    else:
        out["cpu_baseline"] = None
    return out


def bench_serve(args):
    """Poisson arrivals of delta-v requests at each offered rate into one
    server; latency is measured from the scheduled arrival to the result
    (host preparation at submit included)."""
    import threading
    cases = workloads.serve_mix(args.pool, seed=args.seed)
    t0 = time.perf_counter()
    pool = workloads.build_models(cases)
    t_build = time.perf_counter() - t0
    srv = ca.Server(hbm_budget_bytes=args.hbm_budget_gib * 2**30, max_resident=args.max_resident,
                    steps_per_graph=args.steps_per_graph)
    # warm-up: code objects and first captures
    for t in [srv.submit(pool[i]) for i in range(min(8, len(pool)))]:
        srv.wait(t)
    rng = np.random.default_rng(args.seed + 1)
    levels = []
    k_next = 0
    rates = [float(x) for x in args.rates.split(",")]
    for li, rate in enumerate(rates):
        n = args.top_requests if li == len(rates) - 1 else args.requests
        arrive = np.cumsum(rng.exponential(1.0 / rate, size=n))
        idx = [(k_next + i) % len(pool) for i in range(n)]
        k_next += n
        st0 = srv.stats()
        tickets = [None] * n
        lock = threading.Lock()
        nxt = [0]
        base = time.perf_counter() + 0.05

        def submitter():
            while True:
                with lock:
                    i = nxt[0]
                    nxt[0] += 1
                if i >= n:
                    return
                delay = base + arrive[i] - time.perf_counter()
                if delay > 0:
                    time.sleep(delay)
                tickets[i] = srv.submit(pool[idx[i]])

        threads = [threading.Thread(target=submitter) for _ in range(min(8, n))]
        for th in threads:
            th.start()
        done_at = [None] * n
        remaining = set(range(n))
        while remaining:
            for i in list(remaining):
                if tickets[i] is not None and srv.poll(tickets[i]) is not None:
                    done_at[i] = time.perf_counter()
                    remaining.discard(i)
            time.sleep(0.002)
        for th in threads:
            th.join()
        st1 = srv.stats()
        lat = np.array([done_at[i] - (base + arrive[i]) for i in range(n)]) * 1e3
        span = max(done_at) - base
        es = st1.element_steps - st0.element_steps
        # completion rate while the server is loaded: completions between
        # the 10th and 90th percentile completion times (no ramp, no drain)
        tq = np.sort(np.array(done_at) - base)
        i10, i90 = int(0.1 * n), int(0.9 * n) - 1
        steady = (i90 - i10) / max(tq[i90] - tq[i10], 1e-9)
        levels.append({"offered_per_s": rate, "requests": n, "solves_per_s": n / span,
                       "steady_solves_per_s": steady,
                       "latency_ms": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99))},
                       "queue_p50_ms": st1.queue_p50_ms, "gelem_steps_per_s": es / span / 1e9,
                       "driver_ms": {"admit": st1.admit_ms - st0.admit_ms, "lists": st1.lists_ms - st0.lists_ms,
                                     "replay": st1.replay_ms - st0.replay_ms,
                                     "retire": st1.retire_ms - st0.retire_ms},
                       "replays": st1.replays - st0.replays, "captures": st1.captures - st0.captures,
                       "span_ms": span * 1e3})
        print(json.dumps({"level": levels[-1]}), file=sys.stderr, flush=True)
    srv.close()
    best = max(levels, key=lambda x: x["steady_solves_per_s"])
    pg = None if args.no_plain else plain_gpu(pool)
    return {
        "metric": METRIC,
        "value": best["steady_solves_per_s"],
        "value_definition": ("completion rate between the 10th and 90th percentile completion times at the offered "
                             "load where it is highest (the server's capacity under load); solves_per_s per level "
                             "counts from the first arrival to the last result, drain included"),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": len(levels),
        "warmup": 8,
        "ms_per_step": None,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: seeded request mix (60% head-on/offset, 25% T-bone, 15% rear-end) of the four tuned TOMLs",
        "config": {"workload": (f"serving: Poisson arrivals at {args.rates} req/s, {args.requests} requests per rate "
                                f"({args.top_requests} at the highest), "
                                "speeds U(5,45) mph, h in {0.3,0.15} m, retire at separation"),
                   "pool": len(pool), "elements": {"min": min(m.n_hexes for m in pool),
                                                   "max": max(m.n_hexes for m in pool)},
                   "steps_per_graph": args.steps_per_graph, "hbm_budget_gib": args.hbm_budget_gib},
        "levels": levels,
        "plain_gpu": pg,
        "speedup_vs_plain": None if pg is None else best["steady_solves_per_s"] / pg["value"],
        "speedup_vs_plain_with_drain": None if pg is None else best["solves_per_s"] / pg["value"],
        "host": {"build_models_s": t_build},
        "roofline": None,
        "cpu_baseline": None,
    }


def bench_large(args):
    """The large single instance (76,662 hexes at 0.075 m): the batched
    path with one instance (graph-replayed steps), the one-model path, and
    the lane CPU path with OpenMP threads inside the solve."""
    import ctypes as C
    t0 = time.perf_counter()
    m = workloads.large_pair(end_time=args.end_time)
    t_build = time.perf_counter() - t0
    n_hc = workloads.honeycomb_count(m)
    opt = ca.batch_options(threads_per_instance=args.threads_per_instance, force_tile=args.force_tile,
                           steps_per_graph=args.steps_per_graph, autotune=int(args.autotune))
    batch = ca.Batch([m], opt)
    if args.eager:
        batch.time_kernels(True)
    runs = timed_runs(batch, args.steps, args.warmup, args.budget_s, "batch")
    n_timed = len(runs)
    r = batch.result(0, fields=False)
    steps = r.steps
    t_run = sum(x.wall_ms for x in runs) / 1e3
    value = n_timed / t_run
    if args.eager:
        tk = runs[-1]
    else:
        batch.time_kernels(True)
        batch.reset()
        tk = batch.run()
        batch.time_kernels(False)
    layout = LAYOUT_HC_ELEM + LAYOUT_HC_NODE * hc_node_count(m) / max(n_hc, 1)
    es = float(m.n_hexes * steps)
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": n_timed,
        "warmup": min(args.warmup, 1),
        "ms_per_step": 1e3 * t_run / n_timed,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: the reference's crash presets remeshed, generated meshes",
        "config": {"workload": ("large single instance: Silverado vs Neon presets remeshed at 0.075 m, +-30 mph, "
                                "no road, 8-knot tabulated crush curves with lock-up, mass_scaling 0.25, "
                                f"{args.end_time} s"),
                   "elements": m.n_hexes, "nodes": m.n_nodes, "honeycomb": n_hc, "solver_steps": int(steps),
                   "force_tile": batch.options.force_tile, "steps_per_graph": batch.options.steps_per_graph},
        "gelem_steps_per_s": es * n_timed / t_run / 1e9,
        "us_per_solver_step": 1e6 * t_run / n_timed / max(steps, 1),
        "delta_v": [g.delta_v[0] for g in r.groups],
        "host": {"build_model_s": t_build},
        "roofline": force_roofline(tk, layout),
        "kernels": {"ms": {name: tk.k_ms[j] for j, name in enumerate(KERNELS)}, "launches_each": tk.launches},
    }
    if not args.no_plain:
        s0 = time.perf_counter()
        pr = ca.run(m)
        tp = time.perf_counter() - s0
        out["plain_gpu"] = {"value": 1.0 / tp, "unit": "solves/s", "steps": int(pr.steps),
                            "gelem_steps_per_s": m.n_hexes * pr.steps / tp / 1e9}
        out["speedup_vs_plain"] = value * tp
    if not args.no_cpu:
        sys.path.insert(0, str(ROOT / "tests"))
        import oracle_ffi as of
        lib = of.load()
        host = host_cpus()
        cores = host["threads"]
        done = C.c_int32()
        n_s = 40
        s0 = time.perf_counter()
        lib.oracle_run_steps_lanes(m.ptr, 0, cores, None, None, None, None, None, C.byref(done))
        s1 = time.perf_counter()
        lib.oracle_run_steps_lanes(m.ptr, n_s, cores, None, None, None, None, None, C.byref(done))
        s2 = time.perf_counter()
        per_step = max((s2 - s1) - (s1 - s0), 1e-9) / max(done.value, 1)
        out["cpu_baseline"] = {
            "value": 1.0 / (per_step * steps), "unit": "solves/s", "cores": cores, "kind": "port",
            "sample": (f"AVX2 8-lane f32 restatement of the reference's default element path, {done.value} steps of "
                       f"this model with {cores} OpenMP threads in the element pass (the rest of the step "
                       f"sequential, as in the reference) on {host['cpu_model']}; {per_step * 1e3:.1f} ms per step, "
                       f"set-up ({s1 - s0:.1f} s) excluded, scaled to the GPU's {steps} steps"),
            "gelem_steps_per_s": m.n_hexes / per_step / 1e9,
            "host": host,
        }
        out["speedup_vs_cpu"] = value / out["cpu_baseline"]["value"]
    else:
        out["cpu_baseline"] = None
    return out


def bench_crash(args):
    m = ca.Model.crash(30.0, 30.0)
    for _ in range(args.warmup):
        ca.run(m)
    times = []
    res = None
    t0 = time.perf_counter()
    for _ in range(args.steps):
        s = time.perf_counter()
        res = ca.run(m)
        times.append(time.perf_counter() - s)
    wall = time.perf_counter() - t0
    elem_steps = m.n_hexes * res.steps
    out = {
        "metric": METRIC,
        "value": args.steps / wall,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * wall / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: generated vehicle blocks (reference presets), no external data",
        "config": {"workload": "plain single solve: crash 30 30 (Silverado vs Neon presets, 1290 hexes, no road)",
                   "elements": m.n_hexes, "solver_steps": int(res.steps)},
        "gelem_steps_per_s": elem_steps * args.steps / wall / 1e9,
        "latency_ms": {"p50": float(np.percentile(times, 50) * 1e3), "p99": float(np.percentile(times, 99) * 1e3)},
        "roofline": None,
        "cpu_baseline": None if args.no_cpu else cpu_baseline([m], float(elem_steps), budget_s=10.0),
    }
    return out


def bench_pulse_fit(args):
    """The pulse calibration driver as a batched-solver client: rounds of
    calibrate_pulse's pattern search on NHTSA test 2320 (the pulse-tuned
    Neon from its measured-curve start, 35 mph rigid barrier, 0.15 s), each
    round one batch of 18 barrier probes.  Timed end to end per round: the
    host builds the probe models, the batch uploads and solves them, the
    pulses come back and are scored."""
    from crushrs_amd import calibration as cal
    from crushrs_amd import signal as sg
    meas = sg.Pulse.read_nhtsa_tsv(ROOT / "tests" / "golden" / "nhtsa" / "v02320tsv.078", True)
    speed = 35.0 * ca.MPH
    v = ca.vehicle("neon-pulse")
    v.n_force_table = 0
    timing = {"build": 0.0, "upload": 0.0, "solve": 0.0, "results": 0.0}
    solves = [0]

    def timed_solve(models):
        t0 = time.perf_counter()
        b = ca.Batch(models, steps_per_graph=args.steps_per_graph)
        t1 = time.perf_counter()
        b.run()
        t2 = time.perf_counter()
        out = [b.result(i, fields=False) for i in range(len(models))]
        t3 = time.perf_counter()
        timing["upload"] += t1 - t0
        timing["solve"] += t2 - t1
        timing["results"] += t3 - t2
        solves[0] += len(models)
        return out

    # warm-up: the first round (start + probes) loads kernels
    cal.calibrate_pulse(v, meas, speed, rounds=max(args.warmup, 1))
    for k in timing:
        timing[k] = 0.0
    solves[0] = 0
    t0 = time.perf_counter()
    fit = cal.calibrate_pulse(v, meas, speed, rounds=args.steps, solve=timed_solve)
    wall = time.perf_counter() - t0
    value = solves[0] / wall
    # one round's probes of the fitted vehicle (each parameter up and down)
    probe_models = [ca.Model.barrier(cal.pulse_candidate(fit.vehicle, p, f) or fit.vehicle, speed=speed)
                    for p in range(cal.N_KNOTS + 2) for f in (1.1, 1.0 / 1.1)]
    steps = ca.run(probe_models[0]).steps
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * wall / fit.rounds,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "measured: NHTSA test 2320 rear-seat X pulse (the reference's test record); generated barrier models",
        "config": {"workload": ("pulse calibration driver: calibrate_pulse pattern search on NHTSA 2320, pulse-tuned "
                                "Neon from its measured-curve start, 35 mph rigid barrier, 0.15 s; one step = one "
                                "round = one batch of 18 probes (the first round also solves the start)"),
                   "rounds": fit.rounds, "solves": solves[0], "elements": probe_models[0].n_hexes,
                   "solver_steps": int(steps)},
        "round_ms": {k: 1e3 * t / fit.rounds for k, t in timing.items()},
        "objective": {"start": fit.start_objective, "end": fit.objective},
        "roofline": None,
    }
    out["round_ms"]["host_search"] = 1e3 * wall / fit.rounds - sum(out["round_ms"].values())
    if not args.no_plain:
        s0 = time.perf_counter()
        for m in probe_models:
            ca.run(m)
        tp = (time.perf_counter() - s0) / len(probe_models)
        out["plain_gpu"] = {"value": 1.0 / tp, "unit": "solves/s",
                            "sample": f"{len(probe_models)} probe models solved one at a time"}
        out["speedup_vs_plain"] = value * tp
    if not args.no_cpu:
        es = float(probe_models[0].n_hexes * steps)
        out["cpu_baseline"] = cpu_baseline(probe_models, es, budget_s=15.0)
        out["speedup_vs_cpu"] = value / out["cpu_baseline"]["value"]
    else:
        out["cpu_baseline"] = None
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=None)
    ap.add_argument("--workload", default="sweep", choices=["sweep", "calib", "large", "crash", "serve", "pulse-fit"])
    ap.add_argument("--rates", default="10,50,200,800,2000", help="serve: offered loads, requests/s")
    ap.add_argument("--requests", type=int, default=400, help="serve: requests per offered load")
    ap.add_argument("--top-requests", type=int, default=4096, help="serve: requests at the highest offered load")
    ap.add_argument("--pool", type=int, default=256, help="serve: distinct request models")
    ap.add_argument("--hbm-budget-gib", type=float, default=32.0, help="serve: device memory for resident requests")
    ap.add_argument("--max-resident", type=int, default=4096, help="serve: requests on the device at once")
    ap.add_argument("--instances", type=int, default=None, help="sweep: 4096, calib: 1024")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--threads-per-instance", type=int, default=256)
    ap.add_argument("--force-tile", type=int, default=256)
    ap.add_argument("--steps-per-graph", type=int, default=None)
    ap.add_argument("--autotune", action=argparse.BooleanOptionalAction, default=None,
                    help="time the launch shapes and per-size-class grid scales before the run (default: on for "
                         "sweep and calib, off with --eager, whose profiler runs cannot follow graph launches)")
    ap.add_argument("--budget-s", type=float, default=90.0,
                    help="batched workloads: timed runs stop once the next would pass this much timed wall time")
    ap.add_argument("--end-time", type=float, default=0.15,
                    help="simulated time per instance (0.15 s; shorter only for profiling runs)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="BYTES",
                    help="only run a plain device-to-device copy of BYTES (HBM counter calibration)")
    ap.add_argument("--eager", action="store_true",
                    help="launch every step directly with kernel timing (no hipGraph replay), e.g. under a "
                         "kernel-trace profiler")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    args = ap.parse_args()
    if args.autotune is None:
        args.autotune = args.workload in ("sweep", "calib") and not args.eager
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; independent instances would shard across GPUs "
                                   "with no communication, but only --gpus 1 is built and measured"}))
        sys.exit(2)
    if ca.device_count() < 1:
        print(json.dumps({"error": "no HIP device"}))
        sys.exit(1)
    if args.copy_probe:
        copy_probe(args.copy_probe)
        return
    if args.workload == "serve":
        args.steps_per_graph = 16 if args.steps_per_graph is None else args.steps_per_graph
        out = bench_serve(args)
        print(json.dumps(out), flush=True)
        return
    args.steps_per_graph = 32 if args.steps_per_graph is None else args.steps_per_graph
    if args.workload in ("sweep", "calib"):
        if args.instances is None:
            args.instances = 4096 if args.workload == "sweep" else 1024
        args.steps = 1 if args.steps is None else args.steps
        args.warmup = 1 if args.warmup is None else args.warmup
        out = bench_sweep(args)
    elif args.workload == "large":
        args.steps = 3 if args.steps is None else args.steps
        args.warmup = 1 if args.warmup is None else args.warmup
        out = bench_large(args)
    elif args.workload == "pulse-fit":
        args.steps = 4 if args.steps is None else args.steps
        args.warmup = 1 if args.warmup is None else args.warmup
        out = bench_pulse_fit(args)
    else:
        args.steps = 20 if args.steps is None else args.steps
        args.warmup = 2 if args.warmup is None else args.warmup
        out = bench_crash(args)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
