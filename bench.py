"""Benchmark of the crash-solver core on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--mode plain]

One bench step is one complete solve of the workload.  Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "tests"))

import numpy as np  # noqa: E402

import crushrs_amd as ca  # noqa: E402

METRIC = "batched crash solves/s + p50/p99 latency vs plain GPU; Gelem-steps/s, %HBM"
HBM_PEAK_GBS = 8000.0


def cpu_baseline_plain(model, budget_s: float = 15.0):
    """The f64 oracle (single core) on the same model, bounded sample."""
    import oracle_ffi as of
    t0 = time.perf_counter()
    n = 0
    steps = 0
    while time.perf_counter() - t0 < budget_s:
        r = of.run(model)
        steps = r.steps
        n += 1
        if n >= 3:
            break
    dt = time.perf_counter() - t0
    return {"value": n / dt, "unit": "solves/s", "cores": 1, "kind": "port",
            "sample": f"{n} full solves of the same model with the f64 scalar oracle ({steps} steps each)"}


def bench_plain(args):
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
        "cpu_baseline": cpu_baseline_plain(m) if not args.no_cpu else None,
    }
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--mode", default="plain", choices=["plain"])
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; independent instances would shard across GPUs "
                                   "with no communication (see DESIGN.md), but only --gpus 1 is measured"}))
        sys.exit(2)
    if ca.device_count() < 1:
        print(json.dumps({"error": "no HIP device"}))
        sys.exit(1)
    out = bench_plain(args)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
