"""Benchmark of the sphere DEM step path on one MI355X.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload c2|c1] [--no-cpu]

Workload (default): dem_c2 — 4,194,304 polydisperse spheres in a walled column,
packed 64-bit (cell, index) keys sorted every step, dt = 5e-6 s (see
stardust_amd/workloads.py). Inputs are generated on the host, uploaded once, and
are resident in HBM before the timed region; the timed region replays the
captured step graph K times on the scene's stream, bracketed by device
synchronisation and HIP events.

Prints ONE JSON line with the throughput (particle-steps/s), the roofline of the
dominant kernel (its time measured with HIP events on the scene stream in an
eager profiling pass right after the timed region) and a CPU baseline (the
oracle's full step on the host cores, bounded sample of the same workload).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "particle-steps/s and GJK pairs/s, batched+learned vs plain, 1×MI355X"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def stage_bytes(w, n, contacts_per_sphere, n_cells):
    """Algorithmic (compulsory) HBM bytes per launch of each stage."""
    import math
    key_b = 8 if w.key64 else 4
    val_b = 0 if w.key64 else 4
    bits = max(1, int(math.ceil(math.log2(n_cells))))
    passes = (bits + 7) // 8
    c = contacts_per_sphere
    return {
        # one histogram read of the keys, then per pass read + write of keys (and values)
        "sort": n * (key_b + passes * 2 * (key_b + val_b)),
        # read every sorted key once, clear + write start/end of every cell
        "ranges": n * key_b + 8 * n_cells,
        # read the permutation (key/value), gather pos4, write sorted pos4
        "permute": n * ((key_b + val_b) + 2 * 16),
        # own state 52 B + old slot + history count in/out 8 B + state out 52 B + next key;
        # per contact: old history row 16 B read + new row 16 B written
        "contact": n * (52 + (key_b + val_b) + 8 + 52 + (key_b + val_b)) + n * c * 32,
    }


def cpu_baseline(w, s, steps, threads):
    from oracle import oracle as O
    g, c = w.grid(), w.contact()
    st = O.SphereState(s["x"], s["y"], s["z"], s["r"], s["m"], max_contacts=c.max_contacts)
    t0 = time.perf_counter()
    st.step(g, c, nsteps=steps, nthreads=threads)
    dt = time.perf_counter() - t0
    return w.n * steps / dt, dt


def calibrate_copy(nbytes=256 << 20, reps=5):
    """Runs k_copy_u32 over nbytes: each dispatch reads nbytes and writes nbytes."""
    import ctypes
    from stardust_amd import _lib as L
    L.require_gpu()
    a, b = L.DeviceBuffer(nbytes), L.DeviceBuffer(nbytes)
    L.call("sd_memset_u32", a.ptr, 1, nbytes // 4, None)
    for _ in range(reps):
        L.call("sd_memcpy_dtod", b.ptr, a.ptr, nbytes, None)
    L.synchronize()
    ev0, ev1, ms = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_float()
    L.call("sd_event_create", ctypes.byref(ev0))
    L.call("sd_event_create", ctypes.byref(ev1))
    L.call("sd_event_record", ev0, None)
    L.call("sd_memcpy_dtod", b.ptr, a.ptr, nbytes, None)
    L.call("sd_event_record", ev1, None)
    L.call("sd_event_elapsed_ms", ev0, ev1, ctypes.byref(ms))
    print(json.dumps({"copy_bytes_read": nbytes, "copy_bytes_written": nbytes, "dispatches": reps + 1,
                      "copy_GBs": 2 * nbytes / (ms.value / 1000.0) / 1e9}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="c2", choices=["c1", "c2"])
    ap.add_argument("--profile-steps", type=int, default=20)
    ap.add_argument("--cpu-steps", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run the plain copy kernel on 256 MiB (reference for HBM byte counters)")
    args = ap.parse_args()
    if args.calibrate_copy:
        calibrate_copy()
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: every box here has one GPU and the path has no multi-GPU mode "
              "(--gpus %d requested)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    from stardust_amd import _lib as L
    from stardust_amd import workloads as W
    from stardust_amd.scene import SphereScene

    L.require_gpu()
    w = W.dem_c2() if args.workload == "c2" else W.dem_c1()
    steps = args.steps if args.steps is not None else w.steps
    log("generating", w.name)
    t0 = time.perf_counter()
    s = w.generate()
    log("generated in %.1f s" % (time.perf_counter() - t0))
    g, c = w.grid(), w.contact()
    sc = SphereScene(g, c, key64=w.key64, use_graph=True)
    sc.upload(s["x"], s["y"], s["z"], s["r"], s["m"])
    log("uploaded", w.n, "spheres; warmup", args.warmup)
    sc.step(args.warmup)
    stream = sc.stream()
    ev0, ev1 = __import__("ctypes").c_void_p(), __import__("ctypes").c_void_p()
    L.call("sd_event_create", __import__("ctypes").byref(ev0))
    L.call("sd_event_create", __import__("ctypes").byref(ev1))
    L.synchronize(stream)
    log("timed region:", steps, "steps")
    t0 = time.perf_counter()
    L.call("sd_event_record", ev0, stream)
    sc.step(steps)  # returns after the stream is synchronised
    L.call("sd_event_record", ev1, stream)
    L.synchronize(stream)
    wall = time.perf_counter() - t0
    ms = __import__("ctypes").c_float()
    L.call("sd_event_elapsed_ms", ev0, ev1, __import__("ctypes").byref(ms))
    gpu_s = ms.value / 1000.0
    t = max(wall, gpu_s)
    value = w.n * steps / t
    log("timed: wall %.3f s, events %.3f s -> %.3e particle-steps/s" % (wall, gpu_s, value))
    contacts = sc.contact_count()
    diag = sc.diagnostics()

    prof = sc.profile(args.profile_steps)
    per_launch = {k: v / args.profile_steps for k, v in prof.items()}
    log("per-step stage ms:", {k: round(v, 4) for k, v in per_launch.items()})
    sbytes = stage_bytes(w, w.n, contacts / w.n, g.n_cells)
    dominant = max(per_launch, key=per_launch.get)
    achieved = sbytes[dominant] / (per_launch[dominant] / 1000.0) / 1e9
    stages = {k: {"ms": round(per_launch[k], 5), "bytes": int(sbytes[k]),
                  "GBs": round(sbytes[k] / (per_launch[k] / 1000.0) / 1e9, 1)} for k in per_launch}
    sc.close()

    cpu = None
    if not args.no_cpu:
        threads = min(os.cpu_count() or 1, 16)
        log("cpu baseline: oracle full step, %d steps, %d threads" % (args.cpu_steps, threads))
        v_cpu, dt_cpu = cpu_baseline(w, s, args.cpu_steps, threads)
        cpu = {"value": v_cpu, "unit": "particle-steps/s", "cores": threads, "kind": "port",
               "sample": "%d full steps of %s (%d spheres) by the C oracle (keys, stable LSD sort, ranges, permute, "
                         "contact gather, integrate), OpenMP over spheres; %.1f s" % (args.cpu_steps, w.name, w.n,
                                                                                       dt_cpu)}
        log("cpu: %.3e particle-steps/s" % v_cpu)

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": 1000.0 * t / steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded SplitMix64 scene, generated on the host)",
        "config": {"workload": w.name, "spheres": w.n, "steps_per_run": w.steps, "dt": w.dt,
                   "keys": "u64 (cell<<32|index)" if w.key64 else "u32 cell + u32 index",
                   "cells": g.n_cells, "form": "plain single-scene, graph-replayed step"},
        "roofline": {"bound": "hbm", "kernel": dominant, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None},
        "stages": stages,
        "contacts_per_sphere": contacts / w.n,
        "energy_J": {"kinetic": diag["kinetic"], "potential": diag["potential"]},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
