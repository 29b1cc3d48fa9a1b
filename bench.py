"""Benchmark of the sphere DEM step path on one MI355X.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload c2|c1|c2_dense] [--no-cpu]

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

Two host-side legs ride inside "cpu_baseline" (off with --no-cpu):
  fp64   — the GPU's FP32 step against FP64, per post-step quantity: spheres of
           the contact-dense packing (32x32x16 lattice cells) after 3 steps, and
           one rigid-clump step (13,824 clumps) integrated in float64.
  cpu_plan — the C oracle on the host: configuration 1 (65,536 spheres) for
           1,000 steps on all host threads (median of 3) and on one thread;
           serial sweep-and-prune on the C1 sphere boxes and the C4 hull boxes;
           cold GJK/EPA on the C4 pairs on one and on all threads; nproc and the
           CPU model.

Further sections ride on the same line (--sections, default
"c1,c2_dense,fp64,batch,gjk,batch_gjk,clumps,cpu_plan"):
  c1     — dem_c1 (65,536 spheres, 32-bit cell keys): particle-steps/s of 1,000
           graph-replayed steps (--steps when given) and their stages (the
           CPU-runnable configuration).
  c2_dense — dem_c2_dense: the configuration-2 law on a contact-dense FCC packing of
           the same 4,194,304 spheres (about 3 contacts per sphere): particle-steps/s
           and the contact kernel's roofline, beside the dilute headline.
  batch  — batch_c5 (512 independent scenes, 7.7M spheres): particle-steps/s of
           the batched graph-replayed step (after the autotuner picks the launch
           shapes) vs the plain form (every scene alone in its own context); then
           the 128 hull-carrying scenes' rigid bodies (544k hulls on lattices in
           walled boxes) attached, stepped inside the same graph.
  gjk    — hulls_c4 (131,072 convex hulls): GPU sweep-and-prune, then GJK/EPA
           pairs/s cold-started ("plain") and with the learned warm start
           (features + BF16 MFMA inference + warm GJK, "learned"), after training
           the network on pairs logged from plain GJK on seeded scenes; iteration
           histograms; CPU serial SAP and CPU GJK (oracle) beside them.
  batch_gjk — the 128 hull scenes of batch_c5: GJK pairs/s of one batched pass
           (batched SAP, learned warm start) vs every scene alone (plain).
  clumps — clumps_c3 (262,144 rigid clumps, LBVH broadphase): sphere-steps/s.
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

# HBM bytes per launch of the dominant kernel, from rocprofv3 FETCH_SIZE / WRITE_SIZE
# runs on MI355X (FETCH x 2.0 + WRITE x 1.0 KiB, factors from the plain copy kernel);
# the runs and the counters are in profiles/contact_kernel_c2_counters.md.
MEASURED_TRAFFIC = {
    ("dem_c2_column_4194304", "contact"): (136485.2 * 2.0 + 262144.0) * 1024.0,
    ("dem_c2_dense_4194304", "contact"): (1475726.9 * 2.0 + 717303.8) * 1024.0,
}


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
    if w.key64:
        # re-sort from the previous order: read keys and pos4, write sorted keys
        # and pos4; the permutation and the cell-start update ride on it (the
        # changed-cell lists are tiny and only cells between a changed sphere's
        # old and new cell are rewritten)
        sort_b, ranges_b, permute_b = n * 2 * (key_b + 16), 0, 0
    else:
        # one histogram read of the keys, then per pass read + write of keys and
        # values; cell starts: read every sorted key once, write the start of
        # every cell; permute: read the permutation, gather pos4, write sorted pos4
        sort_b = n * (key_b + passes * 2 * (key_b + val_b))
        ranges_b, permute_b = n * key_b + 4 * (n_cells + 1), n * ((key_b + val_b) + 2 * 16)
    return {
        "sort": sort_b,
        "ranges": ranges_b,
        "permute": permute_b,
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


def event_pair():
    import ctypes
    from stardust_amd import _lib as L
    ev = [ctypes.c_void_p(), ctypes.c_void_p()]
    for e in ev:
        L.call("sd_event_create", ctypes.byref(e))
    return ev


def elapsed_ms(ev):
    import ctypes
    from stardust_amd import _lib as L
    ms = ctypes.c_float()
    L.call("sd_event_elapsed_ms", ev[0], ev[1], ctypes.byref(ms))
    return ms.value


def timed(fn, reps, stream=None):
    """Mean ms per call of fn() over reps calls, HIP events on `stream`."""
    from stardust_amd import _lib as L
    ev = event_pair()
    L.synchronize(stream)
    L.call("sd_event_record", ev[0], stream)
    for _ in range(reps):
        fn()
    L.call("sd_event_record", ev[1], stream)
    L.synchronize(stream)
    return elapsed_ms(ev) / reps


_MODEL = {}


def trained_model(args, n_hulls):
    """The warm-start network trained on plain GJK's converged directions on seeded
    scenes of the hulls_c4 generator (trained once per run)."""
    if "model" in _MODEL:
        return _MODEL["model"], _MODEL["train"]
    from stardust_amd import hulls as H
    from stardust_amd import warmstart as WS
    t0 = time.perf_counter()
    data = WS.PairData(args.train_samples)
    seed = 0x5EED1000
    while data.n < args.train_samples and time.perf_counter() - t0 < args.train_log_seconds:
        got = data.append_scene(H.generate_hulls(seed, n_hulls))
        seed += 1
        log("gjk: logged scene %d: %d samples (total %d)" % (seed - 0x5EED1000, got, data.n))
    log_s = time.perf_counter() - t0
    model = WS.WarmStartModel(seed=11)
    t0 = time.perf_counter()
    hist = model.train(data, epochs=args.train_epochs, batch=8192, lr=3e-3, lr_final=3e-4,
                       log=lambda *a: log("gjk:", *a))
    train_s = time.perf_counter() - t0
    log("gjk: trained on %d samples (%d scenes, logged in %.1f s), %.1f s" % (data.n, seed - 0x5EED1000, log_s,
                                                                               train_s))
    train = {"samples": data.n, "epochs": args.train_epochs, "loss": [round(x, 5) for x in hist],
             "log_s": round(log_s, 2), "train_s": round(train_s, 2)}
    _MODEL["model"], _MODEL["train"] = model, train
    del data
    return model, train


def batch_gjk_section(args):
    """batch_c5 hull scenes (128 scenes, 1k-8k hulls each): narrowphase pairs/s of one
    batched pass (batched SAP, features, BF16 inference, warm GJK) vs the plain form
    (each scene alone: SAP, cold GJK)."""
    import ctypes

    import numpy as np
    from stardust_amd import _lib as L
    from stardust_amd import hulls as H
    from stardust_amd import warmstart as WS
    from stardust_amd import workloads as W

    w = W.batch_c5()
    t0 = time.perf_counter()
    sets = [w.generate_hull_scene(j) for j in range(len(w.hull_scenes))]
    log("batch_gjk: %d hull scenes, %d hulls, generated in %.1f s" % (len(sets), int(w.hull_n.sum()),
                                                                     time.perf_counter() - t0))
    model, train = trained_model(args, 131072)
    reps = args.gjk_reps

    # plain: every scene alone (AABBs + SAP with its count read-back, then cold GJK)
    plain_gjk_ms, plain_all_s, plain_pairs = 0.0, 0.0, 0
    for h in sets:
        d = H.DeviceHulls(h)
        d.aabbs()
        pairs, npairs, _ = d.sap()
        out = d.gjk(pairs, npairs)
        L.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            d.aabbs()
            d.sap(capacity=npairs)
            d.gjk(pairs, npairs, out=out)
        L.synchronize()
        plain_all_s += (time.perf_counter() - t0) / reps
        plain_gjk_ms += timed(lambda: d.gjk(pairs, npairs, out=out), reps)
        plain_pairs += npairs
    # batched: all scenes in one set
    b = H.BatchedHulls(sets)
    b.aabbs()
    pairs, npairs = b.sap_batched()
    assert npairs == plain_pairs, (npairs, plain_pairs)
    X, desc = WS.pair_features(b, pairs, npairs)
    warm = model.infer(X, npairs, pairs=pairs, quat=b.buf["quat4"])
    out_c = b.gjk(pairs, npairs)
    out_w = b.gjk(pairs, npairs, warm=warm)
    bb = b.buf
    cold_ms = timed(lambda: b.gjk(pairs, npairs, out=out_c), reps)
    feat_ms = timed(lambda: L.call("sd_mlp_features", pairs.ptr, npairs, bb["pos4"].ptr, bb["quat4"].ptr, desc.ptr,
                                   X.ptr, None), reps)
    infer_ms = timed(lambda: model.infer(X, npairs, pairs=pairs, quat=b.buf["quat4"], out=warm), reps)
    warm_ms = timed(lambda: b.gjk(pairs, npairs, warm=warm, out=out_w), reps)
    L.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        b.aabbs()
        b.sap_batched(capacity=npairs)
        L.call("sd_mlp_features", pairs.ptr, npairs, bb["pos4"].ptr, bb["quat4"].ptr, desc.ptr, X.ptr, None)
        model.infer(X, npairs, pairs=pairs, quat=b.buf["quat4"], out=warm)
        b.gjk(pairs, npairs, warm=warm, out=out_w)
    L.synchronize()
    batched_all_s = (time.perf_counter() - t0) / reps
    c = b.download_gjk(out_c, npairs)
    wv = b.download_gjk(out_w, npairs)
    learned_ms = feat_ms + infer_ms + warm_ms
    res = {"workload": "batch_c5 hull scenes", "scenes": len(sets), "hulls": int(b.n), "pairs": int(npairs),
           "plain_gjk_pairs_per_s": plain_pairs / (plain_gjk_ms / 1000.0),
           "batched_cold_gjk_pairs_per_s": npairs / (cold_ms / 1000.0),
           "batched_learned_gjk_pairs_per_s": npairs / (learned_ms / 1000.0),
           "batched_learned_over_plain": plain_gjk_ms / learned_ms,
           "ms": {"plain_gjk_sum": round(plain_gjk_ms, 4), "batched_cold_gjk": round(cold_ms, 4),
                  "batched_features": round(feat_ms, 4), "batched_inference": round(infer_ms, 4),
                  "batched_warm_gjk": round(warm_ms, 4)},
           "with_broadphase": {"plain_pairs_per_s": plain_pairs / plain_all_s,
                               "batched_learned_pairs_per_s": npairs / batched_all_s,
                               "batched_over_plain": plain_all_s / batched_all_s},
           "iters_mean": {"cold": float(c["iters"].mean()), "learned": float(wv["iters"].mean())},
           "verdicts_identical": bool(np.array_equal(c["flags"] & 1, wv["flags"] & 1)),
           "model": train}
    log("batch_gjk: plain %.3e pairs/s, batched cold %.3e, batched learned %.3e (%.1fx); with broadphase %.1fx" % (
        res["plain_gjk_pairs_per_s"], res["batched_cold_gjk_pairs_per_s"], res["batched_learned_gjk_pairs_per_s"],
        res["batched_learned_over_plain"], res["with_broadphase"]["batched_over_plain"]))
    return res


def gjk_section(args):
    """hulls_c4: SAP + GJK/EPA pairs/s, plain cold start vs learned warm start."""
    import ctypes

    import numpy as np
    from stardust_amd import _lib as L
    from stardust_amd import hulls as H
    from stardust_amd import warmstart as WS

    log("gjk: generating hulls_c4")
    t0 = time.perf_counter()
    h = H.hulls_c4()
    log("gjk: generated in %.1f s" % (time.perf_counter() - t0))
    d = H.DeviceHulls(h)
    d.aabbs()
    pairs, npairs, axis = d.sap()
    cnt, ax = ctypes.c_int(), ctypes.c_int()
    sap_ms = timed(lambda: L.call("sd_sap_pairs", d.lo4.ptr, d.hi4.ptr, d.n, pairs.ptr, npairs, ctypes.byref(cnt),
                                  ctypes.byref(ax), d.scratch.ptr, d.scratch_bytes, None), 5)
    aabb_ms = timed(d.aabbs, 5)
    log("gjk: %d SAP pairs on axis %d; aabb %.3f ms, sap %.3f ms" % (npairs, axis, aabb_ms, sap_ms))

    model, train = trained_model(args, len(h["pos4"]))

    # hull descriptors depend on the shapes only: computed once with the scene
    X, desc = WS.pair_features(d, pairs, npairs)
    warm = model.infer(X, npairs, pairs=pairs, quat=d.buf["quat4"])
    out_c = d.gjk(pairs, npairs)
    out_w = d.gjk(pairs, npairs, warm=warm)
    reps = args.gjk_reps
    b = d.buf
    cold_ms = timed(lambda: d.gjk(pairs, npairs, out=out_c), reps)
    feat_ms = timed(lambda: L.call("sd_mlp_features", pairs.ptr, npairs, b["pos4"].ptr, b["quat4"].ptr, desc.ptr,
                                   X.ptr, None), reps)
    infer_ms = timed(lambda: model.infer(X, npairs, pairs=pairs, quat=d.buf["quat4"], out=warm), reps)
    warm_ms = timed(lambda: d.gjk(pairs, npairs, warm=warm, out=out_w), reps)
    learned_ms = feat_ms + infer_ms + warm_ms
    # ablation: the same seeded warm query started from the centre difference
    # (no network): what the seeding gains without the learned direction
    hp = d_pairs_host(pairs, npairs)
    dcv = (h["pos4"][hp[:, 1], :3] - h["pos4"][hp[:, 0], :3]).astype(np.float32)
    cdir = L.DeviceBuffer.from_array(np.ascontiguousarray(dcv))
    out_s = d.gjk(pairs, npairs, warm=cdir)
    seeded_ms = timed(lambda: d.gjk(pairs, npairs, warm=cdir, out=out_s), reps)
    c = d.download_gjk(out_c, npairs)
    w = d.download_gjk(out_w, npairs)
    cs = d.download_gjk(out_s, npairs)
    L.synchronize()
    wdir = warm.download(np.float32, 3 * npairs).reshape(-1, 3)
    same_verdict = bool(np.array_equal(c["flags"] & 1, w["flags"] & 1))
    a64, b64 = w["sep"].astype(np.float64), c["sep"].astype(np.float64)
    rel = np.where(a64 == b64, 0.0, np.abs(a64 - b64) / np.maximum(np.abs(b64), 1e-300))
    dcn = dcv / np.linalg.norm(dcv, axis=1, keepdims=True)
    ang = lambda u: float(np.degrees(np.arccos(np.clip((u * c["normal"]).sum(1), -1, 1))).mean())
    hmax = int(max(c["iters"].max(), w["iters"].max())) + 1
    log("gjk: cold %.3f ms (%.2f it), learned %.3f ms = feat %.3f + infer %.3f + warm %.3f (%.2f it); "
        "centre-seeded %.3f ms (%.2f it); max rel sep diff %.2e" % (
            cold_ms, c["iters"].mean(), learned_ms, feat_ms, infer_ms, warm_ms, w["iters"].mean(), seeded_ms,
            cs["iters"].mean(), rel.max()))

    cpu = None
    if not args.no_cpu:
        from oracle import oracle as O
        lo, hi = d.download_aabbs()
        t0 = time.perf_counter()
        cpairs, _ = O.sap_pairs(lo, hi, axis)
        sap_cpu_s = time.perf_counter() - t0
        threads = host_cpu()["threads_all"]
        pr = d_pairs_host(pairs, npairs)
        pr = np.ascontiguousarray(pr[:args.cpu_gjk_pairs] if args.cpu_gjk_pairs > 0 else pr)
        t0 = time.perf_counter()
        O.gjk(h, pr, nthreads=threads)
        gjk_cpu_s = time.perf_counter() - t0
        cpu = {"sap_serial_ms": 1000 * sap_cpu_s, "sap_pairs": int(len(cpairs)), "sap_cores": 1,
               "gjk_pairs_per_s": len(pr) / gjk_cpu_s, "gjk_cores": threads,
               "gjk_sample": "%d SAP pairs of hulls_c4, cold GJK/EPA by the C oracle (OpenMP)" % len(pr)}
        log("gjk: cpu serial SAP %.1f ms, cpu GJK %.3e pairs/s (%d threads)" % (1000 * sap_cpu_s,
                                                                               cpu["gjk_pairs_per_s"], threads))
    return {
        "workload": "hulls_c4", "hulls": d.n, "vertices": d.n_verts, "sap_axis": axis, "pairs": npairs,
        "aabb_ms": round(aabb_ms, 4), "sap_ms": round(sap_ms, 4),
        "plain_pairs_per_s": npairs / (cold_ms / 1000.0), "plain_ms": round(cold_ms, 4),
        "learned_pairs_per_s": npairs / (learned_ms / 1000.0),
        "learned_ms": {"features": round(feat_ms, 4), "inference_bf16": round(infer_ms, 4),
                       "gjk_warm": round(warm_ms, 4), "total": round(learned_ms, 4)},
        "learned_over_plain": cold_ms / learned_ms,
        "mlp_inference_tflops": 19584.0 * npairs / (infer_ms / 1000.0) / 1e12,
        "iters_mean": {"plain": float(c["iters"].mean()), "learned": float(w["iters"].mean())},
        "iters_hist": {"plain": np.bincount(c["iters"], minlength=hmax).tolist(),
                       "learned": np.bincount(w["iters"], minlength=hmax).tolist()},
        "warm_used_frac": float(((w["flags"] & L.GJK_WARM_USED) != 0).mean()),
        "verdicts_identical": same_verdict, "max_rel_sep_diff": float(rel.max()) if len(rel) else 0.0,
        "sep_bit_identical_frac": float((w["sep"] == c["sep"]).mean()),
        "ablation_centre_seeded": {
            "note": "the same seeded warm query from the centre difference instead of the network",
            "pairs_per_s": npairs / (seeded_ms / 1000.0), "ms": round(seeded_ms, 4),
            "iters_mean": float(cs["iters"].mean()),
            "mean_angle_deg_to_converged_normal": {"network": ang(wdir), "centre_difference": ang(dcn)}},
        "hit_frac": float((c["flags"] & 1).mean()),
        "train": train,
        "cpu": cpu,
    }


def d_pairs_host(pairs, npairs):
    import numpy as np
    return pairs.download(np.int32, 2 * npairs).reshape(-1, 2)


def clump_section(args):
    """clumps_c3: rigid clumps with the LBVH broadphase, sphere-steps/s."""
    from stardust_amd import clumps as CL
    log("clumps: generating clumps_c3")
    t0 = time.perf_counter()
    s = CL.clumps_c3()
    log("clumps: generated in %.1f s" % (time.perf_counter() - t0))
    sc = CL.ClumpScene(CL.contact_c3())
    sc.upload(s)
    sc.step(args.clump_warmup)
    t0 = time.perf_counter()
    sc.step(args.clump_steps)  # returns after the stream is synchronised
    wall = time.perf_counter() - t0
    prof = sc.profile(args.profile_steps)
    per = {k: v / args.profile_steps for k, v in prof.items()}
    npairs = len(sc.pairs())
    sc.close()
    ms = 1000.0 * wall / args.clump_steps
    log("clumps: %.3f ms/step, stages %s" % (ms, {k: round(v, 4) for k, v in per.items()}))
    return {"workload": "clumps_c3", "bodies": sc.nb, "spheres": sc.ns, "steps": args.clump_steps,
            "ms_per_step": round(ms, 4), "sphere_steps_per_s": sc.ns / (ms / 1000.0),
            "body_steps_per_s": sc.nb / (ms / 1000.0), "stage_ms": {k: round(v, 4) for k, v in per.items()},
            "broadphase_pairs": npairs,
            "form": "graph-replayed step pairs; partners found in the LBVH inside the contact kernel"}


def batch_section(args):
    """batch_c5: 512 scenes stepped as one batch vs the same scenes stepped one by one."""
    import numpy as np
    from stardust_amd import _lib as L
    from stardust_amd import workloads as W
    from stardust_amd.scene import BatchScene, SphereScene

    w = W.batch_c5()
    log("batch: generating %s (%d spheres, %d cells)" % (w.name, w.n, w.cells()))
    t0 = time.perf_counter()
    s, scenes = w.generate()
    log("batch: generated in %.1f s" % (time.perf_counter() - t0))
    b = BatchScene(use_graph=True)
    b.upload_scenes(scenes, s["x"], s["y"], s["z"], s["r"], s["m"])
    tune = b.autotune(trial_steps=4)
    log("batch: autotuned", tune)
    b.step(args.batch_warmup)
    stream = b.stream()
    ev = event_pair()
    L.synchronize(stream)
    t0 = time.perf_counter()
    L.call("sd_event_record", ev[0], stream)
    b.step(args.batch_steps)
    L.call("sd_event_record", ev[1], stream)
    L.synchronize(stream)
    wall = time.perf_counter() - t0
    t = max(wall, elapsed_ms(ev) / 1000.0)
    batched = w.n * args.batch_steps / t
    prof = b.profile(args.profile_steps)
    per = {k: round(v / args.profile_steps, 5) for k, v in prof.items()}
    contacts = b.contact_count()
    log("batch: %.3f ms/step -> %.3e particle-steps/s; stages %s" % (1000 * t / args.batch_steps, batched, per))
    # the persistent form of the fused step kernel, same steps (results identical)
    b.set_persistent(True)
    b.step(3)
    L.synchronize(stream)
    t0 = time.perf_counter()
    L.call("sd_event_record", ev[0], stream)
    b.step(args.batch_steps)
    L.call("sd_event_record", ev[1], stream)
    L.synchronize(stream)
    tp = max(time.perf_counter() - t0, elapsed_ms(ev) / 1000.0)
    pprof = b.profile(args.profile_steps)
    b.set_persistent(False)
    persistent = {"ms_per_step": round(1000 * tp / args.batch_steps, 4),
                  "particle_steps_per_s": w.n * args.batch_steps / tp,
                  "contact_ms": round(pprof["contact"] / args.profile_steps, 5),
                  "form": "fused step kernel as resident workgroups claiming 64-slot tiles per XCD"}
    log("batch: persistent form %.3f ms/step (contact %.4f ms)" % (persistent["ms_per_step"],
                                                                  persistent["contact_ms"]))
    hulls = batch_hulls(args, w, b, t / args.batch_steps) if args.batch_hull_steps > 0 else None
    b.close()

    # plain: every scene alone in its own context (graph-replayed steps), one after the other
    tot_t, tot_work = 0.0, 0
    firsts = np.concatenate([[0], np.cumsum([d[0] for d in scenes])])
    for k, (n, g, c) in enumerate(scenes if args.plain_steps > 0 else []):
        lo, hi = int(firsts[k]), int(firsts[k + 1])
        sc = SphereScene(g, c, key64=True, use_graph=True)
        sc.upload(s["x"][lo:hi], s["y"][lo:hi], s["z"][lo:hi], s["r"][lo:hi], s["m"][lo:hi])
        sc.step(2)
        t0 = time.perf_counter()
        sc.step(args.plain_steps)  # returns after the scene's stream is synchronised
        tot_t += time.perf_counter() - t0
        tot_work += n * args.plain_steps
        sc.close()
        if k % 128 == 127:
            log("batch: plain form, %d scenes done" % (k + 1))
    plain = tot_work / tot_t if tot_t > 0 else None
    if plain:
        log("batch: plain %.3e particle-steps/s, batched/plain %.1fx" % (plain, batched / plain))
    return {"workload": w.name, "scenes": w.nscenes, "spheres": w.n, "cells": w.cells(), "steps": args.batch_steps,
            "batched_particle_steps_per_s": batched, "batched_ms_per_step": round(1000 * t / args.batch_steps, 4),
            "plain_particle_steps_per_s": plain, "plain_steps_per_scene": args.plain_steps,
            "batched_over_plain": batched / plain if plain else None, "autotune": tune, "stage_ms": per,
            "contacts_per_sphere": contacts / w.n, "persistent": persistent, "with_hull_scenes": hulls,
            "form": "batched: one graph-replayed step for all scenes; plain: each scene alone in its own "
                    "graph-replayed context, scenes one after the other"}


def batch_hulls(args, w, b, sphere_s):
    """The hull-carrying scenes of batch_c5 as rigid bodies (hulls on a 20 mm
    lattice in walled boxes, about 0.5 contacts per hull) attached to the
    batch: one graph-replayed step advances the spheres and the bodies together
    (the body step on a forked stream), its GJK queries warm-started by the
    learned network. The bodies' context is also timed alone, cold and warm,
    from the same state; the warm and cold steps' narrowphase verdicts and
    contact sets are compared."""
    import numpy as np
    from stardust_amd import _lib as L
    from stardust_amd import hullbody as HB
    t0 = time.perf_counter()
    bodies, first = w.hull_set()
    n_hulls = int((bodies["nv"] > 3).sum())
    log("batch: %d hull scenes, %d bodies (%d hulls) generated in %.1f s" % (len(first) - 1, len(bodies["pos4"]),
                                                                          n_hulls, time.perf_counter() - t0))
    model, _ = trained_model(args, 131072)
    c = L.ContactParams.make(2e5, 0.3, 0.5, w.dt, gravity=(0.0, 0.0, -9.81), walls=None, max_contacts=16)
    nsteps = args.batch_hull_steps

    def timed_alone(hs):
        L.synchronize(None)
        t0 = time.perf_counter()
        hs.step(nsteps)  # returns after its stream is synchronised
        return (time.perf_counter() - t0) / nsteps

    def stats(hs):
        st = hs.narrowphase_stats()
        st["iters_mean"] = round(st["iters"] / max(st["pairs"], 1), 3)
        return st

    cold = HB.HullBodyScene(c)
    cold.upload(bodies, first)
    cold.step(1)
    first_cold = (stats(cold), cold.contacts()[0])
    cold.step(1)
    alone_cold = timed_alone(cold)
    cold.close()
    hs = HB.HullBodyScene(c)
    hs.upload(bodies, first)
    hs.set_warm(model, warm_budget=16)
    hs.step(1)
    first_warm = (stats(hs), hs.contacts()[0])
    same = (first_cold[0]["pairs"] == first_warm[0]["pairs"] and first_cold[0]["hits"] == first_warm[0]["hits"] and
            np.array_equal(first_cold[1], first_warm[1]))
    log("batch: bodies' first step cold %s, warm %s; same pairs, verdicts and contacts: %s" % (
        first_cold[0], first_warm[0], same))
    hs.step(1)
    alone_warm = timed_alone(hs)
    ev = event_pair()
    b.attach_hulls(hs)
    b.step(2)
    bstream = b.stream()
    L.synchronize(bstream)
    t0 = time.perf_counter()
    L.call("sd_event_record", ev[0], bstream)
    b.step(nsteps)
    L.call("sd_event_record", ev[1], bstream)
    L.synchronize(bstream)
    wall = time.perf_counter() - t0
    both = max(wall, elapsed_ms(ev) / 1000.0) / nsteps
    ncont = len(hs.contacts()[0])
    last = stats(hs)
    b.attach_hulls(None)
    hs.close()
    log("batch: with hull scenes %.3f ms/step (spheres alone %.3f, bodies alone cold %.3f, warm %.3f); %d body "
        "contacts" % (1000 * both, 1000 * sphere_s, 1000 * alone_cold, 1000 * alone_warm, ncont))
    nb = len(bodies["pos4"])
    return {"hull_scenes": len(first) - 1, "bodies": nb, "hulls": n_hulls, "steps": nsteps,
            "ms_per_step": round(1000 * both, 4), "spheres_alone_ms_per_step": round(1000 * sphere_s, 4),
            "bodies_alone_ms_per_step": {"cold": round(1000 * alone_cold, 4), "warm": round(1000 * alone_warm, 4)},
            "particle_steps_per_s": w.n / both, "body_steps_per_s": nb / both,
            "body_contacts_last_step": ncont, "narrowphase_last_step": last,
            "first_step": {"cold": first_cold[0], "warm": first_warm[0], "same_pairs_verdicts_contacts": same},
            "form": "the bodies' step (learned warm start) forked onto its own stream inside the batch's captured "
                    "step graph"}


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


def sphere_run(w, steps, warmup, autotune, profile_steps, s=None):
    """Times `steps` graph-replayed steps of sphere workload w (after the
    autotuner and `warmup` steps), then profiles the stages eagerly. Returns
    the measurements and the generated arrays."""
    import ctypes

    from stardust_amd import _lib as L
    from stardust_amd.scene import SphereScene
    if s is None:
        log("generating", w.name)
        t0 = time.perf_counter()
        s = w.generate()
        log("generated in %.1f s" % (time.perf_counter() - t0))
    g, c = w.grid(), w.contact()
    sc = SphereScene(g, c, key64=w.key64, use_graph=True)
    sc.upload(s["x"], s["y"], s["z"], s["r"], s["m"])
    tune = None
    if autotune:
        # launch shapes only (sort tile, contact block); the state is unchanged
        tune = sc.autotune(trial_steps=8)
        log("autotuned", tune)
    log("uploaded", w.n, "spheres; warmup", warmup)
    sc.step(warmup)
    stream = sc.stream()
    ev0, ev1 = ctypes.c_void_p(), ctypes.c_void_p()
    L.call("sd_event_create", ctypes.byref(ev0))
    L.call("sd_event_create", ctypes.byref(ev1))
    L.synchronize(stream)
    log("timed region:", steps, "steps")
    t0 = time.perf_counter()
    L.call("sd_event_record", ev0, stream)
    sc.step(steps)  # returns after the stream is synchronised
    L.call("sd_event_record", ev1, stream)
    L.synchronize(stream)
    wall = time.perf_counter() - t0
    ms = ctypes.c_float()
    L.call("sd_event_elapsed_ms", ev0, ev1, ctypes.byref(ms))
    gpu_s = ms.value / 1000.0
    t = max(wall, gpu_s)
    value = w.n * steps / t
    log("timed: wall %.3f s, events %.3f s -> %.3e particle-steps/s" % (wall, gpu_s, value))
    contacts = sc.contact_count()
    changes, full_next = sc.cell_changes()
    diag = sc.diagnostics()
    prof = sc.profile(profile_steps)
    per_launch = {k: v / profile_steps for k, v in prof.items()}
    log("per-step stage ms:", {k: round(v, 4) for k, v in per_launch.items()})
    sbytes = stage_bytes(w, w.n, contacts / w.n, g.n_cells)
    dominant = max(per_launch, key=per_launch.get)
    achieved = sbytes[dominant] / (per_launch[dominant] / 1000.0) / 1e9
    stages = {k: {"ms": round(per_launch[k], 5), "bytes": int(sbytes[k]),
                  "GBs": round(sbytes[k] / (per_launch[k] / 1000.0) / 1e9, 1) if per_launch[k] > 0 else None}
              for k in per_launch}
    sc.close()
    r = {"value": value, "t": t, "tune": tune, "grid": g, "contacts": contacts, "changes": changes,
         "full_next": full_next, "diag": diag, "stages": stages, "dominant": dominant, "achieved": achieved,
         "sbytes": sbytes}
    return r, s


def c1_section(args):
    """dem_c1 (65,536 spheres, u32 cell keys, the full radix sort every step):
    particle-steps/s of the graph-replayed step and its stages."""
    from stardust_amd import workloads as W
    w = W.dem_c1()
    steps = args.steps if args.steps is not None else 1000
    r, _ = sphere_run(w, steps, args.warmup, not args.no_autotune, args.profile_steps)
    return {"workload": w.name, "spheres": w.n, "steps": steps, "ms_per_step": round(1000.0 * r["t"] / steps, 4),
            "particle_steps_per_s": r["value"], "contacts_per_sphere": r["contacts"] / w.n,
            "launch_shapes": r["tune"], "stages": r["stages"]}


def c2_dense_section(args):
    """dem_c2_dense: the configuration-2 law on a contact-dense packing of the same
    sphere count (about 3 contacts per sphere), particle-steps/s and the contact
    kernel's roofline."""
    from stardust_amd import workloads as W
    w = W.dem_c2_dense()
    steps = args.steps if args.steps is not None else 20
    r, _ = sphere_run(w, steps, args.warmup, not args.no_autotune, args.profile_steps)
    dom = r["dominant"]
    return {"workload": w.name, "spheres": w.n, "steps": steps, "ms_per_step": round(1000.0 * r["t"] / steps, 4),
            "particle_steps_per_s": r["value"], "contacts_per_sphere": r["contacts"] / w.n,
            "cells": r["grid"].n_cells, "launch_shapes": r["tune"], "stages": r["stages"],
            "roofline": {"bound": "hbm", "kernel": dom, "achieved": round(r["achieved"], 1), "peak": HBM_PEAK_GBS,
                         "unit": "GB/s", "frac": round(r["achieved"] / HBM_PEAK_GBS, 4),
                         "traffic": MEASURED_TRAFFIC.get((w.name, dom)), "algorithmic_bytes": r["sbytes"][dom]},
            "packing": w.describe()["packing"]}


def host_cpu():
    """nproc, the threads this process may use, and the CPU model name."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.lower().startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    avail = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    # the host threads the run may use: the process's CPU share (OMP_NUM_THREADS
    # on a GPU box, which shares its host between GPUs), else every core it sees
    share = int(os.environ.get("OMP_NUM_THREADS", avail) or avail)
    return {"nproc": os.cpu_count(), "affinity": avail, "threads_all": max(1, min(share, avail)), "model": model}


def cpu_plan_section(args):
    """The CPU baselines the plan names (the reference has no CPU path; the C
    oracle restates it): the full configuration-1 DEM step for
    --cpu-c1-steps steps on all host threads (median of --cpu-reps runs) and on
    one thread (median of --cpu-reps-1t runs), serial sweep-and-prune on the C1
    sphere boxes and the C4 hull boxes, and cold GJK/EPA on the C4 pairs on one
    thread and on all threads."""
    import numpy as np
    from oracle import oracle as O
    from stardust_amd import hulls as H
    from stardust_amd import workloads as W
    hc = host_cpu()
    thr = hc["threads_all"]
    w = W.dem_c1()
    s = w.generate()
    g, c = w.grid(), w.contact()

    def c1_run(threads, steps):
        st = O.SphereState(s["x"], s["y"], s["z"], s["r"], s["m"], max_contacts=c.max_contacts)
        t0 = time.perf_counter()
        st.step(g, c, nsteps=steps, nthreads=threads)
        return time.perf_counter() - t0

    steps = args.cpu_c1_steps
    all_t = sorted(c1_run(thr, steps) for _ in range(args.cpu_reps))
    one_t = sorted(c1_run(1, steps) for _ in range(args.cpu_reps_1t))
    med = lambda v: v[len(v) // 2]
    log("cpu plan: C1 %d steps: %d threads %.2f s, 1 thread %.2f s (median)" % (steps, thr, med(all_t), med(one_t)))
    lo = np.stack([s["x"] - s["r"], s["y"] - s["r"], s["z"] - s["r"], 0 * s["r"]], 1).astype(np.float32)
    hi = np.stack([s["x"] + s["r"], s["y"] + s["r"], s["z"] + s["r"], 0 * s["r"]], 1).astype(np.float32)
    t0 = time.perf_counter()
    p1, _ = O.sap_pairs(lo, hi)
    sap_c1 = time.perf_counter() - t0
    h = H.hulls_c4()
    hlo, hhi = O.hull_aabbs(h)
    t0 = time.perf_counter()
    p4, _ = O.sap_pairs(hlo, hhi)
    sap_c4 = time.perf_counter() - t0
    sample = p4[:args.cpu_gjk_1t_pairs]
    t0 = time.perf_counter()
    O.gjk(h, sample, nthreads=1)
    gjk_1t = len(sample) / (time.perf_counter() - t0)
    t0 = time.perf_counter()
    O.gjk(h, p4, nthreads=thr)
    gjk_all = len(p4) / (time.perf_counter() - t0)
    log("cpu plan: serial SAP C1 %.0f ms, C4 %.0f ms; GJK %.3e pairs/s (1 thread), %.3e (%d threads)" % (
        1000 * sap_c1, 1000 * sap_c4, gjk_1t, gjk_all, thr))
    return {"host": hc,
            "note": "C oracle (gcc -O3 -ffp-contract=off, OpenMP); threads_all is the host share of this process "
                    "(OMP_NUM_THREADS on a GPU box), nproc what the OS reports",
            "dem_c1": {"steps": steps, "spheres": w.n,
                       "all_threads": {"threads": thr, "runs_s": [round(x, 3) for x in all_t],
                                       "particle_steps_per_s": w.n * steps / med(all_t)},
                       "one_thread": {"threads": 1, "runs_s": [round(x, 3) for x in one_t],
                                      "particle_steps_per_s": w.n * steps / med(one_t)}},
            "sap_serial": {"c1_spheres": {"boxes": w.n, "pairs": int(len(p1)), "ms": round(1000 * sap_c1, 2)},
                           "c4_hulls": {"boxes": len(hlo), "pairs": int(len(p4)), "ms": round(1000 * sap_c4, 2)},
                           "threads": 1},
            "gjk_cold_c4": {"one_thread_pairs_per_s": gjk_1t, "one_thread_sample_pairs": int(len(sample)),
                            "all_threads_pairs_per_s": gjk_all, "all_threads": thr, "pairs": int(len(p4))}}


def fp64_error_section(args):
    """Error of the GPU's FP32 step against FP64, per post-step quantity
    (norm-relative: max |a - b| over max |b|, and the 99.9th percentile of the
    per-element error): spheres from the contact-dense packing (an FP64 step of
    the same state: FP64 contact law and FP64 integration), rigid clumps from
    the configuration-3 generator (FP64 integration of the same step's forces
    and torques)."""
    import math

    import numpy as np
    from oracle import oracle as O
    from stardust_amd import _lib as L
    from stardust_amd import clumps as CL
    from stardust_amd import workloads as W
    from stardust_amd.scene import SphereScene

    def err(a, b):
        e = np.abs(np.asarray(a, np.float64) - b) / max(float(np.abs(b).max()), 1e-300)
        return {"max": float(e.max()), "p99_9": float(np.quantile(e, 0.999))}

    w = W.dem_c2_dense(tuple(args.fp64_lattice))
    s = w.generate()
    g, c = w.grid(), w.contact()
    sc = SphereScene(g, c, key64=True, use_graph=True)
    sc.upload(s["x"], s["y"], s["z"], s["r"], s["m"])
    sc.step(3)
    got = sc.download()
    sc.close()
    st = O.SphereState(s["x"], s["y"], s["z"], s["r"], s["m"], max_contacts=c.max_contacts)
    st.step(g, c, nsteps=2, nthreads=host_cpu()["threads_all"])
    keys, vals = O.hash_keys(st.pos4[:, 0], st.pos4[:, 1], st.pos4[:, 2], g)
    bits = max(1, int(math.ceil(math.log2(g.n_cells))))
    k, v = O.sort_pairs_u32(keys, vals, 0, bits)
    ranges = O.cell_ranges_u32(k, g.n_cells)
    pos, vel, omg = (st.pos4[v].astype(np.float64), st.vel4[v].astype(np.float64), st.omg4[v].astype(np.float64))
    h64 = st.hist.astype(np.float64)
    h64[:, 3] = st.hist[:, 3].view(np.int32)
    r = O.contact_forces_f64(pos, vel, omg, st.ids[v], g, ranges, c, v.astype(np.int32), st.hcount, h64)
    grav = np.array([c.gravity[0], c.gravity[1], c.gravity[2]], np.float64)
    live = vel[:, 3] > 0
    vel[live, :3] += (r["force"][live, :3] * vel[live, 3:4] + grav) * float(c.dt)
    pos[live, :3] += vel[live, :3] * float(c.dt)
    omg[live, :3] += (r["torque"][live, :3] * omg[live, 3:4]) * float(c.dt)
    inv = np.empty(st.n, np.int64)
    inv[st.ids[v]] = np.arange(st.n)
    pos, vel, omg = pos[inv], vel[inv], omg[inv]
    spheres = {k2: err(got[k2], arr[:, col]) for k2, (arr, col) in
               {"x": (pos, 0), "y": (pos, 1), "z": (pos, 2), "vx": (vel, 0), "vy": (vel, 1), "vz": (vel, 2),
                "wx": (omg, 0), "wy": (omg, 1), "wz": (omg, 2)}.items()}
    # clumps: one step of the GPU, against float64 integration of the oracle's
    # forces and torques of that step (quaternion by the exact exponential)
    cs = CL.generate_clumps(0x5EED0003, args.fp64_clumps)
    cc = CL.contact_c3()
    ost = O.ClumpState(cs)
    scn = CL.ClumpScene(cc)
    scn.upload(cs)
    scn.step(1)
    d = scn.download()
    pairs = scn.pairs()
    scn.close()
    gp, gv, gq = d["pos4"], d["vel4"], d["quat4"]
    F, T = ost.forces(cc, pairs)
    P0, V0, Q0, L0 = (np.array(x, np.float64) for x in (ost.P, ost.V, ost.Q, ost.L))
    dt = float(cc.dt)
    gvec = np.array([cc.gravity[0], cc.gravity[1], cc.gravity[2]], np.float64)
    im = P0[:, 3:4]
    V1 = V0[:, :3] + (F[:, :3] * im + gvec) * dt * (im > 0)
    P1 = P0[:, :3] + V1 * dt * (im > 0)
    L1 = L0[:, :3] + T[:, :3] * dt * (im > 0)
    q = Q0 / np.linalg.norm(Q0, axis=1, keepdims=True)
    w0, x0, y0, z0 = q.T
    Rm = np.stack([np.stack([1 - 2 * (y0 * y0 + z0 * z0), 2 * (x0 * y0 - w0 * z0), 2 * (x0 * z0 + w0 * y0)], -1),
                   np.stack([2 * (x0 * y0 + w0 * z0), 1 - 2 * (x0 * x0 + z0 * z0), 2 * (y0 * z0 - w0 * x0)], -1),
                   np.stack([2 * (x0 * z0 - w0 * y0), 2 * (y0 * z0 + w0 * x0), 1 - 2 * (x0 * x0 + y0 * y0)], -1)], 1)
    I0, I1 = np.asarray(ost.I0, np.float64), np.asarray(ost.I1, np.float64)
    Iinv = np.stack([np.stack([I0[:, 0], I0[:, 3], I1[:, 0]], -1), np.stack([I0[:, 3], I0[:, 1], I1[:, 1]], -1),
                     np.stack([I1[:, 0], I1[:, 1], I0[:, 2]], -1)], 1)
    wv = np.einsum("nij,njk,nlk,nl->ni", Rm, Iinv, Rm, L1)
    wn = np.linalg.norm(wv, axis=1)
    half = 0.5 * wn * dt
    k = np.where(wn > 0, np.sin(half) / np.where(wn > 0, wn, 1), 0)
    dq = np.concatenate([np.cos(half)[:, None], wv * k[:, None]], 1)
    dw, dx, dy, dz = dq.T
    Q1 = np.stack([dw * w0 - (dx * x0 + dy * y0 + dz * z0), dw * x0 + w0 * dx + (dy * z0 - dz * y0),
                   dw * y0 + w0 * dy + (dz * x0 - dx * z0), dw * z0 + w0 * dz + (dx * y0 - dy * x0)], 1)
    Q1 = np.where((im > 0), Q1 / np.linalg.norm(Q1, axis=1, keepdims=True), Q0)
    clumps = {"x": err(gp[:, 0], P1[:, 0]), "vx": err(gv[:, 0], V1[:, 0]), "qw": err(gq[:, 0], Q1[:, 0]),
              "qx": err(gq[:, 1], Q1[:, 1]), "qy": err(gq[:, 2], Q1[:, 2]), "qz": err(gq[:, 3], Q1[:, 3])}
    log("fp64: spheres %s" % {k2: "%.1e" % v2["max"] for k2, v2 in spheres.items()})
    log("fp64: clumps %s" % {k2: "%.1e" % v2["max"] for k2, v2 in clumps.items()})
    return {"spheres": {"workload": w.name, "spheres": w.n, "after_steps": 3, "per_quantity": spheres,
                        "reference": "FP64 contact law and integration of the same state (oracle twin)"},
            "clumps": {"bodies": int(len(cs["pos4"])), "contact_pairs": int(len(pairs)), "per_quantity": clumps,
                       "reference": "FP64 integration (exact quaternion exponential) of the step's FP32 forces "
                                    "and torques"},
            "note": "max is set by pairs at contact onset (the overlap of a just-touching pair rounds to either "
                    "side of zero and the unclamped damping switches a finite force); p99_9 is the bulk"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="c2", choices=["c1", "c2", "c2_dense"],
                    help="headline workload: the dilute configuration-2 column (default), configuration 1, or the "
                         "contact-dense packing of the configuration-2 law")
    ap.add_argument("--profile-steps", type=int, default=20)
    ap.add_argument("--cpu-steps", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default launch shapes of the C2 context")
    ap.add_argument("--sections", default="c1,c2_dense,fp64,batch,gjk,batch_gjk,clumps,cpu_plan",
                    help="extra sections: c2_dense, fp64, batch, gjk, batch_gjk, clumps, cpu_plan (comma list, "
                         "or none)")
    ap.add_argument("--batch-steps", type=int, default=200)
    ap.add_argument("--batch-warmup", type=int, default=10)
    ap.add_argument("--plain-steps", type=int, default=20, help="steps per scene of the plain form (0: skip it)")
    ap.add_argument("--train-samples", type=int, default=10_000_000)
    ap.add_argument("--train-log-seconds", type=float, default=90.0)
    ap.add_argument("--train-epochs", type=int, default=6)
    ap.add_argument("--gjk-reps", type=int, default=5)
    ap.add_argument("--cpu-gjk-pairs", type=int, default=0, help="CPU GJK sample size (0: every SAP pair)")
    ap.add_argument("--clump-warmup", type=int, default=5)
    ap.add_argument("--clump-steps", type=int, default=50)
    ap.add_argument("--batch-hull-steps", type=int, default=10,
                    help="steps of batch_c5 with its hull scenes attached (0: skip)")
    ap.add_argument("--cpu-c1-steps", type=int, default=1000)
    ap.add_argument("--cpu-reps", type=int, default=3, help="runs of the all-threads C1 baseline (median)")
    ap.add_argument("--cpu-reps-1t", type=int, default=1, help="runs of the one-thread C1 baseline (median)")
    ap.add_argument("--cpu-gjk-1t-pairs", type=int, default=100000)
    ap.add_argument("--fp64-lattice", type=int, nargs=3, default=[32, 32, 16])
    ap.add_argument("--fp64-clumps", type=int, default=13824)
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
    w = {"c2": W.dem_c2, "c1": W.dem_c1, "c2_dense": W.dem_c2_dense}[args.workload]()
    steps = args.steps if args.steps is not None else w.steps
    r, s = sphere_run(w, steps, args.warmup, not args.no_autotune, args.profile_steps)
    value, t, tune, g = r["value"], r["t"], r["tune"], r["grid"]
    cpu = None
    if not args.no_cpu:
        threads = host_cpu()["threads_all"]
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
                   "cells": g.n_cells, "form": "plain single-scene, graph-replayed step",
                   "launch_shapes": tune if tune else "defaults (sort tile by key count, contact_block 128)"},
        "roofline": {"bound": "hbm", "kernel": r["dominant"], "achieved": round(r["achieved"], 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(r["achieved"] / HBM_PEAK_GBS, 4),
                     "traffic": MEASURED_TRAFFIC.get((w.name, r["dominant"])),
                     "algorithmic_bytes": r["sbytes"][r["dominant"]]},
        "stages": r["stages"],
        "contacts_per_sphere": r["contacts"] / w.n,
        "cell_changes_last_timed_step": {"spheres": r["changes"], "full_sort_next": r["full_next"]},
        "energy_J": {"kinetic": r["diag"]["kinetic"], "potential": r["diag"]["potential"]},
        "cpu_baseline": cpu,
    }
    sections = [x for x in args.sections.split(",") if x and x != "none"]
    if "c1" in sections and args.workload != "c1":
        out["c1"] = c1_section(args)
    if "c2_dense" in sections:
        out["c2_dense"] = c2_dense_section(args)
    # the host-side legs: the oracle runs here only as the CPU baseline and as
    # the FP64 checker of the GPU's post-step quantities
    if cpu is not None and "fp64" in sections:
        cpu["fp64_error"] = fp64_error_section(args)
    if cpu is not None and "cpu_plan" in sections:
        cpu["plan"] = cpu_plan_section(args)
    if "batch" in sections:
        out["batch"] = batch_section(args)
    if "gjk" in sections:
        out["gjk"] = gjk_section(args)
    if "batch_gjk" in sections:
        out["batch_gjk"] = batch_gjk_section(args)
    if "clumps" in sections:
        out["clumps"] = clump_section(args)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
