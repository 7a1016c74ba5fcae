#!/usr/bin/env python3
"""PTP geodesics benchmark on one MI355X.

Workload: 1024 x 1024 grid height field (1,048,576 vertices, 2,093,058 faces,
seed 0, 8 sinusoids, random diagonals), one source drawn from the seed.  One
step = one complete distance field computed from device-resident mesh tables
(gp_plan_run); the mesh build, toplesets, orderings and upload happen before
the timed region, as the reference harness computes toplesets outside its
timer.  Times are device times (hipEvent around the launches).

Rows:
  * plain_ref_f32 / _f64: the reference window schedule (K/2 freeze,
    PTP_TOL 1e-3, 2*|limits| cap) -- parity with the FP64 oracle of the same
    schedule; it stops ~1e-2 short of the relaxation's fixed point;
  * plain_uncapped_f32: the reference's uncapped loop at tol 1e-6, the plain
    ring schedule that reaches the fixed point within 1e-4 ("plain at bar");
  * band_euclid / band_learned: bands of the Euclidean or the learned
    distance prediction (width picked by the library's band-width tuner from
    0.45-1.0 mean edge lengths) in place of the rings in the same uncapped
    loop at tol 1e-6, then check sweeps (band_learned when a trained model
    exists);
  * patches: the patch schedule (LDS-iterated patches, active list) at tol 1e-6.
value = ms per field of band_learned when it meets the 1e-4 bar against the
FP64 fixed point, else of the fastest row that does; the JSON carries every
row with its errors against both FP64 references, the end-to-end time of a
reference-schedule solve on a fresh mesh handle (plan with host- or
device-built rings and tables, upload, run, result download), the 16-source
field of the same mesh, and (unless --no-c3c5 / --no-next) the batched
256-mesh launch (c3), the precision stress grid (c5) and the rows around the
path (next: farthest-point sampling, cluster labels, iteration-error loop,
mesh readers).
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

B_RELAX_F32 = 72   # bytes per vertex relaxation: 2*4 dist + 3*4 xyz + 4 EVT + 8*(3F/V) VT/OT
B_RELAX_F64 = 92
PEAK_HBM_GBS = 8000.0
ERR_BAR = 1e-4
# HBM bytes per dispatch of the headline kernel on the default workload, from
# two rocprofv3 --pmc passes (FETCH_SIZE, WRITE_SIZE in KiB per dispatch)
# scaled by what the same counters read for a 256 MiB device-to-device copy
# (FETCH_SIZE x 2.00, WRITE_SIZE x 1.00); see
# profiles/k_sweep_f32_height_field_1M.md.  Keyed by (row, grid side, sources):
# (bytes per dispatch, how they were measured).
MEASURED_TRAFFIC = {("band_learned", 1024, 1): ((8263.80 * 2.0 + 766.20 * 1.0) * 1024,
                                                "rocprofv3 FETCH_SIZE x 2.00 + WRITE_SIZE x 1.00 per dispatch, "
                                                "mean of 6,640 dispatches of band_learned (0.5-edge bands, 8-byte "
                                                "FP32 slot geometry, contiguous XCD eighths), kernels launched one "
                                                "by one under the profiler")}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def algorithmic_bytes(st, f64=False):
    return st["relaxations"] * (B_RELAX_F64 if f64 else B_RELAX_F32) + st["cond_checks"] * 2 * (8 if f64 else 4)


def rel_err(a, b):
    m = np.isfinite(b) & (b > 0)
    r = np.abs(a[m] - b[m]) / b[m]
    return float(r.max()), float(r.mean())


def time_plan(plan, steps, warmup):
    for _ in range(warmup):
        plan.run()
    kms, walls, st = [], [], None
    for _ in range(steps):
        t0 = time.perf_counter()
        st = plan.run()
        walls.append((time.perf_counter() - t0) * 1e3)
        kms.append(st["kernel_ms"])
    return dict(st=st, kernel_ms=float(np.mean(kms)), kernel_ms_min=float(np.min(kms)),
                wall_ms=float(np.mean(walls)), dist=plan.result())


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=1024, help="grid side")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--sources", type=int, default=1, help="sources of the one field (1 or 16 in the workload)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline and the oracle error checks")
    ap.add_argument("--cpu-budget", type=float, default=30.0, help="seconds of CPU baseline work")
    ap.add_argument("--model", default=os.path.join(ROOT, "gproshan_amd", "models", "band_mlp.npy"))
    ap.add_argument("--band-width", type=float, default=0.6,
                    help="band width in mean edge lengths (0.6: measured with margin above 0.45, where the "
                         "1M field stops meeting the 1e-4 bar)")
    ap.add_argument("--local-iters", type=int, default=4, help="patch schedule: iterations per visit")
    ap.add_argument("--patch", type=int, default=512, help="patch schedule: patch size (256 or 512)")
    ap.add_argument("--stress", action="store_true", help="precision stress case instead (2048^2 1:100 grid, 4 sources)")
    ap.add_argument("--batch", action="store_true", help="batched case instead (256 meshes, one launch)")
    ap.add_argument("--only", default="", help="comma-separated row names to run (default: all; with --next: "
                                                   "fps, clusters, iter_error, readers)")
    ap.add_argument("--next", action="store_true",
                    help="rows around the hot path instead (FPS, clusters, iteration error, mesh readers)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a device-to-device copy of 256 MiB five times (to scale FETCH_SIZE/WRITE_SIZE)")
    ap.add_argument("--no-c3c5", action="store_true",
                    help="leave out the batched (C3) and precision-stress (C5) results of the default line")
    ap.add_argument("--no-next", action="store_true", help="leave the rows around the path (FPS, clusters, "
                    "iteration-error loop, readers) out of the default line")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1"}))
        sys.exit(2)
    if args.copy_calibration:
        import torch
        n = 64 << 20                                  # 64 Mi floats = 256 MiB each way
        x = torch.ones(n, device="cuda:0")
        y = torch.empty_like(x)
        for _ in range(5):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": 4 * n, "copy_bytes_written": 4 * n, "launches": 5}), flush=True)
        return

    from gproshan_amd import _lib as L
    from gproshan_amd import geodesics as P
    from gproshan_amd import meshgen as G
    if args.stress:
        return stress(args, L, P, G)
    if args.batch:
        return batch(args, L, P, G)
    if args.next:
        return next_rows(args, L, P, G)

    from gproshan_amd import workloads as W
    t = time.time()
    xyz, faces, src = W.c2(args.seed, args.sources, args.n)
    mesh = P.che(xyz, faces)
    log(f"[bench] mesh V={mesh.n_vertices()} F={mesh.n_faces()} source={src.tolist()} ({time.time() - t:.1f}s)")

    net = None
    if os.path.exists(args.model):
        from gproshan_amd import learned as Lr
        net = Lr.MLP(params=np.load(args.model, allow_pickle=False))

    rows = {}
    patch_kw = dict(block=args.patch, local_iters=args.local_iters, tol=1e-6)
    band_kw = dict(band_width=args.band_width, tol=1e-6)
    configs = [
        ("plain_ref_f32", dict(precision=L.PREC_F32)),
        ("plain_ref_f64", dict(precision=L.PREC_F64)),
        ("plain_uncapped_f32", dict(precision=L.PREC_F32, tol=1e-6, uncapped=True)),
        ("band_euclid", dict(schedule=L.SCHED_EUCLID_BAND, **band_kw)),
        ("patches", dict(schedule=L.SCHED_PATCHES, **patch_kw)),
    ]
    only = set(filter(None, args.only.split(",")))
    tuned = {}
    for name, kw in configs:
        if only and name not in only:
            continue
        t = time.time()
        if name == "band_euclid":
            # Bands of the Euclidean distance, width and launch shape tuned.
            e = np.min(np.linalg.norm(xyz[:, None, :] - xyz[src][None, :, :], axis=2), axis=1)
            bkw = {k: v for k, v in kw.items() if k != "band_width"}
            plan, w = P.Plan.tuned_bands(mesh, src, e, **bkw)
            tuned[name] = {"band_width": w, **plan.autotune()}
        else:
            plan = P.Plan(mesh, src, **kw)
            if name != "patches":
                tuned[name] = plan.autotune()
        setup_s = time.time() - t
        rows[name] = time_plan(plan, args.steps, args.warmup)
        rows[name]["setup_s"] = setup_s
        plan.close()
        st = rows[name]["st"]
        log(f"[bench] {name}: {rows[name]['kernel_ms']:.3f} ms/field, sweeps {st['sweeps']}, "
            f"relaxations {st['relaxations']}, visits {st['patch_visits']} (plan {time.time() - t:.1f}s)")
    # The same reference-schedule field with its rings and ring-numbered tables
    # built on the device (gp_ptp_opts.device_setup) instead of on the host.
    dev_setup = None
    if not only or "plain_ref_f32" in only:
        t = time.time()
        plan = P.Plan(mesh, src, precision=L.PREC_F32, device_setup=True)
        setup_s = time.time() - t
        r = time_plan(plan, args.steps, args.warmup)
        plan.close()
        dev_setup = {"row": "plain_ref_f32", "host_setup_s": round(setup_s, 4), "kernel_ms": round(r["kernel_ms"], 4),
                     "sweeps": r["st"]["sweeps"],
                     "same_as_host_built": bool(np.array_equal(r["dist"], rows["plain_ref_f32"]["dist"]))
                     if "plain_ref_f32" in rows else None}
        log(f"[bench] plain_ref_f32 with device setup: {dev_setup}")
    if net is not None and (not only or "band_learned" in only):
        t_pred = time.perf_counter()
        dhat, pred_ms = net.predict(mesh, src)
        pred_wall_ms = (time.perf_counter() - t_pred) * 1e3
        t = time.time()
        # Band width and launch shape by the library's autotuners.
        plan, w = P.Plan.tuned_bands(mesh, src, dhat, schedule=L.SCHED_LEARNED_BAND, tol=1e-6)
        tuned["band_learned"] = {"band_width": w, **plan.autotune()}
        band_kw = dict(band_width=w, tol=1e-6)
        setup_s = time.time() - t
        rows["band_learned"] = time_plan(plan, args.steps, args.warmup)
        rows["band_learned"]["setup_s"] = setup_s
        rows["band_learned"]["predict_ms"] = pred_ms
        rows["band_learned"]["predict_wall_ms"] = pred_wall_ms
        plan.close()
        st = rows["band_learned"]["st"]
        log(f"[bench] band_learned: {rows['band_learned']['kernel_ms']:.3f} ms/field (+{pred_ms:.2f} ms features+MLP kernels, "
            f"{pred_wall_ms:.1f} ms host call with toplesets), "
            f"sweeps {st['sweeps']}, visits {st['patch_visits']}")

    # End to end: a fresh mesh handle (its CHE built beforehand, outside the
    # timer, as the reference builds its che), then plan (rings, renumbering,
    # tables, upload), one run and the result download.
    e2e = {}
    if not only or "e2e" in only:
        for name, kw in (("host_setup", {}), ("device_setup", {"device_setup": True})):
            walls, kms = [], []
            for _ in range(3):
                C = P.che(xyz, faces)
                t0 = time.perf_counter()
                plan = P.Plan(C, src, precision=L.PREC_F32, **kw)
                plan.autotune()            # the launch shape tuned above, from the cache
                stx = plan.run()
                plan.result()
                plan.close()
                walls.append((time.perf_counter() - t0) * 1e3)
                kms.append(stx["kernel_ms"])
                del C
            e2e[name] = {"ms_min": round(min(walls), 2), "ms_mean": round(float(np.mean(walls)), 2),
                         "kernel_ms": round(float(np.mean(kms)), 4)}
            log(f"[bench] end to end, plain_ref_f32, {name}: {e2e[name]}")
        # The drop-in call (gp_ptp: the front end's PTP_GPU), which also
        # returns the reference's BFS order (sorted_index).
        walls = []
        for _ in range(3):
            C = P.che(xyz, faces)
            t0 = time.perf_counter()
            P.ptp(C, src, device_setup=True)
            walls.append((time.perf_counter() - t0) * 1e3)
            del C
        e2e["gp_ptp_with_sorted_index"] = {"ms_min": round(min(walls), 2), "ms_mean": round(float(np.mean(walls)), 2)}
        log(f"[bench] end to end, gp_ptp with sorted_index: {e2e['gp_ptp_with_sorted_index']}")
        # The learned-band field end to end: prediction (features need the
        # rings), band plan (bands, renumbering, tables, upload), one run,
        # result download.
        for tag, dev in (("band_learned", False), ("band_learned_device_setup", True)):
            if net is None:
                break
            walls, parts = [], []
            for _ in range(3):
                C = P.che(xyz, faces)
                t0 = time.perf_counter()
                dhat, _ = net.predict(C, src)
                t1 = time.perf_counter()
                plan = P.Plan(C, src, dhat=dhat, schedule=L.SCHED_LEARNED_BAND, device_setup=dev, **band_kw)
                plan.autotune()            # the launch shape tuned above, from the cache
                t2 = time.perf_counter()
                stx = plan.run()
                plan.result()
                plan.close()
                walls.append((time.perf_counter() - t0) * 1e3)
                parts.append(((t1 - t0) * 1e3, (t2 - t1) * 1e3, stx["kernel_ms"]))
                del C
            k = int(np.argmin(walls))
            e2e[tag] = {"ms_min": round(walls[k], 2), "predict_ms": round(parts[k][0], 2),
                        "plan_ms": round(parts[k][1], 2), "kernel_ms": round(parts[k][2], 4)}
            log(f"[bench] end to end, {tag}: {e2e[tag]}")
        # The same in one call (gp_plan_create_learned with device_setup): one
        # BFS feeds the features and the plan, the prediction stays on the GPU.
        if net is not None:
            walls, parts = [], []
            for _ in range(3):
                C = P.che(xyz, faces)
                t0 = time.perf_counter()
                plan = P.Plan.learned(C, src, net, device_setup=True, **band_kw)
                plan.autotune()
                t1 = time.perf_counter()
                stx = plan.run()
                plan.result()
                plan.close()
                walls.append((time.perf_counter() - t0) * 1e3)
                parts.append(((t1 - t0) * 1e3, stx["kernel_ms"]))
                del C
            k = int(np.argmin(walls))
            e2e["band_learned_one_call"] = {"ms_min": round(walls[k], 2), "predict_and_plan_ms": round(parts[k][0], 2),
                                            "kernel_ms": round(parts[k][1], 4)}
            log(f"[bench] end to end, band_learned_one_call: {e2e['band_learned_one_call']}")
    # The 16-source field of the same mesh (one multi-source field).
    src16 = W.c2(args.seed, 16, args.n)[2]
    multi = {}
    if not only or "src16" in only:
        for name, kw in (("plain_ref_f32", dict(precision=L.PREC_F32)),
                         ("plain_uncapped_f32", dict(precision=L.PREC_F32, tol=1e-6, uncapped=True))):
            plan = P.Plan(mesh, src16, **kw)
            multi[name] = time_plan(plan, args.steps, args.warmup)
            plan.close()
            log(f"[bench] 16 sources, {name}: {multi[name]['kernel_ms']:.3f} ms/field, "
                f"sweeps {multi[name]['st']['sweeps']}")

    errs = {}
    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
        ncores = os.cpu_count() or 1
        om = O.Mesh(xyz, faces)
        ts = om.toplesets(src)                     # outside the timer, like the reference harness
        times, t_start = [], time.time()
        while True:
            t0 = time.perf_counter()
            om.ptp(src, O.CPU_PLAIN, toplesets=ts, n_threads=ncores)
            times.append(time.perf_counter() - t0)
            if time.time() - t_start > args.cpu_budget * 0.5 or len(times) >= 10:
                break
        cpu = {"value": round(min(times) * 1e3, 3), "unit": "ms", "cores": ncores, "kind": "port",
               "sample": f"full 1M-vertex field, min of {len(times)} runs of the FP64 OpenMP restatement of "
                         f"parallel_toplesets_propagation_cpu (toplesets outside the timer)"}
        ref_sched, _ = om.ptp(src, O.GPU_COAL, toplesets=ts, n_threads=ncores)
        t0 = time.time()
        fixed = om.fixed_point(src)
        log(f"[bench] oracle fixed point {time.time() - t0:.1f}s")
        for name, r in rows.items():
            e_sched = rel_err(r["dist"], ref_sched)
            e_fixed = rel_err(r["dist"], fixed)
            errs[name] = {"max_vs_ref_schedule": e_sched[0], "mean_vs_ref_schedule": e_sched[1],
                          "max_vs_fixed_point": e_fixed[0], "mean_vs_fixed_point": e_fixed[1]}
            log(f"[bench] {name}: vs reference schedule {e_sched[0]:.2e}, vs fixed point {e_fixed[0]:.2e}")
        if multi:
            ref16, _ = om.ptp(src16, O.GPU_COAL, n_threads=ncores)
            for name, r in multi.items():
                e = rel_err(r["dist"], ref16)
                r["max_vs_ref_schedule"], r["mean_vs_ref_schedule"] = e
                log(f"[bench] 16 sources, {name}: vs reference schedule {e[0]:.2e}")

    if not rows:
        print(json.dumps({"metric": "end to end and 16-source rows only (no timed field rows selected)",
                          "end_to_end": e2e or None,
                          "sources16": {n: {"kernel_ms": round(r["kernel_ms"], 4), "sweeps": r["st"]["sweeps"]}
                                        for n, r in multi.items()} or None}), flush=True)
        return

    # Headline: learned bands if they meet the bar, else the fastest row that does.
    def meets(name):
        return name in errs and errs[name]["max_vs_fixed_point"] <= ERR_BAR
    candidates = [n for n in rows if n.startswith("band_") or n == "patches" or n == "plain_uncapped_f32"] or list(rows)
    headline = "band_learned" if ("band_learned" in rows and (args.no_cpu or meets("band_learned"))) else None
    if headline is None:
        ok = [n for n in candidates if meets(n)] or candidates
        headline = min(ok, key=lambda n: rows[n]["kernel_ms"])
    h = rows[headline]
    st = h["st"]
    patch_row = headline == "patches"
    # Algorithmic bytes: B_relax per vertex relaxation that reads the vertex's
    # tables from memory (window vertices per sweep; for the patch schedules,
    # patch vertices per visit -- the local iterations run in LDS).
    loads = st["cond_checks"] if patch_row else st["relaxations"]
    bytes_alg = loads * B_RELAX_F32 + (0 if patch_row else st["cond_checks"] * 8)
    launches = max(1, st["sweeps"] + st["check_sweeps"])
    avg_launch_s = h["kernel_ms"] * 1e-3 / launches          # hipEvent time / launches
    achieved = bytes_alg / launches / avg_launch_s / 1e9
    traffic = MEASURED_TRAFFIC.get((headline, args.n, args.sources)) if args.seed == 0 else None
    out = {
        "metric": "ms per 1M-vertex distance field, learned-band vs plain PTP at <=1e-4 err; GB/s",
        "value": round(h["kernel_ms"], 4),
        "unit": "ms",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(h["wall_ms"], 4),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"height field {args.n}x{args.n} ({mesh.n_vertices()} vertices), {args.sources} "
                               f"source(s), seed "
                               f"{args.seed}; value = {headline}", "sources": src.tolist(),
                   "band_width": tuned.get(headline, {}).get("band_width"), "local_iters": args.local_iters,
                   "patch": args.patch},
        "autotuned": tuned,
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": round(achieved / PEAK_HBM_GBS, 5),
                     "traffic": round(traffic[0]) if traffic else None,
                     "traffic_note": traffic[1] if traffic else None,
                     "algorithmic_bytes_per_launch": round(bytes_alg / launches),
                     "avg_launch_us": round(avg_launch_s * 1e6, 3),
                     "kernel": f"gp::k_patch_sweep<{args.patch}u>" if patch_row else L.sweep_kernel_name(st["kernel"])},
        "speedup_vs_plain_at_bar": (round(rows["plain_uncapped_f32"]["kernel_ms"] / h["kernel_ms"], 3)
                                    if "plain_uncapped_f32" in rows else None),
        "speedup_vs_plain_ref": (round(rows["plain_ref_f32"]["kernel_ms"] / h["kernel_ms"], 3)
                                 if "plain_ref_f32" in rows else None),
        "rows": {n: {"kernel_ms": round(r["kernel_ms"], 4), "kernel_ms_min": round(r["kernel_ms_min"], 4),
                     "sweeps": r["st"]["sweeps"], "check_sweeps": r["st"]["check_sweeps"],
                     "precision": "f64" if r["st"]["precision"] == L.PREC_F64 else "f32",
                     "relaxations": r["st"]["relaxations"],
                     "f64_updates": r["st"]["f64_updates"], "patch_visits": r["st"]["patch_visits"],
                     "host_setup_s": round(r["setup_s"], 3),
                     "deferred": r["st"]["deferred"], **errs.get(n, {}),
                     **({"predict_ms": round(r["predict_ms"], 4),
                         "predict_wall_ms": round(r["predict_wall_ms"], 2)} if "predict_ms" in r else {})}
                 for n, r in rows.items()},
    }
    if dev_setup:
        out["device_setup"] = dev_setup
    if e2e:
        out["end_to_end"] = e2e
    if multi:
        out["sources16"] = {"sources": src16.tolist(),
                            **{n: {"kernel_ms": round(r["kernel_ms"], 4), "sweeps": r["st"]["sweeps"],
                                   "relaxations": r["st"]["relaxations"],
                                   **({"max_vs_ref_schedule": r["max_vs_ref_schedule"]}
                                      if "max_vs_ref_schedule" in r else {})}
                               for n, r in multi.items()}}
    if cpu:
        out["cpu_baseline"] = cpu
    # The batched (C3) and precision-stress (C5) configurations, each timed
    # over 3 launches after 1 warm-up and checked against the FP64 oracle.
    if not args.no_c3c5 and not only:
        out["c3"] = batch(args, L, P, G, steps=3, warmup=1, emit=False)
        out["c5"] = stress(args, L, P, G, steps=3, warmup=1, emit=False)
    if not args.no_next and not only:
        out["next"] = next_rows(args, L, P, G, emit=False)
    print(json.dumps(out), flush=True)


def next_rows(args, L, P, G, emit=True):
    """The rows around the hot path on the default height field: farthest-point
    sampling (32 samples from one seed, FP32 and FP64; the FP64 oracle
    restatement's first rounds timed on the host cores and compared sample for
    sample), cluster labels (16 sources; FP64 labels and distances against the
    oracle), the uncapped iteration-error loop (one source, the FP64 field of
    the same loop as the exact input) and the mesh readers (OFF text and
    binary PLY files of the same mesh, read back and compared)."""
    import tempfile
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    t = time.time()
    xyz, faces = G.height_field(args.n, args.seed)
    src = G.draw_sources(np.random.default_rng(args.seed + 1000), len(xyz), 1)
    mesh = P.che(xyz, faces)
    om = O.Mesh(xyz, faces)
    nv = mesh.n_vertices()
    log(f"[next] mesh V={nv} ({time.time() - t:.1f}s)")
    rows = {}
    only = set(filter(None, args.only.split(",")))   # fps, clusters, iter_error, readers

    def want(k):
        return not only or k in only
    # Farthest-point sampling.
    n_fps, n_cpu = 32, 5
    fps = {}
    for name, prec in (("fps_f32", L.PREC_F32), ("fps_f64", L.PREC_F64)) if want("fps") else ():
        P.farthest_point_sampling_ptp_gpu(mesh, src, 3, 0.0, prec)
        smp, r, secs = P.farthest_point_sampling_ptp_gpu(mesh, src, n_fps, 0.0, prec)
        fps[name] = smp
        rows[name] = {"ms_per_sample": round(secs * 1e3 / (len(smp) - len(src)), 3), "samples": len(smp),
                      "radius": r, "seconds": round(secs, 4)}
        log(f"[next] {name}: {len(smp)} samples in {secs * 1e3:.1f} ms, radius {r:.6f}")
    if want("fps") and not args.no_cpu:
        t0 = time.perf_counter()
        cpu_smp, _ = om.fps(src, n_cpu)
        cpu_s = time.perf_counter() - t0
        rows["fps_f64"]["oracle_prefix_identical"] = bool(np.array_equal(fps["fps_f64"][:n_cpu], cpu_smp))
        rows["fps_cpu_oracle"] = {"ms_per_sample": round(cpu_s * 1e3 / (len(cpu_smp) - len(src)), 1),
                                  "samples": len(cpu_smp), "cores": os.cpu_count()}
        log(f"[next] fps oracle: {len(cpu_smp)} samples in {cpu_s:.1f}s; GPU FP64 prefix identical: "
            f"{rows['fps_f64']['oracle_prefix_identical']}")
    # Cluster labels.
    src16 = G.draw_sources(np.random.default_rng(args.seed + 2000), nv, 16)
    for name, prec in (("clusters_f32", L.PREC_F32), ("clusters_f64", L.PREC_F64)) if want("clusters") else ():
        plan = P.Plan(mesh, src16, precision=prec, clusters=True)
        r = time_plan(plan, args.steps, args.warmup)
        labels = plan.clusters()
        plan.close()
        rows[name] = {"kernel_ms": round(r["kernel_ms"], 4), "sweeps": r["st"]["sweeps"]}
        if prec == L.PREC_F64 and not args.no_cpu:
            t0 = time.perf_counter()
            dref, cref, nsw = om.ptp_clusters(src16, O.GPU_COAL)
            rows[name]["oracle_seconds"] = round(time.perf_counter() - t0, 2)
            rows[name]["labels_identical"] = bool(np.array_equal(labels, cref))
            rows[name]["distances_identical"] = bool(np.array_equal(r["dist"], dref))
        log(f"[next] {name}: {r['kernel_ms']:.2f} ms/field, sweeps {r['st']['sweeps']} {rows[name]}")
    # Iteration-error loop (one sweep per launch with a host read-back, as the
    # reference's instrumented test runs it).
    if want("iter_error"):
        exact, _, _ = P.ptp(mesh, src, precision=L.PREC_F64, tol=1e-6, uncapped=True)
        t0 = time.perf_counter()
        it, err, band = P.iter_error_parallel_toplesets_propagation_coalescence_gpu(mesh, src, exact, L.PREC_F32)
        rows["iter_error_f32"] = {"seconds": round(time.perf_counter() - t0, 3), "sweeps": int(len(band)),
                                  "error_rows": int(len(it)),
                                  "last_error_pct": float(err[-1]) if len(err) else None}
        log(f"[next] iter_error: {rows['iter_error_f32']}")
    # Mesh readers.
    if want("readers"):
        with tempfile.TemporaryDirectory() as d:
            off, ply = os.path.join(d, "m.off"), os.path.join(d, "m.ply")
            with open(off, "w") as fh:
                fh.write(f"OFF\n{nv} {len(faces)} 0\n")
                np.savetxt(fh, xyz, fmt="%.17g")
                np.savetxt(fh, np.hstack([np.full((len(faces), 1), 3), faces]), fmt="%d")
            with open(ply, "wb") as fh:
                fh.write((f"ply\nformat binary_little_endian 1.0\nelement vertex {nv}\nproperty double x\n"
                          f"property double y\nproperty double z\nelement face {len(faces)}\n"
                          f"property list uchar int vertex_indices\nend_header\n").encode())
                xyz.astype("<f8").tofile(fh)
                rec = np.zeros(len(faces), dtype=[("n", "u1"), ("v", "<i4", 3)])
                rec["n"] = 3
                rec["v"] = faces
                rec.tofile(fh)
            for name, path in (("read_off", off), ("read_ply", ply)):
                t0 = time.perf_counter()
                m2 = P.che.read(path)
                secs = time.perf_counter() - t0
                same = bool(np.array_equal(m2.gt(), xyz) and np.array_equal(m2.tables()[0].reshape(-1, 3), faces))
                rows[name] = {"ms": round(secs * 1e3, 1), "MB_per_s": round(os.path.getsize(path) / secs / 1e6, 1),
                              "file_MB": round(os.path.getsize(path) / 1e6, 1), "identical": same}
                log(f"[next] {name}: {rows[name]}")
    out = {"metric": "rows around the hot path: FPS ms/sample, clusters ms/field, iteration-error loop, mesh readers",
           "unit": "ms", "higher_is_better": False, "n_gpus": 1, "data": "synthetic",
           "config": {"workload": f"height field {args.n}x{args.n} ({nv} vertices), seed {args.seed}; FPS from "
                                  f"{src.tolist()}, 32 samples; clusters with 16 sources"},
           "rows": rows}
    if emit:
        print(json.dumps(out), flush=True)
    return out


def stress(args, L, P, G, steps=None, warmup=None, emit=True):
    """C5: 2048 x 2048 grid stretched 1:100 in x, jitter U(-0.4, 0.4) cells, 4
    sources, the reference schedule, against the FP64 oracle (it runs on the
    host cores).  Rows: f32_request (an FP32 request; over a quarter of this
    mesh's triangles are flagged, so it runs FP64 sweeps), f32_forced (FP32
    sweeps forced with the FP32 geometry table, the FP64 update on flagged
    triangles), f64."""
    from gproshan_amd import workloads as W
    steps = args.steps if steps is None else steps
    warmup = args.warmup if warmup is None else warmup
    t = time.time()
    xyz, faces, src = W.c5(args.seed)
    mesh = P.che(xyz, faces)
    log(f"[stress] V={mesh.n_vertices()} sources={src.tolist()} ({time.time() - t:.1f}s)")
    rows, traces = {}, {}
    for name, kw in (("f32_request", dict(precision=L.PREC_F32)),
                     ("f32_forced", dict(precision=L.PREC_F32, fp32_geometry=L.GEO_FP32)),
                     ("f64", dict(precision=L.PREC_F64))):
        plan = P.Plan(mesh, src, **kw)
        rows[name] = time_plan(plan, steps, warmup)
        traces[name] = plan.trace()
        plan.close()
        log(f"[stress] {name}: {rows[name]['kernel_ms']:.2f} ms, sweeps {rows[name]['st']['sweeps']}, "
            f"precision {rows[name]['st']['precision']}")
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    om = O.Mesh(xyz, faces)
    t0 = time.perf_counter()
    ref, nsw = om.ptp(src, O.GPU_COAL, n_threads=os.cpu_count() or 1)
    cpu_s = time.perf_counter() - t0
    # Where the forced-FP32 run's schedule first departs from the FP64 run's
    # (a ring test flipped by rounding moves values by about PTP_TOL).
    t32, t64 = traces["f32_forced"][:, 1:7], traces["f64"][:, 1:7]
    n = min(len(t32), len(t64))
    diff = np.flatnonzero(np.any(t32[:n] != t64[:n], axis=1))
    first_diff = int(diff[0]) + 1 if len(diff) else (None if len(t32) == len(t64) else n + 1)
    out = {"metric": "C5 precision stress: max relative error vs the FP64 oracle", "unit": "relative error",
           "f32_forced_schedule_matches_f64": first_diff is None, "f32_forced_first_differing_sweep": first_diff,
           "config": {"workload": "anisotropic grid 2048x2048 (4194304 vertices), stretch 1:100, jitter 0.4, "
                                  f"4 sources, seed {args.seed}", "sources": src.tolist()},
           "oracle": {"sweeps": nsw, "seconds": round(cpu_s, 2), "cores": os.cpu_count()},
           "rows": {}}
    for name, r in rows.items():
        mx, mn = rel_err(r["dist"], ref)
        st = r["st"]
        gbs = algorithmic_bytes(st, st["precision"] == L.PREC_F64) / (r["kernel_ms"] * 1e-3) / 1e9
        out["rows"][name] = {"kernel_ms": round(r["kernel_ms"], 3), "GB_per_s": round(gbs, 1),
                             "sweeps": st["sweeps"], "precision": "f64" if st["precision"] == L.PREC_F64 else "f32",
                             "max_rel_err": mx, "mean_rel_err": mn, "meets_1e-4": mx <= ERR_BAR,
                             "f64_updates": st["f64_updates"], "relaxations": st["relaxations"]}
        log(f"[stress] {name}: max {mx:.3e} mean {mn:.3e} vs FP64 oracle ({nsw} sweeps, {cpu_s:.1f}s on the host)")
    if emit:
        print(json.dumps(out), flush=True)
    return out


def batch(args, L, P, G, steps=None, warmup=None, emit=True):
    """C3: 256 meshes (grid height fields and displaced icospheres), vertex
    counts log-uniform in [2k, 200k], U{1..64} sources each; every (mesh,
    source) pair is its own one-source field, all in one persistent launch."""
    from gproshan_amd import workloads as W
    steps = args.steps if steps is None else steps
    warmup = args.warmup if warmup is None else warmup
    t = time.time()
    raw, jobs = W.c3(args.seed)
    meshes = [P.che(xyz, f) for xyz, f in raw]
    nvf = sum(meshes[m].n_vertices() for m, _ in jobs)

    def max_neighbours(f):
        e = np.sort(np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]]), axis=1)
        e = np.unique(e, axis=0)
        return int(np.bincount(e.ravel()).max())
    # The kernel holds 8 slots in registers when every fan has at most 8
    # neighbours (closed fans store one slot per neighbour), else 10.
    ms_slots = 8 if max(max_neighbours(f) for _, f in raw) <= 8 else 10
    log(f"[batch] {len(meshes)} meshes, {len(jobs)} fields, {nvf} vertex-fields ({time.time() - t:.1f}s)")
    t = time.time()
    b = P.Batch(meshes, jobs, block=0)
    log(f"[batch] created and autotuned in {time.time() - t:.1f}s")
    for _ in range(warmup):
        b.run()
    ms = []
    for _ in range(steps):
        st = b.run()
        ms.append(st["kernel_ms"])
    kms = float(np.mean(ms))
    gbs = st["relaxations"] * B_RELAX_F32 / (kms * 1e-3) / 1e9
    log(f"[batch] {kms:.2f} ms per launch, block {st['block']}, grid {st['grid']}, sweeps {st['sweeps']}, "
        f"relaxations {st['relaxations']}, {gbs:.1f} GB/s algorithmic")
    # Check fields against the FP64 oracle: 24 drawn from the batch (seeded).
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    pick = np.random.default_rng(args.seed + 3).choice(len(jobs), 24, replace=False)
    worst, mean_sum, t0 = 0.0, 0.0, time.time()
    for q in pick:
        m, s = jobs[q]
        ref, _ = O.Mesh(*raw[m]).ptp(np.array([s], np.uint32), O.GPU_PLAIN)
        d, _ = b.result(int(q))
        mx, mn = rel_err(d, ref)
        worst = max(worst, mx)
        mean_sum += mn
    log(f"[batch] max relative error of 24 drawn fields vs the FP64 oracle: {worst:.2e} ({time.time() - t0:.1f}s)")
    b.close()
    out = {"metric": "C3 batched PTP: ms per launch of all fields", "value": round(kms, 3), "unit": "ms",
           "higher_is_better": False, "n_gpus": 1, "steps": steps, "warmup": warmup, "dtype": "f32",
           "data": "synthetic",
           "config": {"workload": f"256 meshes 2k-200k vertices, {len(jobs)} one-source fields, seed {args.seed}",
                      "vertex_fields": nvf, "block": st["block"], "grid": st["grid"]},
           "kernel_ms": round(kms, 3), "GB_per_s": round(gbs, 1), "sweeps": st["sweeps"],
           "relaxations": st["relaxations"],
           "fields_per_s": round(len(jobs) / (kms * 1e-3), 1),
           "vertex_fields_per_s": round(nvf / (kms * 1e-3), 1),
           "roofline": {"bound": "hbm", "achieved": round(gbs, 2), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                        "frac": round(gbs / PEAK_HBM_GBS, 5), "traffic": None, "kernel": f"gp::k_batch<{st['block']}u, {ms_slots}u>"},
           "max_rel_err_24_fields": worst, "mean_rel_err_24_fields": mean_sum / len(pick),
           "meets_1e-4": worst <= ERR_BAR}
    if emit:
        print(json.dumps(out), flush=True)
    return out


if __name__ == "__main__":
    main()
