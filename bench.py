"""Benchmark: 2M-face mesh -> 20k faces on one MI355X.

Workload (synthetic): marching-tetrahedra contour of the seed-1 union of 16
rotated ellipsoids and 4 boxes on [-1,1]^3 at grid 296 (2,006,528 faces,
1,003,254 vertices, fp32 positions, int32 faces, watertight and manifold),
simplified with ratio 0.01, tolerance 4, threshold 1e-3, iter 100000,
min_verts 0, as PaMO(mesh, use_stage1=False, use_stage3=True).run(...).

One step = one full PaMO.run from device-resident input arrays to the numpy
result the reference returns: centring, every stage-2 round with its guard
and undo loop, and stage 3 (five steps of ten Newton iterations against the
2M-face input surface), each step from a fresh object's state.  Before the
timed steps the launch-shape autotuner picks the guard-tree reuse and guard
split on this mesh (results unchanged).  Prints one JSON line with:

- roofline: the dominant kernel's algorithmic bytes over its HIP-event time,
  and its HBM bytes per launch from two rocprofv3 --pmc passes (FETCH_SIZE,
  WRITE_SIZE) over one stage-2 run, scaled by a plain copy's counters;
- path_roofline: the whole timed step's algorithmic bytes (the per-round and
  per-Newton-iteration byte model of the hot path) over its wall time;
- cpu_baseline: the CPU oracle on this box's cores -- the whole stage-2 run
  (also compared bit for bit with this run's stage 2: stage2_vs_oracle) and
  stage 3 timed per Newton iteration and scaled;
- legs: the headline mesh through the whole pipeline (stage-1 remeshing
  included, the form the reference's published figure most likely covers),
  the icosphere workload (stage 2, against the oracle), a batch of meshes
  against the plain form, and a short served Poisson stream against plain
  FIFO.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# The copy that calibrates the HBM counters: FETCH_SIZE and WRITE_SIZE of a
# plain device-to-device copy of this size convert the counters to bytes.
CALIB_COPY_MB = 512


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=3)
    p.add_argument("--warmup", type=int, default=1)
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--stage3", type=int, default=1, help="1/0: stage 3 on/off")
    p.add_argument("--exact-guard", action="store_true",
                   help="exact intersection guard instead of the reference's predicate (results differ from the reference)")
    p.add_argument("--iters", type=int, default=100000,
                   help="cap on stage-2 rounds (the reference's iter; the workload uses 100000)")
    p.add_argument("--stress-iters", type=int, default=400,
                   help="stress: cap on stage-2 rounds (the sheets crawl under the guard; uncapped takes minutes)")
    p.add_argument("--workload", choices=["headline", "batch", "serve", "stress"], default="headline",
                   help="headline: one 2M-face mesh; batch: many meshes, batched vs plain; serve: a request stream")
    p.add_argument("--batch-n", type=int, default=96, help="batch: number of meshes")
    p.add_argument("--slots", type=int, default=8, help="batch / serve: concurrent slots")
    p.add_argument("--serve-seconds", type=float, default=5.0, help="serve: length of the arrival stream")
    p.add_argument("--serve-rate", type=float, default=40.0, help="serve: Poisson arrival rate (jobs/s)")
    p.add_argument("--no-legs", action="store_true", help="headline only (no icosphere / batch / serve legs)")
    p.add_argument("--no-autotune", action="store_true", help="keep the default launch shapes")
    p.add_argument("--leg-batch-n", type=int, default=24, help="legs: meshes in the batch leg")
    p.add_argument("--leg-serve-seconds", type=float, default=1.5, help="legs: length of the served stream")
    p.add_argument("--no-traffic", action="store_true",
                   help="skip the two rocprofv3 --pmc passes that measure the dominant kernel's HBM traffic")
    p.add_argument("--traffic-probe", action="store_true",
                   help="(child of the traffic passes) one stage-2 run of the workload, then the calibration copy")
    p.add_argument("--calib-copy-mb", type=int, default=0,
                   help="after the run, one device-to-device copy of this many MiB (HBM counter calibration)")
    return p.parse_args()


def bench_batch(args, n_meshes=None):
    """Batched stages 2+3 over many meshes (pamo_batch_run) against the plain
    form, one PaMO.run per mesh in sequence, on the same device-resident
    meshes.  Returns the result record."""
    import numpy as np
    import torch
    from pamo_amd import synth
    from pamo_amd.batch import pack_csr, simplify_batch
    from pamo_amd.pamo import PaMO

    meshes, ratios = synth.batch_workload(args.batch_n if n_meshes is None else n_meshes)
    nF = int(sum(len(F) for _, F in meshes))
    print(f"[bench] batch: {len(meshes)} meshes, {nF} faces", file=sys.stderr, flush=True)
    packed = pack_csr(meshes)
    dev = [(torch.from_numpy(V).cuda(), torch.from_numpy(F).cuda()) for V, F in meshes]
    # warm-up: a small batch and one plain run
    simplify_batch(meshes[:4], ratios[:4], True, n_slots=args.slots)
    PaMO(meshes[0], use_stage1=False, use_stage3=True).run(dev[0][0], dev[0][1], float(ratios[0]), 4, 1e-3,
                                                           100000, 0)
    torch.cuda.synchronize()
    print("[bench] warmup done", file=sys.stderr, flush=True)
    # plain form: one PaMO object and one run per mesh, as a caller of the
    # reference does
    t0 = time.perf_counter()
    plain = []
    for k, ((V, F), (v, f), r) in enumerate(zip(meshes, dev, ratios)):
        plain.append(PaMO((V, F), use_stage1=False, use_stage3=True).run(v, f, float(r), 4, 1e-3, 100000, 0))
        if (k + 1) % 16 == 0:
            print(f"[bench] plain {k + 1}/{len(meshes)} meshes, {time.perf_counter() - t0:.1f}s", file=sys.stderr,
                  flush=True)
    torch.cuda.synchronize()
    t_plain = time.perf_counter() - t0
    print(f"[bench] plain {t_plain:.2f}s", file=sys.stderr, flush=True)
    done = []

    def on_done(mesh, status):
        done.append(mesh)
        if len(done) % 8 == 0 or status != 0:
            print(f"[bench] batch {len(done)}/{len(meshes)} meshes (mesh {mesh} status {status})", file=sys.stderr,
                  flush=True)

    t0 = time.perf_counter()
    out, secs = simplify_batch(None, ratios, True, n_slots=args.slots, packed=packed, on_done=on_done)
    torch.cuda.synchronize()
    t_batch = time.perf_counter() - t0
    same_faces = all(np.array_equal(a[1].astype(np.int64), b[1].astype(np.int64)) for a, b in zip(out, plain))
    n = len(meshes)
    res = {"metric": "batched meshes/s (stages 2+3)", "value": round(n / t_batch, 3), "unit": "meshes/s",
           "n_gpus": 1, "higher_is_better": True, "vs_baseline": None, "dtype": "f32", "data": "synthetic",
           "config": {"workload": f"{n} ellipsoid/box unions, faces log-uniform in [20k, 800k], ratio in "
                                  "{0.005, 0.01, 0.02, 0.05}", "faces_in": nF, "slots": args.slots},
           "plain_meshes_per_s": round(n / t_plain, 3), "speedup_vs_plain": round(t_plain / t_batch, 3),
           "batch_s": round(t_batch, 3), "plain_s": round(t_plain, 3), "faces_identical_to_plain": same_faces}
    return res


def bench_serve(args, seconds=None, one_slot=True):
    """A Poisson request stream served by pamo_serve_* with --slots slots,
    against FIFO one-at-a-time service of the same stream, in the plain form
    (one PaMO.run per request) and, with one_slot, through the engine with
    one slot.  Returns the result record."""
    import numpy as np
    import torch
    from pamo_amd import synth, _hip
    from pamo_amd.batch import Server

    seconds = args.serve_seconds if seconds is None else seconds
    pool, jobs = synth.serve_stream(seconds, args.serve_rate)
    dev = [(torch.from_numpy(V).cuda(), torch.from_numpy(F).cuda()) for V, F in pool]
    print(f"[bench] serve: {len(jobs)} jobs over {seconds}s, pool of {len(pool)} meshes",
          file=sys.stderr, flush=True)

    def run_stream(n_slots):
        srv = Server(n_slots=n_slots, stage3=True)
        srv.result(srv.submit(*dev[0], 0.05, True))  # warm-up
        ids = []
        t0 = time.perf_counter()
        for (ta, k, r, s3) in jobs:
            dt = t0 + ta - time.perf_counter()
            if dt > 0:
                time.sleep(dt)
            ids.append(srv.submit(dev[k][0], dev[k][1], r, s3))
        lat, qd, failed = [], [], []
        for k, i in enumerate(ids):
            info = srv.poll(i, wait=True)
            lat.append(info["latency_s"])
            qd.append(info["queue_s"])
            if info["state"] == "failed":
                # the reference raises on the same input (stage 3's closed-manifold assertion)
                ta, m, r, s3 = jobs[k]
                failed.append({"pool_mesh": m, "faces": int(len(pool[m][1])), "ratio": round(r, 5), "stage3": s3})
                try:
                    srv.result(i)
                except _hip.PamoError as err:
                    failed[-1]["error"] = str(err).split(": ")[-1]
            else:
                srv.result(i)
            if (k + 1) % 25 == 0:
                print(f"[bench] {n_slots}-slot stream: {k + 1}/{len(ids)} jobs done", file=sys.stderr, flush=True)
        t_end = time.perf_counter() - t0
        srv.close()
        return np.array(lat), np.array(qd), t_end, failed

    def run_plain_fifo():
        # the plain form: one PaMO object and one run per request, in arrival
        # order (a request that arrives while another runs waits for it)
        from pamo_amd.pamo import PaMO
        lat, failed = [], 0
        t0 = time.perf_counter()
        for k, (ta, m, r, s3) in enumerate(jobs):
            dt = t0 + ta - time.perf_counter()
            if dt > 0:
                time.sleep(dt)
            try:
                PaMO(pool[m], use_stage1=False, use_stage3=s3).run(dev[m][0], dev[m][1], r, 4, 1e-3, 100000, 0)
            except Exception:
                failed += 1
            lat.append(time.perf_counter() - (t0 + ta))
            if (k + 1) % 25 == 0:
                print(f"[bench] plain fifo: {k + 1}/{len(jobs)} jobs done", file=sys.stderr, flush=True)
        torch.cuda.synchronize()
        return np.array(lat), time.perf_counter() - t0, failed

    lat, qd, t_all, failed = run_stream(args.slots)
    print(f"[bench] served in {t_all:.1f}s", file=sys.stderr, flush=True)
    if one_slot:
        lat1, qd1, t1, failed1 = run_stream(1)
        print(f"[bench] one-slot fifo in {t1:.1f}s", file=sys.stderr, flush=True)
    latp, tp, failedp = run_plain_fifo()
    print(f"[bench] plain fifo in {tp:.1f}s", file=sys.stderr, flush=True)
    n = len(jobs)
    res = {"metric": "served meshes/s and p50/p99 latency (stages 2+3 on half the jobs)",
           "value": round(n / t_all, 3), "unit": "meshes/s", "n_gpus": 1, "higher_is_better": True,
           "vs_baseline": None, "dtype": "f32", "data": "synthetic",
           "config": {"workload": f"Poisson {args.serve_rate}/s for {seconds}s, {n} jobs, faces "
                                  f"log-uniform in [10k, 2M] over a pool of {len(pool)} meshes, ratio U[0.002, 0.05]",
                      "slots": args.slots},
           "p50_latency_s": round(float(np.percentile(lat, 50)), 3),
           "p99_latency_s": round(float(np.percentile(lat, 99)), 3),
           "plain_fifo_meshes_per_s": round(n / tp, 3),
           "plain_fifo_p50_latency_s": round(float(np.percentile(latp, 50)), 3),
           "plain_fifo_p99_latency_s": round(float(np.percentile(latp, 99)), 3),
           "speedup_vs_plain_fifo": round(tp / t_all, 3),
           "failed_jobs": failed, "plain_fifo_failed_jobs": failedp}
    if one_slot:
        res.update({"fifo_meshes_per_s": round(n / t1, 3),
                    "fifo_p50_latency_s": round(float(np.percentile(lat1, 50)), 3),
                    "fifo_p99_latency_s": round(float(np.percentile(lat1, 99)), 3),
                    "speedup_vs_fifo": round(t1 / t_all, 3), "fifo_failed_jobs": len(failed1)})
    return res


def bench_stress(args):
    """Intersection stress: two closely separated sheets (about 8M faces,
    ratio 0.0025, tolerance 8, stages 2+3), one mesh in the plain form and
    four (seeds 500..503) batched against the same four run one by one.
    Stage 2 is capped at --stress-iters rounds: the guard undoes most
    collapses between sheets this close, so the loop reaches the face target
    only after tens of thousands of rounds."""
    import numpy as np
    import torch
    from pamo_amd import synth
    from pamo_amd.batch import pack_csr, simplify_batch
    from pamo_amd.pamo import PaMO

    ratio, tol, it = 0.0025, 8, args.stress_iters
    s3 = bool(args.stage3)
    meshes = []
    for seed in (500, 501, 502, 503):
        V, F, res = synth.stress_sheets(seed, 8e6)
        meshes.append((V, F))
        print(f"[bench] stress sheets seed {seed}: grid {res}, {len(F)} faces", file=sys.stderr, flush=True)
    dev = [(torch.from_numpy(V).cuda(), torch.from_numpy(F).cuda()) for V, F in meshes]

    def plain(k):
        pm = PaMO(meshes[k], use_stage1=False, use_stage3=s3, exact_guard=args.exact_guard)
        out = pm.run(dev[k][0], dev[k][1], ratio, tol, 1e-3, it, 0)
        return out, pm

    plain(0)  # warm-up
    torch.cuda.synchronize()
    print("[bench] warmup done", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    (vo, fo), pm = plain(0)
    torch.cuda.synchronize()
    t_single = time.perf_counter() - t0
    print(f"[bench] single: {t_single:.2f}s, {len(fo)} faces", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    outs = [plain(k)[0] for k in range(4)]
    torch.cuda.synchronize()
    t_plain4 = time.perf_counter() - t0
    print(f"[bench] plain x4: {t_plain4:.2f}s", file=sys.stderr, flush=True)
    packed = pack_csr(meshes)
    t0 = time.perf_counter()
    bout, _ = simplify_batch(None, ratio, s3, n_slots=4, tolerance=tol, iters=it, packed=packed,
                             exact_guard=args.exact_guard)
    torch.cuda.synchronize()
    t_batch4 = time.perf_counter() - t0
    print(f"[bench] batched x4: {t_batch4:.2f}s", file=sys.stderr, flush=True)
    same = all(np.array_equal(a[1].astype(np.int64), b[1].astype(np.int64)) for a, b in zip(outs, bout))
    nF = len(meshes[0][1])
    res = {"metric": "stress sheets wall-time (stages 2+3)" if s3 else "stress sheets wall-time (stage 2)",
           "value": round(t_single, 4), "unit": "s", "n_gpus": 1, "steps": 1, "warmup": 1,
           "ms_per_step": round(t_single * 1e3, 2), "higher_is_better": False, "scaling": "strong",
           "vs_baseline": None, "dtype": "f32", "data": "synthetic",
           "config": {"workload": "two level-set sheets of the ellipsoid/box union (level 0 and -1e-3), seed 500",
                      "faces_in": nF, "faces_out": int(len(fo)), "ratio": ratio, "tolerance": tol,
                      "stage2_round_cap": it, "rounds": int(pm.rounds), "stage3": s3,
                      "guard": "exact" if args.exact_guard else "reference predicate"},
           "stage2_s": round(pm.timings["stage2_s"], 4), "stage3_s": round(pm.timings.get("stage3_s", 0.0), 4),
           "batched_x4_s": round(t_batch4, 3), "plain_x4_s": round(t_plain4, 3),
           "batched_x4_meshes_per_s": round(4 / t_batch4, 3), "plain_x4_meshes_per_s": round(4 / t_plain4, 3),
           "batched_x4_faces_identical_to_plain": same}
    print(json.dumps(res))


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(args, pm, V, F, ratio, tol, thr, iters, use_s3):
    """The CPU oracle on the same job.  Returns (cpu_baseline record,
    stage-2 parity record)."""
    import numpy as np
    from oracle import oracle
    oracle.build()
    v2, f2 = pm.stage2_out
    mean = pm.tris_mean
    # the reference's centring, in numpy
    mean_np = V[F].mean(axis=1).mean(axis=0)
    Vc = (V - mean_np).astype(np.float32)
    print("[bench] cpu baseline: stage-2 oracle, whole run", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    vo, fo, ho = oracle.run(Vc, F, ratio, tol, thr, iters=iters, min_verts=0)
    t_s2 = time.perf_counter() - t0
    parity = {"centring_bit_exact": bool(np.array_equal(np.asarray(mean, np.float32).view(np.uint32),
                                                          mean_np.astype(np.float32).view(np.uint32))),
              "face_history_equal": bool(np.array_equal(pm.face_hist, ho)),
              "faces_bit_exact": bool(np.array_equal(f2.cpu().numpy(), fo)),
              "vertices_bit_exact": bool(np.array_equal(v2.cpu().numpy().view(np.uint32), vo.view(np.uint32))),
              "oracle_rounds": int(len(ho))}
    print(f"[bench] stage-2 oracle {t_s2:.1f}s, parity {parity}", file=sys.stderr, flush=True)
    cores = oracle.lib().orc_num_threads()
    rec = {"value": round(t_s2, 3), "unit": "s", "cores": cores, "kind": "port", "cpu_model": cpu_model(),
           "stage2_s": round(t_s2, 3),
           "sample": f"stage 2: the whole oracle run ({len(ho)} rounds, same input and parameters)"}
    if use_s3:
        from pamo_amd import stage3
        from oracle import stage3 as s3o
        sV = (vo + mean_np).astype(np.float32)
        scale, t = stage3.normalization_transform(V.astype(np.float64))
        gn = (V.astype(np.float64) * scale + t).astype(np.float32)
        vn = (sV.astype(np.float64) * scale + t).astype(np.float32)
        vor = stage3.voronoi_areas(sV.astype(np.float64) * scale + t, fo).astype(np.float32)
        E, H = stage3.edges_unique(fo), stage3.hinges(fo)
        # the closest points on the input surface come from the device: the
        # oracle finds them by brute force over the 2M input faces, which is
        # not a CPU implementation worth timing
        sysg = stage3.Stage3System()
        sysg.set_target(torch_from(gn), torch_from(F), 1.0, np.zeros(3))
        sysg.set_mesh(torch_from(vn), torch_from(fo), 1.0, np.zeros(3))
        from pamo_amd import _hip
        tg = sysg.debug(_hip.PAMO_S3_DBG_TARGET).reshape(-1, 3)
        times = {}
        for n_newton in (1, 2):
            o = s3o.System(fp64=True, n_newton=n_newton)
            o.set_target(gn, F)
            o.set_mesh(vn, fo, E, H, vor)
            o.set_targets(tg)
            t0 = time.perf_counter()
            o.step(1, keep_target=True)
            times[n_newton] = time.perf_counter() - t0
            print(f"[bench] stage-3 oracle: one step of {n_newton} Newton iterations {times[n_newton]:.1f}s",
                  file=sys.stderr, flush=True)
        t_newton = max(times[2] - times[1], 0.0)
        t_s3 = 50 * t_newton
        rec["stage3_s"] = round(t_s3, 3)
        rec["value"] = round(t_s2 + t_s3, 3)
        rec["sample"] += (f"; stage 3: one step of 1 and of 2 Newton iterations (FP64) on the oracle's stage-2 "
                          f"output ({times[1]:.1f}s, {times[2]:.1f}s) give {t_newton:.2f}s per Newton iteration, "
                          f"scaled to the job's 5 steps x 10 iterations; the 5 closest-point updates against the "
                          f"2M-face input are not included (targets taken from the device)")
    return rec, parity


def torch_from(a):
    import torch
    return torch.from_numpy(a).cuda()


def leg_icosphere(args):
    """The icosphere workload (level 5, spherical-harmonic displacement,
    20,480 faces, ratio 0.05, stage 2 only) through PaMO.run, against the
    oracle on the reference-centred input."""
    import numpy as np
    import torch
    from pamo_amd import synth
    from pamo_amd.pamo import PaMO
    from oracle import oracle
    V, F = synth.icosphere_sh(5, 0)
    Vd, Fd = torch.from_numpy(V).cuda(), torch.from_numpy(F).cuda()
    pm = PaMO((V, F), use_stage1=False, use_stage3=False)
    for _ in range(2):
        pm.run(Vd, Fd, 0.05, 4, 1e-3, 100000, 0)
    ts = []
    for _ in range(7):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        vo, fo = pm.run(Vd, Fd, 0.05, 4, 1e-3, 100000, 0)
        ts.append(time.perf_counter() - t0)
    mean = V[F].mean(axis=1).mean(axis=0)
    t0 = time.perf_counter()
    vr, fr, hr = oracle.run((V - mean).astype(np.float32), F, 0.05, min_verts=0)
    t_cpu = time.perf_counter() - t0
    return {"workload": "icosphere level 5 + spherical-harmonic displacement (seed 0), 20,480 faces, ratio 0.05, "
                        "stage 2 only", "value": round(float(np.median(ts)), 5), "unit": "s",
            "faces_out": int(len(fo)), "rounds": int(pm.rounds), "runs": len(ts),
            "cpu_oracle_s": round(t_cpu, 4), "cpu_cores": oracle.lib().orc_num_threads(),
            "bit_exact_vs_oracle": bool(np.array_equal(fo, fr) and
                                        np.array_equal(vo.view(np.uint32), (vr + mean).view(np.uint32)))}


def leg_full_pipeline(V, F):
    """The headline mesh through the whole PaMO.run of the reference's default
    form, PaMO(mesh, use_stage1=True, use_stage3=True): remeshing at the
    resolution its rule picks (256 for a 20k-face target), then stages 2 and
    3; one warm-up run, then one timed run from a reset stage-2 context.
    The reference's only published figure, "just 3 seconds" for 2M -> 20k
    faces on an RTX 4090 (README.md:12), most likely covers this form."""
    import torch
    from pamo_amd import _hip
    from pamo_amd.pamo import PaMO
    Vd, Fd = torch.from_numpy(V).cuda(), torch.from_numpy(F).cuda()
    pm = PaMO((V, F), use_stage1=True, use_stage3=True)
    pm.run(Vd, Fd, 0.01, 4, 1e-3, 100000, 0)
    _hip.check(_hip.lib().pamo_ctx_reset(pm.ctx.ptr), "pamo_ctx_reset")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    vo, fo = pm.run(Vd, Fd, 0.01, 4, 1e-3, 100000, 0)
    torch.cuda.synchronize()
    t = time.perf_counter() - t0
    # output quality against the input surface, untimed (GPU metrics)
    from pamo_amd import metrics
    cd, hd = metrics.compute_igl_CD_HD(V, F, vo, fo)
    return {"workload": "the 2M-face headline mesh, PaMO(use_stage1=True, use_stage3=True).run, ratio 0.01",
            "value": round(t, 4), "unit": "s", "resolution": pm.R, "hausdorff": float(hd), "chamfer": float(cd),
            "remeshed_faces": int(len(pm.stage1_out[1])), "faces_out": int(len(fo)),
            "stage1_s": round(pm.timings["stage1_s"], 4), "stage2_s": round(pm.timings["stage2_s"], 4),
            "stage3_s": round(pm.timings.get("stage3_s", 0.0), 4),
            "published": {"value": 3.0, "unit": "s", "hardware": "RTX 4090", "source": "README.md:12",
                          "scope": "not stated"}}


def legs(args, V=None, F=None):
    """The other workloads, at sizes that keep the default run within a few
    minutes: the headline mesh through the whole pipeline (stage 1
    included), the icosphere, a batch of --leg-batch-n meshes against the
    plain form, and a --leg-serve-seconds Poisson stream against plain
    FIFO."""
    out = {}
    if V is not None:
        print("[bench] leg: full pipeline", file=sys.stderr, flush=True)
        try:
            out["full_pipeline"] = leg_full_pipeline(V, F)
        except Exception as err:  # reported, not fatal to the other legs
            out["full_pipeline"] = {"error": str(err)}
    print("[bench] leg: icosphere", file=sys.stderr, flush=True)
    out["icosphere_stage2"] = leg_icosphere(args)
    print("[bench] leg: batch", file=sys.stderr, flush=True)
    out["batch"] = bench_batch(args, n_meshes=args.leg_batch_n)
    print("[bench] leg: serve", file=sys.stderr, flush=True)
    out["serve"] = bench_serve(args, seconds=args.leg_serve_seconds, one_slot=False)
    return out


def traffic_probe():
    """The program the two counter passes profile: one stage-2 run of the
    headline workload on a fresh object, then a plain copy of CALIB_COPY_MB."""
    import torch
    from pamo_amd import synth
    from pamo_amd.pamo import PaMO
    V, F = synth.headline_mesh(1)
    Vd, Fd = torch.from_numpy(V).cuda(), torch.from_numpy(F).cuda()
    pm = PaMO((V, F), use_stage1=False, use_stage3=False)
    pm.run(Vd, Fd, 0.01, 4, 1e-3, 100000, 0)
    torch.cuda.synchronize()
    src = torch.empty(CALIB_COPY_MB << 18, dtype=torch.float32, device="cuda").fill_(1.0)
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    dst.copy_(src)
    torch.cuda.synchronize()
    print(f"[probe] {pm.rounds} rounds, copy of {CALIB_COPY_MB} MiB", file=sys.stderr, flush=True)


def measure_traffic(kernel):
    """HBM bytes per launch of `kernel` (a kernel-name fragment): two
    rocprofv3 --pmc passes over traffic_probe, FETCH_SIZE then WRITE_SIZE,
    each scaled by what the counter reads for the calibration copy.  Returns
    (bytes per launch, detail) or (None, reason)."""
    import csv
    import glob
    import shutil
    import subprocess
    import tempfile
    prof = shutil.which("rocprofv3")
    if prof is None:
        return None, "rocprofv3 not found"
    per = {}
    detail = {}
    for counter in ("FETCH_SIZE", "WRITE_SIZE"):
        out = tempfile.mkdtemp(prefix="pamo_pmc_")
        cmd = [prof, "--pmc", counter, "--kernel-include-regex", f"{kernel}|copyBuffer", "--output-format", "csv",
               "-d", out, "-o", "probe", "--", sys.executable, os.path.abspath(__file__), "--traffic-probe"]
        print(f"[bench] traffic pass: {counter}", file=sys.stderr, flush=True)
        try:
            r = subprocess.run(cmd, stdout=subprocess.DEVNULL, stderr=subprocess.PIPE, timeout=180, text=True)
        except subprocess.TimeoutExpired:
            return None, f"{counter} pass timed out"
        files = glob.glob(os.path.join(out, "**", "*counter_collection.csv"), recursive=True)
        if not files:
            return None, f"{counter} pass wrote no counters (rc {r.returncode}): {r.stderr[-300:]}"
        vals, copy = [], []
        for row in csv.DictReader(open(files[0])):
            if row.get("Counter_Name") != counter:
                continue
            name, v = row["Kernel_Name"], float(row["Counter_Value"])
            if kernel in name:
                vals.append(v)
            elif "copyBuffer" in name:
                copy.append(v)
        shutil.rmtree(out, ignore_errors=True)
        if not vals or not copy:
            return None, f"{counter}: {len(vals)} kernel and {len(copy)} copy dispatches"
        # the calibration copy is the largest copy dispatch
        calib = max(copy)
        expect = (CALIB_COPY_MB << 20) / 1024.0  # KiB the copy reads and writes
        factor = expect / calib if calib > 0 else float("nan")
        per[counter] = sum(vals) / len(vals) * 1024.0 * factor
        detail[counter] = {"launches": len(vals), "avg_kib": round(sum(vals) / len(vals), 1),
                           "calib_kib": round(calib, 1), "factor": round(factor, 4)}
    return per["FETCH_SIZE"] + per["WRITE_SIZE"], detail


def main():
    args = parse()
    if args.traffic_probe:
        return traffic_probe()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {args.gpus}); multi-GPU runs are not part of this "
              "build", file=sys.stderr)
        sys.exit(2)
    if args.workload == "batch":
        print(json.dumps(bench_batch(args)))
        return
    if args.workload == "serve":
        print(json.dumps(bench_serve(args)))
        return
    if args.workload == "stress":
        return bench_stress(args)
    import numpy as np
    import torch
    from pamo_amd import synth, _hip
    from pamo_amd.pamo import PaMO

    ratio, tol, thr, iters = 0.01, 4, 1e-3, args.iters
    V, F = synth.headline_mesh(1)
    print(f"[bench] mesh ready: {len(F)} faces", file=sys.stderr, flush=True)
    nF0, nV0 = len(F), len(V)
    dev = torch.device("cuda", 0)
    Vd = torch.from_numpy(V).to(dev)
    Fd = torch.from_numpy(F).to(dev)
    use_s3 = bool(args.stage3)
    pm = PaMO((V, F), use_stage1=False, use_stage3=use_s3, exact_guard=args.exact_guard)
    ctx = pm.ctx
    prof = _hip.Profiler(ctx)
    tune = None
    if not args.no_autotune:
        # launch shapes (guard-tree reuse, guard subtree split) picked on this
        # mesh before the timed runs; they change timings, never results
        from pamo_amd import autotune
        from pamo_amd.pamo import face_centroid_mean
        mean = face_centroid_mean(ctx, Vd, Fd)
        Vc = (Vd - torch.from_numpy(mean).to(dev)).contiguous()
        tune = autotune.tune_stage2(Vc, Fd, ratio, tol, thr, repeats=2,
                                    log=lambda m: print(m, file=sys.stderr, flush=True))
        autotune.apply(ctx, tune["settings"])
        tune = {"settings": tune["settings"],
                "stage2_ms": {k: {str(v): round(t * 1e3, 2) for v, t in d.items()} for k, d in tune["timings"].items()}}
        del Vc

    def one_run():
        # each run starts from a fresh object's stage-2 state: the context
        # keeps its allocations but restarts its persistent buffers at zero
        # (a second run() on one reference PaMO object would start from the
        # stale buffers the first left, and take different decisions)
        _hip.check(_hip.lib().pamo_ctx_reset(ctx.ptr), "pamo_ctx_reset")
        vo, fo = pm.run(Vd, Fd, ratio, tol, thr, iters, 0)
        return vo, fo, pm.face_hist

    for _ in range(args.warmup):
        one_run()
    torch.cuda.synchronize()
    print("[bench] warmup done", file=sys.stderr, flush=True)
    # untimed pass with every kernel bracketed by HIP events: per-kernel table
    prof.reset()
    prof.enable(True)
    one_run()
    torch.cuda.synchronize()
    prof.enable(False)
    stats = prof.stats()
    print("[bench] per-kernel pass done", file=sys.stderr, flush=True)
    # dominant path kernel; the generic prefix-scan and fill primitives are
    # listed in "kernels" but not chosen: they run ~13k launches per job, and
    # an event bracket around each would itself add ~30 ms to the timed run
    primitives = ("scan_tiles", "scan_sums", "other")
    dom_name = max((s for s in stats if s["name"] not in primitives), key=lambda s: s["total_ms"])["name"]
    # timed region: only the dominant kernel is bracketed
    prof.reset()
    prof.enable(True, kernels=[dom_name])
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    s2, s3 = 0.0, 0.0
    for _ in range(args.steps):
        vo, fo, hist = one_run()
        s2 += pm.timings["stage2_s"]
        s3 += pm.timings.get("stage3_s", 0.0)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    prof.enable(False)
    per_step = (t1 - t0) / args.steps
    face_rounds = int(nF0 + hist[:-1].sum())  # faces entering each round

    # roofline of the dominant kernel (live HIP-event timing inside the library)
    dom = [s for s in prof.stats() if s["name"] == dom_name][0]
    achieved = dom["bytes"] / (dom["total_ms"] * 1e-3) / 1e9 if dom["total_ms"] > 0 else 0.0
    traffic, tdetail = (None, "not measured (--no-traffic)")
    if not args.no_traffic:
        # the profiler's name is a fragment of every variant's kernel name
        # (guard_query_k, guard_query_w4_k, edge_cost_group_k, ...)
        traffic, tdetail = measure_traffic(dom["name"])
        print(f"[bench] traffic of {dom['name']}: {traffic} bytes per launch ({tdetail})", file=sys.stderr,
              flush=True)
    roofline = {"bound": "hbm", "kernel": dom["name"], "achieved": round(achieved, 1), "peak": 8000.0,
                "unit": "GB/s", "frac": round(achieved / 8000.0, 4),
                "traffic": round(traffic) if traffic is not None else None,
                "traffic_source": ("rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE over one stage-2 run of the same workload "
                                   f"(bench.py --traffic-probe), scaled by a {CALIB_COPY_MB} MiB copy")
                if traffic is not None else None,
                "traffic_detail": tdetail,
                "algorithmic_bytes_per_launch": round(dom["bytes"] / max(dom["launches"], 1)),
                "avg_launch_us": round(1e3 * dom["total_ms"] / max(dom["launches"], 1), 2),
                "launches": dom["launches"]}

    # whole-path roofline with the byte model of the hot-path contract:
    # stage 2 sum over rounds of 231 F_r + 128 V_r (+ 80 C_r, the guard's
    # candidate pairs, not counted here: a lower bound), V_r = F_r / 2 on
    # these closed meshes; stage 3 per Newton iteration 40 CG iterations of
    # 160V + 32F + 24E + 68A, one full energy pass 28V + 32F + 24E + 20K and
    # ten line-search passes 28V + 32F + 24E + 4K + 20A, the gradient pass
    # 52V + 32F + 24E + 68A and the detection 12V + 24F + 24E + 16K, plus 24V
    # per closest-point update, with the stage-3 mesh's sizes, the contacts K
    # found at each detection and the active ones A (distance below d_hat):
    # only those enter the gradient and the products, and the line-search
    # passes read the others' distances only
    Fr = np.concatenate([[nF0], hist[:-1]]).astype(np.float64)
    B2 = float((231.0 * Fr + 128.0 * (Fr / 2.0)).sum())
    B3 = 0.0
    if use_s3:
        st3 = pm.system.stats()
        v3, f3 = pm.stage3_in
        Vs, Fs = float(len(v3)), float(len(f3))
        Es = 1.5 * Fs
        n_nt, K, A = float(st3["newton"]), float(st3["contacts"]), float(pm.system.active_contacts())
        per_newton = (40 * (160 * Vs + 32 * Fs + 24 * Es) + 11 * (28 * Vs + 32 * Fs + 24 * Es)
                      + (52 * Vs + 32 * Fs + 24 * Es) + (12 * Vs + 24 * Fs + 24 * Es))
        B3 = n_nt * per_newton + K * (20 + 10 * 4 + 16) + A * (40 * 68 + 10 * 20 + 68) + 5 * 24 * Vs
    path_ach = (B2 + B3) / per_step / 1e9
    path_roofline = {"bound": "hbm", "achieved": round(path_ach, 2), "peak": 8000.0, "unit": "GB/s",
                     "frac": round(path_ach / 8000.0, 5), "stage2_bytes": round(B2), "stage3_bytes": round(B3),
                     "model": "stage 2: sum_r 231 F_r + 128 V_r (V_r = F_r / 2; the 80 C_r guard-pair term not "
                              "counted); stage 3 per Newton iteration: 40 CG passes over the active contacts A, "
                              "one full and ten line-search energy passes, the gradient pass and the detection "
                              "over the K contacts found"}

    # output quality, untimed: igl-semantics Hausdorff and sampled Chamfer
    # distances of the output to the input surface (GPU metrics)
    from pamo_amd import metrics
    tq = time.perf_counter()
    cd, hd = metrics.compute_igl_CD_HD(V, F, vo, fo)
    quality = {"hausdorff": float(hd), "chamfer": float(cd), "bbox_diag": float(np.linalg.norm(V.max(0) - V.min(0))),
               "metric_s": round(time.perf_counter() - tq, 3)}
    print(f"[bench] quality: Hausdorff {hd:.4e}, Chamfer {cd:.4e}", file=sys.stderr, flush=True)

    # CPU baseline, stages 2+3, on this box's host cores: the whole stage-2
    # oracle run (its output doubles as a bit-exactness check of this run's
    # stage 2), and the stage-3 oracle timed for one and two Newton
    # iterations of a step on the stage-2 output, scaled to the job's five
    # steps of ten iterations.
    cpu, parity = None, None
    if not args.no_cpu_baseline:
        cpu, parity = cpu_baseline(args, pm, V, F, ratio, tol, thr, iters, use_s3)

    out = {
        "metric": "2M->20k faces wall-time (stages 2+3)" if use_s3 else "2M->20k faces wall-time (stage 2)",
        "value": round(per_step, 4),
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(per_step * 1e3, 2),
        "higher_is_better": False,
        "scaling": "strong",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "ellipsoid/box union, marching tetrahedra grid 296", "faces_in": nF0,
                   "verts_in": nV0, "faces_out": int(len(fo)), "ratio": ratio, "tolerance": tol,
                   "threshold": thr, "rounds": int(len(hist)), "stage3": use_s3,
                   "guard": "exact" if args.exact_guard else "reference predicate"},
        "stage2_s": round(s2 / args.steps, 4),
        "stage3_s": round(s3 / args.steps, 4),
        "faces_per_s": round(nF0 / per_step, 1),
        "face_rounds": face_rounds,
        "roofline": roofline,
        "path_roofline": path_roofline,
        "cpu_baseline": cpu,
        "stage2_vs_oracle": parity,
        "autotune": tune,
        "quality": quality,
        "kernels": stats,
    }
    if not args.no_legs:
        out["legs"] = legs(args, V, F)
    if args.calib_copy_mb > 0:
        # a plain copy of known size: its FETCH_SIZE / WRITE_SIZE counters
        # convert the counters of the kernels above into bytes
        src = torch.empty(args.calib_copy_mb << 18, dtype=torch.float32, device=dev).fill_(1.0)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        out["calib_copy_bytes"] = 2 * src.numel() * 4
    print(json.dumps(out))


if __name__ == "__main__":
    main()
