"""Benchmark (one GPU): primary Mrays/s @1080p on the 2^15 coloured terrain; edited voxels/s of
4096 mixed edits batched (one fused pass) vs plain (one pass per edit), each followed by GC, on
the 2^16 coloured terrain; p50 / p99 latency of a served request stream on the 2^16 terrain.

Workload (one step = one frame for one camera): trace primary paths, resolve colours and
one shadow ray per pixel at 1920x1080, cameras cycling over 64 seeded views placed 5-20 %
of the extent above the terrain. The scene (seeded simplex heightfield shell, per-voxel
RGB8 colours in compressed leaves) is generated and loaded before timing; the timed region
runs with everything resident in HBM.

Prints one JSON line with the metric, a roofline object for the dominant kernel (algorithmic
bytes counted by an instrumented launch of the same kernels, divided by the average kernel
time measured with events on the launch stream) and a CPU baseline (the oracle's host
traversal on a bounded sample of the same frames, all host cores).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=64)
    ap.add_argument("--warmup", type=int, default=4)
    ap.add_argument("--levels", type=int, default=15)
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--cameras", type=int, default=64)
    ap.add_argument("--seed", type=int, default=1337)
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="budget of the CPU baseline sample")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--tile", type=str, default="auto", help="wavefront tile WxH, or auto (autotuned)")
    ap.add_argument("--schedule", type=str, default="auto",
                    help="tracer schedule: tile (one wavefront per tile), persist[:N] (persistent wavefronts "
                         "refilled at N idle lanes), beam[:N] (per tile, walks beyond N steps searched a wavefront "
                         "per ray), resume[:A[:B]] (per tile to A steps, resumed by compacted wavefronts to B, the "
                         "rest searched per ray), or auto (the autotuner's pick with --tile auto, else tile)")
    ap.add_argument("--edits", type=int, default=4096, help="mixed edits of the batched-vs-plain comparison")
    ap.add_argument("--edit-cpu-seconds", type=float, default=10.0, help="budget of the CPU edit baseline sample")
    ap.add_argument("--no-edits", action="store_true")
    ap.add_argument("--bucket-scale", type=int, default=0,
                    help="edit scene buckets: N times the reference's bucket counts at 1/N the size (1 = the "
                         "reference's), or 0 to time 1, 4, 8 and keep the fastest")
    ap.add_argument("--no-plain", action="store_true",
                    help="skip the plain leg of the edit comparison (for counter runs of the fused pass)")
    ap.add_argument("--edit-forms", choices=("fast", "all", "level"), default="fast",
                    help="edit launch forms compared: fast (a launch per step and level, a captured single-edit "
                         "graph per plain edit, one cooperative launch for the fused pass, one workgroup on a "
                         "sample by radius), all (also one workgroup per plain edit over the whole list) or "
                         "level (the first only; for profilers that cannot follow a cooperative launch)")
    ap.add_argument("--edit-levels", type=int, default=16,
                    help="scene depth of the batched-vs-plain edit comparison (0 = the traced scene)")
    ap.add_argument("--serve-requests", type=int, default=400, help="requests of the served stream (0 = skip)")
    ap.add_argument("--serve-rate", type=float, default=100.0, help="Poisson arrival rate, requests/s")
    ap.add_argument("--no-replays", action="store_true",
                    help="skip serving the reference's benchmark replays (edits, small / large edits, copies)")
    ap.add_argument("--c5", action="store_true",
                    help="also run the depth-17 workload: instancing, 512^3 copies, large carves in a frame graph, "
                         "undo kept across a full GC, a 3840x2160 frame")
    args = ap.parse_args()

    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1"}))
        print("bench.py measures a single MI355X; multi-GPU runs are not part of this build", file=sys.stderr)
        return 2

    import numpy as np
    import torch

    from hashdag_amd import DAGInfo, DAGTracer, Geometry, HashDAG, TerrainParams, generate_terrain, ray_params
    from hashdag_amd._lib import check, lib
    from hashdag_amd.workload import cameras

    W, H, D = args.width, args.height, args.levels
    t0 = time.time()
    params = TerrainParams(levels=D, seed=args.seed, colors=True)
    scene = generate_terrain(params)
    log(f"scene 2^{D}: {scene.dag.size} BasicDAG words, {scene.num_voxels} voxels, "
        f"{scene.blocks.size} colour blocks, generated in {time.time() - t0:.1f}s")

    geom = fit_geometry(D, scene.level_words)
    log(f"bucket geometry: {geom.num_top_levels} top-geometry levels, {geom.total_virtual()} virtual words; "
        f"level words {[int(w) for w in scene.level_words]}")
    t0 = time.time()
    dag = HashDAG(geom)
    dag.load_scene(scene)
    info = dag.info()
    log(f"loaded into the device hash table in {time.time() - t0:.1f}s (pool_top {info.pool_top} pages)")
    dinfo = DAGInfo.unit(D)
    cams = cameras(params, args.cameras, seed=args.seed)
    tracer = DAGTracer(W, H)
    rparams = [ray_params(c, dinfo, D, W, H) for c in cams]
    if args.tile == "auto":
        from hashdag_amd.tracer import autotune
        tuned = autotune(tracer, dag, rparams[:4])
        args.tile = tuned["tile"]
        log(f"autotuned tracer tile: {tuned}")
    else:
        tw, th = (int(v) for v in args.tile.split("x"))
        tracer.set_tile(dag, tw, th)
    if args.schedule.startswith("persist"):
        refill = int(args.schedule.split(":")[1]) if ":" in args.schedule else 16
        tracer.set_schedule(dag, 1, refill)
    elif args.schedule.startswith("beam"):
        cap = int(args.schedule.split(":")[1]) if ":" in args.schedule else 128
        tracer.set_schedule(dag, 2, cap)
    elif args.schedule.startswith("resume"):
        parts = args.schedule.split(":")
        c1 = int(parts[1]) if len(parts) > 1 else 32
        c2 = int(parts[2]) if len(parts) > 2 else 128
        tracer.set_schedule(dag, 3, c1, resume_to=c2)
    elif args.schedule == "tile":
        tracer.set_schedule(dag, 0)
    from hashdag_amd.tracer import launch_config
    launch = launch_config(dag)
    log(f"tracer launch: {launch}")

    stream = torch.cuda.current_stream()

    def frame(i: int, ev=None):
        p = rparams[i % len(rparams)]
        if ev is not None:
            ev[0].record(stream)
        tracer.resolve_paths(None, dinfo, dag, params=p)
        if ev is not None:
            ev[1].record(stream)
        tracer.resolve_colors(dag)
        if ev is not None:
            ev[2].record(stream)
        tracer.resolve_shadows(None, dinfo, dag, 1.0, 0.0, params=p)
        if ev is not None:
            ev[3].record(stream)

    for i in range(args.warmup):
        frame(i)
    torch.cuda.synchronize()

    # algorithmic bytes of the frames the timed region runs (instrumented launches, untimed)
    alg = np.zeros(3, np.float64)
    ncount = min(args.steps, len(rparams))
    for i in range(ncount):
        alg += np.array(tracer.count_bytes(dag, rparams[i % len(rparams)]), np.float64)
    alg /= ncount
    torch.cuda.synchronize()

    events = [[torch.cuda.Event(enable_timing=True) for _ in range(4)] for _ in range(args.steps)]
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    for i in range(args.steps):
        frame(i, events[i])
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t_start

    k_ms = np.zeros(3)
    for ev in events:
        k_ms += [ev[0].elapsed_time(ev[1]), ev[1].elapsed_time(ev[2]), ev[2].elapsed_time(ev[3])]
    k_ms /= args.steps
    names = ["trace_paths", "trace_colors", "trace_shadows"]
    dom = int(np.argmax(k_ms))
    achieved = alg[dom] / (k_ms[dom] * 1e-3) / 1e9

    # achievable streaming bandwidth on this device (read + write of 2 GiB)
    nbytes = 1 << 31
    src = torch.empty(nbytes // 4, dtype=torch.int32, device="cuda")
    dst = torch.empty_like(src)
    import ctypes as C
    for _ in range(3):
        check(lib.hd_stream_copy(C.c_void_p(src.data_ptr()), C.c_void_p(dst.data_ptr()), nbytes,
                                 C.c_void_p(stream.cuda_stream)))
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    reps = 10
    for _ in range(reps):
        check(lib.hd_stream_copy(C.c_void_p(src.data_ptr()), C.c_void_p(dst.data_ptr()), nbytes,
                                 C.c_void_p(stream.cuda_stream)))
    e1.record(stream)
    torch.cuda.synchronize()
    stream_gbs = 2 * nbytes * reps / (e0.elapsed_time(e1) * 1e-3) / 1e9
    # and a read-only pass over 4 GiB (the tracer kernels read far more than they write)
    rsrc = torch.cat([src, dst])
    fold = torch.empty(1 << 16, dtype=torch.int32, device="cuda")
    for _ in range(3):
        check(lib.hd_stream_read(C.c_void_p(rsrc.data_ptr()), 2 * nbytes, C.c_void_p(fold.data_ptr()),
                                 C.c_void_p(stream.cuda_stream)))
    e0.record(stream)
    for _ in range(reps):
        check(lib.hd_stream_read(C.c_void_p(rsrc.data_ptr()), 2 * nbytes, C.c_void_p(fold.data_ptr()),
                                 C.c_void_p(stream.cuda_stream)))
    e1.record(stream)
    torch.cuda.synchronize()
    read_gbs = 2 * nbytes * reps / (e0.elapsed_time(e1) * 1e-3) / 1e9
    del src, dst, rsrc, fold

    value = W * H * args.steps / elapsed / 1e6
    result = {
        "metric": "primary Mrays/s @1080p; edited voxels/s batched vs plain; p99 request latency",
        "value": round(value, 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"2^{D} coloured simplex terrain, {W}x{H} primary paths + colours + 1 shadow ray per pixel, "
                        f"one camera per step: cameras 0..{min(args.steps, len(cams)) - 1} of {len(cams)} seeded "
                        f"cameras" + (f" (cycled {args.steps / len(cams):.2f} times)" if args.steps > len(cams) else ""),
            "levels": D, "width": W, "height": H, "cameras_timed": min(args.steps, len(cams)),
            "cameras_generated": len(cams), "seed": args.seed,
            "voxels": int(scene.num_voxels), "tile": launch["tile"], "fetch": launch["fetch"],
            "schedule": launch["schedule"],
            "bucket_geometry": {"num_top_levels": geom.num_top_levels,
                                "top": f"{1 << geom.bucket_bits_top} x {geom.bucket_size_top} words",
                                "low": f"{1 << geom.bucket_bits_low} x {geom.bucket_size_low} words",
                                "page_words": geom.page_size},
        },
        "kernels_ms": {n: round(float(v), 4) for n, v in zip(names, k_ms)},
        "primary_only_mrays_s": round(W * H / (k_ms[0] * 1e-3) / 1e6, 2),
        "algorithmic_bytes_per_launch": {n: int(v) for n, v in zip(names, alg)},
        "roofline": {
            "bound": "hbm",
            "kernel": names[dom],
            "achieved": round(float(achieved), 2),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(float(achieved) / HBM_PEAK_GBS, 4),
            "traffic": measured_traffic(names[dom], D, W, H, launch),
            "traffic_source": "profiles/trace_kernels_2p15_1080p.md (rocprofv3 FETCH_SIZE + WRITE_SIZE, copy-calibrated)",
            "achievable_stream_gbs": round(stream_gbs, 1),
            "achievable_read_gbs": round(read_gbs, 1),
        },
    }

    dag.close()
    if not args.no_cpu_baseline:
        od, oc, result["cpu_baseline"] = cpu_baseline(scene, geom, rparams, W, H, args.cpu_seconds)
        if not args.no_edits:
            edit_cpu = cpu_edit_baseline(od, oc, params, args)
        del od, oc
    if not args.no_edits or args.serve_requests > 0 or not args.no_replays:
        ED = args.edit_levels or D
        if ED != D:
            # the batched-vs-plain comparison on its own (deeper) scene
            del scene
            eparams = TerrainParams(levels=ED, seed=args.seed, colors=True)
            t0 = time.time()
            escene = generate_terrain(eparams)
            egeom = fit_geometry(ED, escene.level_words)
            log(f"edit scene 2^{ED}: {escene.num_voxels} voxels, generated in {time.time() - t0:.1f}s; "
                f"buckets top {egeom.bucket_size_top} low {egeom.bucket_size_low} words")
        else:
            eparams, escene, egeom = params, scene, geom
        bucket_trials = None
        if args.bucket_scale > 0:
            egeom = fit_geometry(eparams.levels, escene.level_words, bucket_scale=args.bucket_scale)
        elif not args.no_edits:
            egeom, bucket_trials = tune_buckets(escene, eparams, args)
        if not args.no_edits:
            result["edits"] = edit_bench(escene, egeom, eparams, W, H, args)
            result["edits"]["bucket_geometry"] = {"low": f"{1 << egeom.bucket_bits_low} x {egeom.bucket_size_low} words",
                                                  "top": f"{1 << egeom.bucket_bits_top} x {egeom.bucket_size_top} words",
                                                  "top_levels": egeom.num_top_levels,
                                                  "fused_ms_by_geometry": bucket_trials}
            if not args.no_cpu_baseline:
                result["edits"]["cpu_baseline"] = edit_cpu
        if args.serve_requests > 0:
            result["serving"] = serve_bench(escene, egeom, eparams, args)
        if not args.no_replays:
            result["replays"] = replay_bench(escene, egeom, eparams, W, H, args)
    if args.c5:
        result["c5"] = c5_bench(args)
    print(json.dumps(result), flush=True)
    return 0


def fit_geometry(levels, level_words, bucket_scale=1):
    """Bucket geometry with 50% headroom for edits when it fits, else the tightest that does."""
    from hashdag_amd import Geometry
    try:
        return Geometry.fit(levels, level_words, headroom=1.5, bucket_scale=bucket_scale)
    except ValueError:
        return Geometry.fit(levels, level_words, headroom=1.0, bucket_scale=bucket_scale)


def tune_buckets(scene, params, args):
    """The bucket-count axis: the edit list's fused pass timed on the scene loaded with the
    reference's bucket counts and with 4x / 8x as many buckets of 1/4 / 1/8 the size
    (the same capacity; a find scans a shorter bucket); a geometry whose buckets overflow drops
    out. Returns the fastest geometry and the trials."""
    from hashdag_amd._lib import HashDAGError
    import torch

    from hashdag_amd import HashDAG
    from hashdag_amd.workload import mixed_edits
    edits = mixed_edits(params, args.edits, seed=args.seed + 1)
    trials, best = {}, None
    for scale in (1, 4, 8):
        try:
            g = fit_geometry(params.levels, scene.level_words, bucket_scale=scale)
        except ValueError:
            continue
        d = HashDAG(g)
        try:
            d.load_scene(scene)
            d.edit_batch(edits, fused=True)  # sizes the scratch
            for _ in range(d.undo_depth()[0]):
                d.undo()
            d.remove_stale_nodes()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            d.edit_batch(edits, fused=True)
            torch.cuda.synchronize()
            ms = (time.perf_counter() - t0) * 1e3
        except HashDAGError as e:
            log(f"bucket geometry x{scale}: {e}")
            continue
        finally:
            d.close()
        key = f"{1 << g.bucket_bits_low} x {g.bucket_size_low} words (low levels)"
        trials[key] = round(ms, 2)
        log(f"bucket geometry x{scale}: {key}, top {1 << g.bucket_bits_top} x {g.bucket_size_top}: "
            f"fused pass {ms:.1f} ms")
        if best is None or ms < best[0]:
            best = (ms, g)
    return best[1], trials


def edit_bench(scene, geom, params, W, H, args):
    """Edited voxels/s of the same seeded mixed edit list applied one pass per edit (plain)
    and as one fused pass (batched), each on a freshly loaded copy of the scene with the edit
    scratch already sized by an untimed run (undone and garbage-collected before timing).
    The two results must trace to identical frames."""
    import numpy as np
    import torch

    from hashdag_amd import DAGInfo, DAGTracer, HashDAG, ray_params
    from hashdag_amd.workload import cameras, mixed_edits

    edits = mixed_edits(params, args.edits, seed=args.seed + 1)
    dinfo = DAGInfo.unit(params.levels)
    rparams = [ray_params(c, dinfo, params.levels, W, H) for c in cameras(params, 2, seed=args.seed)]
    log(f"edits: {len(edits)} mixed edits (70% sphere, 20% box, 10% paint; radius log-uniform 2-256)")

    def frames(d):
        tr = DAGTracer(W, H)
        out = []
        for p in rparams:
            tr.resolve_paths(None, dinfo, d, params=p)
            tr.resolve_colors(d)
            tr.resolve_shadows(None, dinfo, d, 1.0, 0.0, params=p)
            torch.cuda.synchronize()
            out.append(tr.colors.cpu().numpy().copy())
        return out

    def timed(d, fused, warm_edits, launch=(0, 0)):
        """The edit list on the DAG as it is, after an untimed run that sizes the scratch (undone,
        and its nodes collected, so the timed run finds the table as loaded). launch: passes of
        at most launch[0] edits run in one workgroup, else of at most launch[1] edits as one
        cooperative launch, else a launch per step and level."""
        d.set_edit_launch(*launch)
        d.edit_batch(warm_edits, fused=fused)
        for _ in range(d.undo_depth()[0]):
            d.undo()
        d.remove_stale_nodes()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        st = d.edit_batch(edits, fused=fused)
        torch.cuda.synchronize()
        return st, time.perf_counter() - t0

    def gc_timed(d):
        """The GC pass that follows an edit list (mark / compact of the stale nodes)."""
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        gc = d.remove_stale_nodes()
        torch.cuda.synchronize()
        return time.perf_counter() - t0, gc

    def undo_all(d):
        for _ in range(d.undo_depth()[0]):
            d.undo()

    def plain_by_radius(d, sample):
        """Median milliseconds of one plain pass by editor radius, a launch per step and level
        against one single-workgroup launch, on a sample of the edit list (each edit undone)."""
        classes = ((2, 8), (8, 32), (32, 128), (128, 257))
        out = {}
        for name, launch in (("level_synchronous", (0, 0)), ("single_workgroup", (1 << 30, 0))):
            d.set_edit_launch(*launch)
            for e in sample[:8]:
                d.edit(*e)
            undo_all(d)
            times = []
            for kind, center, r in sample:
                t0 = time.perf_counter()
                d.edit(kind, center, r)
                times.append((r, time.perf_counter() - t0))
                d.undo()
            out[name] = {f"r{lo}-{hi - 1}": round(float(np.median([t for r, t in times if lo <= r < hi])) * 1e3, 3)
                         for lo, hi in classes if any(lo <= r < hi for r, _ in times)}
        d.set_edit_launch(0, 0)
        return out

    # fused: one launch per step and level, then (after undo) one cooperative launch
    dag = HashDAG(geom)
    dag.load_scene(scene)
    st_f, t_f = timed(dag, True, edits)
    img_f = frames(dag)
    # algorithmic bytes of the same pass, counted on a replay from the loaded state (the
    # counting kernels are slower, so the timed runs do not count)
    undo_all(dag)
    dag.remove_stale_nodes()
    dag.count_edit_bytes(True)
    dag.edit_batch(edits, fused=True)
    ab = dag.count_edit_bytes(False)
    undo_all(dag)
    if args.edit_forms != "level":
        st_fp, t_fp = timed(dag, True, edits, launch=(0, 1 << 30))
        img_fp = frames(dag)
    else:
        st_fp, t_fp, img_fp = st_f, float("inf"), img_f
    gc_f, gcs_f = gc_timed(dag)
    dag.close()
    abytes = sum(ab.values())
    log(f"fused: {st_f.num_voxels} voxels in {t_f * 1e3:.1f} ms ({st_f.passes} pass; {t_fp * 1e3:.1f} ms as one "
        f"cooperative launch), then GC {gc_f * 1e3:.1f} ms ({gcs_f['words_before']} -> {gcs_f['words_after']} words)")
    log(f"fused pass algorithmic bytes: {ab} = {abytes / 1e9:.3f} GB -> {abytes / t_f / 1e9:.1f} GB/s")
    if st_fp.num_voxels != st_f.num_voxels or not all(np.array_equal(a, b) for a, b in zip(img_f, img_fp)):
        raise RuntimeError("fused pass: cooperative and level-synchronous launches differ")
    if t_fp < t_f:
        st_b, t_b, batched_form = st_fp, t_fp, "one cooperative launch"
    else:
        st_b, t_b, batched_form = st_f, t_f, "a launch per step and level"

    roofline = {"bound": "hbm", "achieved": round(abytes / t_f / 1e9, 2), "peak": 8000.0, "unit": "GB/s",
                "frac": round(abytes / t_f / 8e12, 5),
                "traffic": measured_edit_traffic(params.levels, len(edits), args.seed + 1, geom),
                "traffic_source": "profiles/edit_kernels_2p16_4096edits.md (rocprofv3 FETCH_SIZE + WRITE_SIZE of the "
                                  "fused pass, copy-calibrated)",
                "bytes": ab, "note": "algorithmic bytes of the fused pass (node words read by the descent, "
                "bucket bytes read by the finds, node words written, staged colour markers and explicit "
                "colours written once and read twice, colour leaf words written) over the wall time of its "
                "level-synchronous form including host steps"}
    workload = (f"{len(edits)} mixed edits on the 2^{params.levels} coloured terrain (70% sphere add/carve, "
                f"20% box add/carve, 10% sphere paint; radius log-uniform 2-256; seed {args.seed + 1})")
    if args.no_plain:
        return {"workload": workload, "edited_voxels": int(st_b.num_voxels),
                "batched_voxels_per_s": round(st_b.num_voxels / t_b, 1), "batched_ms": round(t_b * 1e3, 2),
                "batched_form": batched_form, "gc_after_batched_ms": round(gc_f * 1e3, 2),
                "plain": "not run (--no-plain)", "roofline": roofline}

    # plain: one pass per edit, in every launch form; the fastest is the baseline
    plain = HashDAG(geom)
    plain.load_scene(scene)
    inf = float("inf")
    nv_pp, t_pp, img_pp = None, inf, None
    nv_pg, t_pg, img_pg, replays = None, inf, None, 0
    nv_pt, t_pt, img_pt, tuned = None, inf, None, None
    by_radius = None
    if args.edit_forms in ("all", "fast"):
        # the launch autotuner on a sample of the list, then the list with its pick
        tuned = plain.autotune_edit_launch(edits[::64], reps=1)
        plain.remove_stale_nodes()
        log(f"edit launch autotuner: {tuned}")
        st_pt, t_pt = timed(plain, False, edits[:16], launch=(tuned["workgroup_max_items"], 0))
        nv_pt = st_pt.num_voxels
        img_pt = frames(plain)
        undo_all(plain)
        # a captured single-edit pass replayed per edit (HIP graph, edit-only)
        from hashdag_amd.graph import FrameGraph
        plain.set_edit_launch(0, 0)
        fg = FrameGraph(plain, None, 1)
        for e in edits[:16]:
            fg.launch([e])
        undo_all(plain)
        plain.remove_stale_nodes()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        nv_pg = 0
        for e in edits:
            st, ran = fg.launch([e])
            nv_pg += st.num_voxels
            replays += ran
        torch.cuda.synchronize()
        t_pg = time.perf_counter() - t0
        img_pg = frames(plain)
        fg.close()
        undo_all(plain)
        by_radius = plain_by_radius(plain, edits[::16])
    if args.edit_forms == "all":
        st_pp, t_pp = timed(plain, False, edits[:16], launch=(1 << 30, 0))
        nv_pp = st_pp.num_voxels
        img_pp = frames(plain)
        undo_all(plain)
    # the level-synchronous form last, so the GC timed after it finds this list's stale nodes
    st_pl, t_pl = timed(plain, False, edits[:16])
    img_pl = frames(plain)
    gc_p, _ = gc_timed(plain)
    plain.close()
    log(f"plain: {st_pl.num_voxels} voxels in {t_pl * 1e3:.1f} ms ({st_pl.passes} passes, a launch per step and "
        f"level), {t_pt * 1e3:.1f} ms (autotuned: small passes in one workgroup), {t_pp * 1e3:.1f} ms (one "
        f"single-workgroup launch per edit), {t_pg * 1e3:.1f} ms (a captured single-edit graph replayed per edit, "
        f"{replays} replays), then GC {gc_p * 1e3:.1f} ms")
    plain_forms = {"a launch per step and level": t_pl, "one single-workgroup launch per edit": t_pp,
                   "a captured single-edit pass replayed per edit": t_pg,
                   "small passes in one workgroup (autotuned), the rest a launch per step": t_pt}
    plain_form = min(plain_forms, key=plain_forms.get)
    t_p = plain_forms[plain_form]
    st_p = st_pl
    same = st_pl.num_voxels == st_f.num_voxels and all(np.array_equal(a, b) for a, b in zip(img_f, img_pl))
    for nv, img in ((nv_pp, img_pp), (nv_pg, img_pg), (nv_pt, img_pt)):
        if img is not None:
            same = same and nv == st_f.num_voxels and all(np.array_equal(a, b) for a, b in zip(img_f, img))
    if not same:
        raise RuntimeError("fused and plain edit results differ")
    return {
        "workload": workload,
        "edited_voxels": int(st_b.num_voxels),
        "plain_voxels_per_s": round(st_p.num_voxels / t_p, 1),
        "batched_voxels_per_s": round(st_b.num_voxels / t_b, 1),
        "batched_over_plain": round(t_p / t_b, 2),
        "plain_ms": round(t_p * 1e3, 2),
        "batched_ms": round(t_b * 1e3, 2),
        "plain_ms_per_edit": round(t_p * 1e3 / len(edits), 3),
        "plain_form": plain_form,
        "batched_form": batched_form,
        "plain_ms_by_form": {"level_synchronous": round(t_pl * 1e3, 2),
                             "single_workgroup": round(t_pp * 1e3, 2) if t_pp < float("inf") else None,
                             "graph_replay": round(t_pg * 1e3, 2) if t_pg < float("inf") else None,
                             "small_passes_in_a_workgroup": round(t_pt * 1e3, 2) if t_pt < float("inf") else None},
        "edit_launch_autotuner": tuned,
        "plain_pass_ms_by_radius": by_radius,
        "batched_ms_by_form": {"level_synchronous": round(t_f * 1e3, 2),
                               "cooperative": round(t_fp * 1e3, 2) if t_fp < float("inf") else None},
        "gc_after_batched_ms": round(gc_f * 1e3, 2),
        "gc_after_plain_ms": round(gc_p * 1e3, 2),
        "gc_words_before_after": [gcs_f["words_before"], gcs_f["words_after"]],
        "batched_with_gc_voxels_per_s": round(st_b.num_voxels / (t_b + gc_f), 1),
        "batched_over_plain_with_gc": round((t_p + gc_p) / (t_b + gc_f), 2),
        "results_identical": same,
        "roofline": roofline,
    }


def serve_bench(scene, geom, params, args):
    """A served request stream on a fresh copy of the scene: Poisson arrivals, 60% trace
    requests (640x360 .. 3840x2160, seeded cameras), 40% edit requests (1-64 sphere / box
    editors, radius log-uniform 2-128); p50 / p99 latency beside throughput."""
    from hashdag_amd import HashDAG
    from hashdag_amd.serve import latency_summary, request_stream, serve

    reqs = request_stream(params, args.serve_requests, args.serve_rate, seed=args.seed + 2)
    d = HashDAG(geom)
    d.load_scene(scene)
    serve(d, request_stream(params, 8, 1000.0, seed=args.seed + 3), realtime=False)  # warm-up
    res = serve(d, reqs, trace_slots=4, realtime=True)
    d.close()
    summ = latency_summary(reqs, res)
    log(f"serving: {summ}")
    summ["workload"] = (f"{len(reqs)} requests, Poisson {args.serve_rate}/s on the 2^{params.levels} coloured terrain: "
                        f"60% trace (640x360..3840x2160), 40% edit (1-64 sphere/box, radius log-uniform 2-128)")
    return summ


REPLAYS = ("epiccitadel_edits", "epiccitadel_small_edits", "epiccitadel_large_edits", "epiccitadel_copy")


def replay_bench(scene, geom, params, W, H, args):
    """The reference's benchmark replays (tests/golden/replays, recorded at depth 17) served on
    the terrain at 60 frames/s in real time: each frame's edits (scaled to the scene depth) and
    a W x H frame traced from a seeded camera over this terrain (the replays' own cameras
    belong to their scenes). Per replay: p50 / p99 latency of frames and edits, edited voxels."""
    from hashdag_amd import DAGInfo, HashDAG, ray_params
    from hashdag_amd.replay import Replay, replay_requests
    from hashdag_amd.serve import latency_summary, serve
    from hashdag_amd.workload import cameras

    dinfo = DAGInfo.unit(params.levels)
    prms = [ray_params(c, dinfo, params.levels, W, H) for c in cameras(params, 16, seed=args.seed + 4)]
    out = {}
    d = HashDAG(geom)
    d.load_scene(scene)
    for name in REPLAYS:
        rep = Replay.load(os.path.join(ROOT, "tests", "golden", "replays", name + ".csv"), params.levels)
        reqs = replay_requests(rep, W, H, params.levels, fps=60.0, params=prms)
        res = serve(d, reqs, trace_slots=4, realtime=True, keep_undo=True)
        # back to the loaded scene for the next replay
        for _ in range(d.undo_depth()[0]):
            d.undo()
        d.remove_stale_nodes()
        summ = latency_summary(reqs, res)
        summ["frames"] = rep.frames
        summ["replay_edits"] = rep.counts()["edits"]
        summ["rejected"] = int((res["status"] != 0).sum())
        log(f"replay {name}: {summ}")
        out[name] = summ
    d.close()
    out["workload"] = (f"the reference's replays {', '.join(REPLAYS)} at 60 frames/s on the 2^{params.levels} "
                       f"terrain, edits scaled from depth 17, {W}x{H} frames from seeded cameras")
    return out


def c5_bench(args):
    """Depth-17 coloured terrain (a 65536^2 footprint: the full 131072^2 footprint needs about
    1.3 G words at one level, beyond what the u32 virtual address space holds with the uniform
    bucket geometry of the low levels), then, each phase timed on its own:
      2000 instanced boxes and spheres standing on the terrain (heights 16-512), one fused pass;
      32 copies of 512^3 regions to seeded destinations (one pass each);
      64 sphere carves of radius 512-1024 (eagerly once, undone), then the same carves and a
      3840x2160 frame as one replay of a captured graph;
      a full GC that keeps the undo history;
      3840x2160 primary + colour + shadow frames.
    Checked on the way: GC leaves the traced frame unchanged and the undo history usable (undo
    then redo of the carve pass reproduces the frame)."""
    import numpy as np
    import torch

    from hashdag_amd import DAGInfo, DAGTracer, HashDAG, TerrainParams, generate_terrain, ray_params
    from hashdag_amd.graph import FrameGraph
    from hashdag_amd.serve import frame_checksum
    from hashdag_amd.workload import cameras, carves, copy_edits, instances

    D, W, H = 17, 3840, 2160
    params = TerrainParams(levels=D, footprint=1 << 16, seed=args.seed, colors=True)
    t0 = time.time()
    scene = generate_terrain(params)
    geom = fit_geometry(D, scene.level_words)
    log(f"c5 scene 2^{D} (footprint 2^16): {scene.num_voxels} voxels, {scene.dag.size} BasicDAG words, generated in "
        f"{time.time() - t0:.1f}s; buckets top {geom.bucket_size_top} low {geom.bucket_size_low} words, "
        f"{geom.total_virtual()} virtual words")
    t0 = time.time()
    dag = HashDAG(geom)
    dag.load_scene(scene)
    del scene
    log(f"c5 loaded in {time.time() - t0:.1f}s")
    dinfo = DAGInfo.unit(D)
    cam = cameras(params, 1, seed=args.seed)[0]
    prm = ray_params(cam, dinfo, D, W, H)
    tr = DAGTracer(W, H)
    out = {"workload": "depth 17, 65536^2 coloured terrain footprint; 2000 instances (heights 16-512), 32 copies "
                       "of 512^3 regions, 64 carves (radius 512-1024) + 3840x2160 frame in one graph replay, "
                       "GC keeping undo, 3840x2160 frames", "phases_ms": {}, "voxels": {}}

    def timed(name, fn):
        torch.cuda.synchronize()
        t = time.perf_counter()
        r = fn()
        torch.cuda.synchronize()
        out["phases_ms"][name] = round((time.perf_counter() - t) * 1e3, 2)
        return r

    st = timed("instances_2000", lambda: dag.edit_batch(instances(params, 2000, seed=args.seed + 5), fused=True))
    out["voxels"]["instances_2000"] = int(st.num_voxels)
    log(f"c5 instances: {st.num_voxels} voxels in {out['phases_ms']['instances_2000']} ms")
    st = timed("copies_32", lambda: dag.edit_batch(copy_edits(params, 32, seed=args.seed + 6), fused=True))
    out["voxels"]["copies_32"] = int(st.num_voxels)
    out["copy_passes"] = int(st.passes)
    log(f"c5 copies: {st.num_voxels} voxels in {out['phases_ms']['copies_32']} ms ({st.passes} passes)")
    carve_list = carves(params, 64, seed=args.seed + 7)
    # the carves once eagerly (sizes the edit scratch the graph bakes in), then undone
    st = timed("carves_64_eager", lambda: dag.edit_batch(carve_list, fused=True))
    dag.undo()
    fg = FrameGraph(dag, tr, 64)
    timed("graph_capture_empty_batch", lambda: fg.launch([], prm))
    st, replayed = timed("carves_64_and_frame_graph", lambda: fg.launch(carve_list, prm))
    out["voxels"]["carves_64"] = int(st.num_voxels)
    out["graph_replayed"] = bool(replayed)
    log(f"c5 carves + frame (graph replay {replayed}): {st.num_voxels} voxels in "
        f"{out['phases_ms']['carves_64_and_frame_graph']} ms")
    fg.close()
    before = frame_checksum(dag, tr.colors, W, H)
    depth = dag.undo_depth()
    gc = timed("gc_keep_undo", lambda: dag.remove_stale_nodes(keep_undo=True))
    out["gc_words_before_after"] = [gc["words_before"], gc["words_after"]]
    out["undo_depth"] = list(depth)
    log(f"c5 GC: {gc['words_before']} -> {gc['words_after']} words in {out['phases_ms']['gc_keep_undo']} ms, "
        f"undo depth {depth} -> {dag.undo_depth()}")

    def frame():
        tr.resolve_paths(None, dinfo, dag, params=prm)
        tr.resolve_colors(dag)
        tr.resolve_shadows(None, dinfo, dag, 1.0, 0.0, params=prm)

    frame()
    after = frame_checksum(dag, tr.colors, W, H)
    reps = 8
    timed("frames_4k_x8", lambda: [frame() for _ in range(reps)])
    out["frame_mrays_per_s"] = round(W * H * reps / (out["phases_ms"]["frames_4k_x8"] * 1e-3) / 1e6, 1)
    dag.undo()
    frame()
    undone = frame_checksum(dag, tr.colors, W, H)
    dag.redo()
    frame()
    redone = frame_checksum(dag, tr.colors, W, H)
    out["gc_keeps_frame"] = bool(before == after)
    out["undo_redo_after_gc_keeps_frame"] = bool(redone == after and undone != after)
    out["undo_depth_after_gc"] = list(dag.undo_depth())
    dag.close()
    log(f"c5: {out}")
    if not (out["gc_keeps_frame"] and out["undo_redo_after_gc_keeps_frame"] and out["undo_depth_after_gc"] == list(depth)):
        raise RuntimeError("depth-17 workload: GC or undo changed the traced frame")
    return out


def cpu_edit_baseline(od, oc, params, args):
    """The oracle's edit path (a C++ restatement of edit_threads) on the CPU-plain workload: the
    2^12 binary terrain (colour tree down to the leaf level: no colour leaves), 100 sphere edits
    (50 add / 50 carve, radius uniform 8-64) at seeded surface points, one at a time, timed like
    the reference's "edits" stat (hash_dag.h:262, 368). The edit tasks run on one thread and on
    all host cores (the reference's thread pool, NUM_THREADS, typedefs.h:54-56); both runs must
    reach the same canonical DAG. The sample is bounded by --edit-cpu-seconds."""
    import numpy as np

    from hashdag_amd import Geometry, TerrainParams, generate_terrain
    from hashdag_amd.workload import sphere_edits
    from oracle import OracleDAG

    p1 = TerrainParams(levels=12, seed=args.seed, colors=False, color_tree_levels=10)
    sc = generate_terrain(p1)
    g = Geometry.fit(12, sc.level_words, color_tree_levels=10)
    edits = sphere_edits(p1, 100, seed=args.seed + 9)
    threads = len(os.sched_getaffinity(0))
    out = {}
    canon = []
    for th in (1, threads):
        od1 = OracleDAG.from_basic(g, sc.dag, 1 << 19)
        voxels, spent, n = 0, 0.0, 0
        for kind, center, r in edits:
            if spent >= args.edit_cpu_seconds / 2:
                break
            t = time.perf_counter()
            voxels += od1.edit(kind, center, r, None, threads=th)[1]
            spent += time.perf_counter() - t
            n += 1
        out[th] = (voxels / spent, n, voxels, spent)
        if n == len(edits):
            canon.append(od1.canonical())
        del od1
    same = len(canon) == 2 and bool(np.array_equal(canon[0], canon[1]))
    v1, vn = out[1], out[threads]
    return {"value": round(vn[0], 1), "unit": "edited voxels/s", "cores": threads, "kind": "port",
            "single_thread_value": round(v1[0], 1),
            "sample": f"C1: 2^12 binary terrain, {vn[1]} of 100 sphere edits (50 add / 50 carve, radius 8-64), "
                      f"{vn[2]} voxels in {vn[3]:.2f}s on {threads} threads; 1 thread: {v1[1]} edits, {v1[2]} voxels "
                      f"in {v1[3]:.2f}s" + ("; both reach the same canonical DAG" if same else "")}


def measured_edit_traffic(levels: int, n_edits: int, seed: int, geom):
    """HBM bytes of the fused edit pass from the committed rocprofv3 summary, when it covers
    this workload (scene depth, edit count and seed of the edit list, low-level buckets)."""
    path = os.path.join(ROOT, "profiles", "edit_kernels_2p16_4096edits.md")
    if not os.path.exists(path):
        return None
    want = {"levels": str(levels), "edits": str(n_edits), "seed": str(seed),
            "buckets_low": f"{1 << geom.bucket_bits_low}x{geom.bucket_size_low}"}
    for line in open(path):
        if line.startswith("traffic_bytes "):
            kv = dict(tok.split("=") for tok in line.split()[1:])
            if all(kv.get(k) == v for k, v in want.items()):
                return int(kv["fused_pass"])
    return None


def measured_traffic(kernel: str, levels: int, W: int, H: int, launch: dict):
    """HBM bytes per call of `kernel` (every kernel the call launches) from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary (scaled by the copy-kernel calibration), when it
    covers this workload, tile shape, fetch strategy and schedule."""
    path = os.path.join(ROOT, "profiles", "trace_kernels_2p15_1080p.md")
    if not os.path.exists(path):
        return None
    want = {"levels": str(levels), "width": str(W), "height": str(H), "tile": launch["tile"],
            "fetch": launch["fetch"], "schedule": launch["schedule_key"]}
    for line in open(path):
        if line.startswith("traffic_bytes "):
            kv = dict(tok.split("=") for tok in line.split()[1:])
            if all(kv.get(k) == v for k, v in want.items()):
                return int(kv[kernel]) if kernel in kv else None
    return None


def cpu_baseline(scene, geom, rparams, W, H, budget_s):
    """The oracle's host traversal (paths + colours + shadows) on a bounded sample of the same
    frames: whole rows of the 1080p frames, all host cores, until the time budget is used."""
    import numpy as np

    from oracle import Oracle, OracleColors, OracleDAG

    threads = len(os.sched_getaffinity(0))
    t0 = time.time()
    pool_pages = 0x8FFFFFFF // 512 // 4
    od = OracleDAG.from_basic(geom, scene.dag, pool_pages)
    oc = OracleColors.from_scene(scene)
    log(f"oracle loaded the scene in {time.time() - t0:.1f}s")
    rays = 0
    spent = 0.0
    frames = 0
    rows_per_chunk = 64
    while spent < budget_s and frames < len(rparams):
        p = np.array(rparams[frames].as_list())
        for y0 in range(0, H, rows_per_chunk):
            if spent >= budget_s:
                break
            y1 = min(H, y0 + rows_per_chunk)
            t = time.perf_counter()
            paths = Oracle.trace_paths(od, p, W, H, rows=(H - y1, H - y0), threads=threads)
            cols = Oracle.trace_colors(od, oc, paths, W, H, rows=(y0, y1), threads=threads)
            Oracle.trace_shadows(od, p, paths, cols, W, H, rows=(y0, y1), threads=threads)
            spent += time.perf_counter() - t
            rays += W * (y1 - y0)
        frames += 1
    return od, oc, {
        "value": round(rays / spent / 1e6, 3),
        "unit": "Mrays/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{rays} pixels ({rays / (W * H):.2f} frames of {W}x{H}, paths+colours+shadows) in {spent:.1f}s",
    }


if __name__ == "__main__":
    sys.exit(main())
