"""Benchmark: 640x480 scene-flow pairs/s on one MI355X, batched vs plain.

Workload: 640x480 synthetic RGB-D frame pairs (f = 525, c = (319.5, 239.5),
6 moving planar segments, seeds 21, 22, ...), 5-level pyramid at ratio 0.5,
15 outer x 50 inner iterations, boundary-probability regularisation with each
level's maps made from the frame on the device inside the step, as
sceneFlowMultiscale.cpp:116 does (--label-maps: maps from the generator's
segment labels instead), the reference's mixed precision (f64 images,
derivatives and flow; f32 system and increments).  A step is one batch of
--batch pairs solved together (depth stdevs, the six pyramids, boundary maps,
per-level features / upsample / warp / weights and the level solves of every
pair), with the inputs resident in HBM.  Beside it, as sub-objects of the same
line: the plain single-pair GPU path (one pair, one launch per colour per
relaxation iteration, as the reference schedules it), the tiled single pair,
the batch with the other boundary maps, the FP16-storage batch, a batch of 64
mixed-size pairs (320x240 to 1280x720) and 16 served streams of those sizes
at 30 fps with warm starts of 4 and of 15 outer iterations.

Prints one JSON line.  --gpus above 1 is refused: the solve runs on one GPU.
"""
from __future__ import annotations

import argparse
import importlib.util
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG_DIR = os.path.join(ROOT, "rgbd-flow_amd")

METRIC = "640×480 scene-flow pairs/s + p99 latency, batched vs plain; HBM GB/s"
W, H = 640, 480
B_SWEEP = 72.0     # algorithmic bytes per pixel per red-black iteration (f32 system + increments)
B_SWEEP_16 = 48.0  # the same with the 12 system values stored in 16 bits
B_STAGE = 128.0    # algorithmic bytes per pixel per outer iteration (warp, gradients, weights, assembly)
HBM_PEAK_GBS = 8000.0
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "relax_traffic.json")
T_START = time.perf_counter()


def log(msg):
    """A progress line on stderr, with the seconds since the start."""
    print(f"bench [{time.perf_counter() - T_START:6.1f} s]: {msg}", file=sys.stderr, flush=True)


def _package():
    spec = importlib.util.spec_from_file_location("rgbdflow_amd", os.path.join(PKG_DIR, "__init__.py"),
                                                  submodule_search_locations=[PKG_DIR])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["rgbdflow_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def workload(R, n_distinct):
    import rgbdflow_amd.synth as S
    cam = S.Camera(W, H, 525.0, 525.0, 319.5, 239.5)
    levels = R.pyramid_levels(W, H, 0.5, 40)
    pairs = []
    for i in range(n_distinct):
        pair = S.make_pair(W, H, 21 + i, n_segments=6, cam=cam)
        im1, d1, im2, d2 = S.to_solver_inputs(pair)
        hpb = S.segment_boundary_maps(pair["seg1"], [w for w, h in levels], [h for w, h in levels])
        pairs.append(dict(im1=im1, dm1=d1.astype(np.float64), im2=im2, dm2=d2.astype(np.float64), hpb=hpb,
                          gt=pair["flow"].astype(np.float64)))
    return pairs, levels


def flow_error(flow, gt):
    """Mean end-point error of (u, v) in px and mean |dw| in m against gt (H, W, 3)."""
    u, v, w = flow
    return (float(np.mean(np.hypot(u - gt[..., 0], v - gt[..., 1]))), float(np.mean(np.abs(w - gt[..., 2]))))


MIXED_SIZES = [(320, 240), (640, 480), (960, 540), (1280, 720)]


def mixed_workload(R, n, n_distinct, seed=100):
    """n pairs with sizes drawn uniformly from MIXED_SIZES, intrinsics scaled
    by W/640, 2-8 moving segments; n_distinct synthetic pairs per size are
    generated and cycled."""
    import rgbdflow_amd.synth as S
    rng = np.random.default_rng(seed)
    sizes = [MIXED_SIZES[k] for k in rng.integers(0, len(MIXED_SIZES), n)]
    made = {}
    pairs = []
    for i, (w, h) in enumerate(sizes):
        lst = made.setdefault((w, h), [])
        if len(lst) < n_distinct:
            f = 525.0 * w / 640
            cam = S.Camera(w, h, f, f, (w - 1) / 2, (h - 1) / 2)
            segs = int(rng.integers(2, 9))
            pr = S.make_pair(w, h, seed + i, n_segments=segs, cam=cam)
            im1, d1, im2, d2 = S.to_solver_inputs(pr)
            lv = R.pyramid_levels(w, h, 0.5, w // 16)
            hpb = S.segment_boundary_maps(pr["seg1"], [a for a, b in lv], [b for a, b in lv])
            c = R.make_params(w, h, f=f).cam
            lst.append(dict(im1=im1, dm1=d1.astype(np.float64), im2=im2, dm2=d2.astype(np.float64), hpb=hpb,
                            cam=c, min_width=w // 16))
        pairs.append(lst[sum(1 for x in sizes[:i] if x == (w, h)) % len(lst)])
    return pairs, sizes


def params_for(R, variant, graph=True, storage="f32"):
    st = {"f32": R.STORE_F32, "f16": R.STORE_F16, "bf16": R.STORE_BF16}[storage]
    return R.make_params(W, H, f=525.0, cx=319.5, cy=239.5, n_outer=15, n_sor=50, min_width=40,
                         relax_variant=R.RELAX_TILED if variant == "tiled" else R.RELAX_PLAIN, graph=graph,
                         storage=st)


def algorithmic_bytes(stats, b_sweep=B_SWEEP):
    sweep = stage = 0.0
    for st in stats:
        for k in range(st.n_levels):
            n = st.width[k] * st.height[k]
            sweep += n * st.sweeps_exec[k] * b_sweep
            stage += n * st.outer_exec[k] * B_STAGE
    return sweep, stage


def timed_runs(ctx, p, steps):
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        ctx.batch_run(p, stats=False)
    ctx.synchronize()
    return time.perf_counter() - t0


def cpu_threads():
    """Host threads for the all-cores CPU baseline: the cores this process may
    run on, capped by OMP_NUM_THREADS where the host sets a share."""
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    cap = os.environ.get("OMP_NUM_THREADS", "")
    if cap.isdigit() and int(cap) > 0:
        n = min(n, int(cap))
    return max(1, n)


def cpu_baseline(pr, budget_s, threads=1):
    """The oracle (reference-precision restatement; level-solve loops on
    `threads` row partitions, pyramids as concurrent tasks) on the same
    workload: as many whole pairs as fit in the time budget (at least one)."""
    if ROOT not in sys.path:
        sys.path.insert(0, ROOT)
    import oracle as O
    po = O.make_params(W, H, f=525.0, cx=319.5, cy=239.5, n_outer=15, n_sor=50, min_width=40)
    sigma = O.depth_stdevs(pr["dm1"].astype(np.float32), 4)
    O.set_threads(threads)
    try:
        n, t0 = 0, time.perf_counter()
        while True:
            O.multiscale(pr["im1"], pr["dm1"], pr["im2"], pr["dm2"], sigma, pr["hpb"], po)
            n += 1
            el = time.perf_counter() - t0
            if el > budget_s or n >= 24:
                break
    finally:
        O.set_threads(1)
    return n / el, n, el


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10, help="timed batches")
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--batch", type=int, default=16, help="pairs per step")
    ap.add_argument("--distinct", type=int, default=4, help="distinct synthetic pairs (cycled over the batch)")
    ap.add_argument("--variant", choices=["plain", "tiled"], default="tiled")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--save-schedules", help="write the autotuned schedules (JSON) here")
    ap.add_argument("--load-schedules", help="install schedules from this JSON instead of autotuning")
    ap.add_argument("--plain-steps", type=int, default=5, help="steps of the plain single-pair path")
    ap.add_argument("--single-steps", type=int, default=10, help="steps of the tiled single-pair path")
    ap.add_argument("--latency-steps", type=int, default=20)
    ap.add_argument("--cpu-budget", type=float, default=20.0,
                    help="seconds of CPU baseline work (half single-threaded, half on all cores)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-graph", action="store_true", help="launch every kernel instead of replaying a HIP graph")
    ap.add_argument("--storage", choices=["f32", "f16", "bf16"], default="f32",
                    help="storage of the per-pixel systems (f32: the reference's precision)")
    ap.add_argument("--mixed", type=int, default=0,
                    help="instead: a batch of this many pairs of mixed sizes (320x240 .. 1280x720)")
    ap.add_argument("--serve", default="",
                    help="instead: serving, streams as 'WxH:count,...' (e.g. 320x240:4,640x480:4,960x540:4,1280x720:4)")
    ap.add_argument("--fps", type=float, default=30.0, help="serving: frames per second per stream (Poisson)")
    ap.add_argument("--seconds", type=float, default=3.0, help="serving: arrival window")
    ap.add_argument("--max-batch", type=int, default=16, help="serving: pairs per batch")
    ap.add_argument("--frames", type=int, default=8, help="serving: frames of each size's synthetic sequence")
    ap.add_argument("--warm", action="store_true",
                    help="serving: solve each stream's later pairs as sceneFlowWithInit from its last flow")
    ap.add_argument("--warm-outer", type=int, default=0, help="serving: outer iterations of warm starts (0: 15)")
    ap.add_argument("--no-f16", action="store_true", help="skip the FP16-storage run of the batch")
    ap.add_argument("--images", choices=["f64", "f16", "bf16"], default="f64",
                    help="storage of the feature images of the batch (or the served streams); f64: the "
                         "reference's precision, 16 bits: the stage in f32 arithmetic")
    ap.add_argument("--fill-holes", action="store_true",
                    help="serving: hole-fill each frame's depth on the device (the frames here have no holes)")
    ap.add_argument("--device-hpb", action="store_true",
                    help="serving: submit no boundary maps; the server makes them on the device")
    ap.add_argument("--lanes2", action="store_true", help="also time the batch as two halves on two streams")
    ap.add_argument("--relax-level", default="",
                    help="instead: time the tiled relaxation alone on --batch levels of this size ('WxH'), "
                         "one check interval (25 iterations) per timing, --steps timings")
    ap.add_argument("--hpb", default="",
                    help="instead: time computeHeuristicPb on the device for frames of these sizes ('WxH,...')")
    ap.add_argument("--served-streams", type=int, default=16,
                    help="streams of the served measurement beside the batch, a quarter of each of 320x240, "
                         "640x480, 960x540 and 1280x720 (0: skip)")
    ap.add_argument("--served-seconds", type=float, default=4.0)
    ap.add_argument("--served-max-batch", type=int, default=8)
    ap.add_argument("--served-warm-outer", type=int, default=4,
                    help="outer iterations of the served measurement's warm starts (also run at 15)")
    ap.add_argument("--served-compare", type=int, default=4,
                    help="served warm pairs per size solved again at the configured iterations")
    ap.add_argument("--mixed-pairs", type=int, default=64, help="pairs of the mixed-size batch beside it (0: skip)")
    ap.add_argument("--mixed-steps", type=int, default=3)
    ap.add_argument("--label-maps", action="store_true",
                    help="boundary maps from the generator's segment labels instead of made on the device")
    ap.add_argument("--no-other-maps", action="store_true", help="skip the batch with the other boundary maps")
    ap.add_argument("--latency-batch", type=int, default=6,
                    help="pairs of the latency-bounded batch beside the headline (0: skip)")
    ap.add_argument("--coarse-levels", type=int, default=8,
                    help="levels of the coarse-grid correction run beside the headline (device maps; 0: skip)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: --gpus {args.gpus} refused: the scene-flow solve runs on one GPU "
              "(frame pairs are independent: one server or batch per GPU)", file=sys.stderr)
        sys.exit(2)

    R = _package()
    if args.mixed:
        return mixed_main(R, args)
    if args.serve:
        return serve_main(R, args)
    if args.hpb:
        return hpb_main(R, args)
    if args.relax_level:
        return relax_main(R, args)
    pairs, levels = workload(R, max(1, min(args.distinct, args.batch)))
    B = args.batch
    label_batch = [pairs[i % len(pairs)] for i in range(B)]
    # sceneFlowMultiscale.cpp:116 makes every level's boundary maps from the frame: the step does
    # the same on the device; --label-maps takes the generator's segment-label maps instead
    dev_batch = [dict(pr, hpb=None) for pr in label_batch]
    batch = label_batch if args.label_maps else dev_batch
    maps = "generator labels" if args.label_maps else "device (computeHeuristicPb per level, in the step)"
    p = params_for(R, args.variant, not args.no_graph, args.storage)
    p.image_storage = {"f64": R.IMAGES_F64, "f16": R.IMAGES_F16, "bf16": R.IMAGES_BF16}[args.images]
    ctx = R.Context(0)
    raw = {}

    def autotune(nb, params, record=True):
        out = []
        for (w, h) in levels:
            ch = ctx.autotune_relax(w, h, batch=nb, params=params)
            if record:
                raw.setdefault(nb, []).append(dict(W=w, H=h, **ch))
            out.append(dict(level=f"{w}x{h}", tile=f"{ch['tile_w']}x{ch['tile_h']}", iters=ch["tile_iters"],
                            threads=ch["threads"], pairs_per_thread=ch["pairs_per_thread"],
                            us_per_iter=round(ch["us_per_iter"], 2)))
        return out

    def install(nb):
        """The schedules of batch size nb again (a later search installs its own)."""
        sched = raw.get(nb)
        if args.load_schedules:
            with open(args.load_schedules) as f:
                sched = json.load(f).get(str(nb))
        for ch in sched or []:
            ctx.set_relax_schedule(ch["W"], ch["H"], ch)

    tuned = []
    if args.variant == "tiled" and args.load_schedules:
        install(B)
        tuned = [dict(level=f"{ch['W']}x{ch['H']}", loaded=True) for ch in json.load(open(args.load_schedules))[str(B)]]
    elif args.variant == "tiled" and not args.no_autotune:
        # one-time schedule search per level shape (outside the timed region)
        tuned = autotune(B, p)
        log(f"autotuned for batch {B}: {tuned}")
    log(f"{B} x 640x480 pairs, {len(levels)} levels, 15x50, variant {args.variant}, maps {maps}")
    pairs_per_s, elapsed, lat, st = run_batch(ctx, p, batch, args.steps, args.warmup, args.latency_steps)
    log(f"timed {args.steps} batches in {elapsed:.3f} s")
    flow0 = ctx.batch_read(0)

    # dominant kernel: the relaxation, timed launch by launch with HIP events
    ctx.set_relax_timing(True)
    rst = ctx.batch_run(p, stats=True)
    ctx.set_relax_timing(False)
    sweep_bytes, stage_bytes = algorithmic_bytes(rst, B_SWEEP if args.storage == "f32" else B_SWEEP_16)
    relax_ms = rst[0].relax_ms
    launches = rst[0].relax_launches
    achieved = sweep_bytes / (relax_ms / 1e3) / 1e9 if relax_ms > 0 else 0.0
    avg_launch_us = relax_ms * 1e3 / max(launches, 1)

    # the plain single-pair GPU path (the reference's schedule) and the tiled single pair, same maps
    plain_pps = single_pps = None
    single_lat = None
    if args.plain_steps > 0:
        pp = params_for(R, "plain", graph=False)
        ctx.batch_load(pp, batch[:1])
        ctx.batch_run(pp, stats=True)
        plain_pps = args.plain_steps / timed_runs(ctx, pp, args.plain_steps)
        log(f"plain single pair: {plain_pps:.2f} pairs/s")
    if args.single_steps > 0 and args.variant == "tiled":
        if args.load_schedules:
            install(1)
        elif not args.no_autotune:
            autotune(1, p)
        ctx.batch_load(p, batch[:1])
        ctx.batch_run(p, stats=True)
        single_pps = args.single_steps / timed_runs(ctx, p, args.single_steps)
        single_lat = np.array([ctx.batch_run(p, stats=True)[0].device_ms for _ in range(10)])
        log(f"tiled single pair: {single_pps:.2f} pairs/s")
    install(B)

    def speedup(v):
        return round(v / plain_pps, 2) if plain_pps else None

    # the same pairs as two half batches on two contexts (two HIP streams), so
    # one half's coarse, latency-bound levels overlap the other's fine levels
    two_lanes = None
    if args.lanes2 and args.variant == "tiled" and B >= 2:
        half = B // 2
        ctxs = [R.Context(0), R.Context(0)]
        for (w, h) in levels:
            ch = ctxs[0].autotune_relax(w, h, batch=half, params=p)
            ctxs[1].set_relax_schedule(w, h, ch)
        ctxs[0].batch_load(p, batch[:half])
        ctxs[1].batch_load(p, batch[half:])
        for c in ctxs:
            c.batch_run(p, stats=True)
        for _ in range(max(1, args.warmup)):
            for c in ctxs:
                c.batch_run(p, stats=False)
        for c in ctxs:
            c.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            for c in ctxs:
                c.batch_run(p, stats=False)
        for c in ctxs:
            c.synchronize()
        el2 = time.perf_counter() - t0
        two_lanes = {"pairs_per_s": round(B * args.steps / el2, 3), "ms_per_step": round(el2 / args.steps * 1e3, 3),
                     "lanes": 2, "pairs_per_lane": half}
        for c in ctxs:
            c.close()
        log(f"two lanes {two_lanes}")

    # the same f32 batch with the other boundary maps
    other_maps = None
    if args.storage == "f32" and args.variant == "tiled" and not args.no_other_maps:
        ob = dev_batch if args.label_maps else label_batch
        opps, oel, olat, ost = run_batch(ctx, p, ob, args.steps, args.warmup, 0)
        L = ost[0].n_levels
        other_maps = {"boundary_maps": "device" if args.label_maps else "generator labels",
                      "pairs_per_s": round(opps, 3), "ms_per_step": round(oel / args.steps * 1e3, 3),
                      "speedup_vs_plain_single": speedup(opps),
                      "sweeps_exec_pair0": [ost[0].sweeps_exec[k] for k in range(L)],
                      "outer_exec_pair0": [ost[0].outer_exec[k] for k in range(L)]}
        log(f"other boundary maps {other_maps}")

    # the device-map batch with the coarse-grid correction (opt-in): time, sweeps, and pair 0's
    # distance to the generator's truth beside the reference rule's
    coarse = None
    if args.storage == "f32" and args.variant == "tiled" and args.coarse_levels > 0:
        pc = params_for(R, args.variant, not args.no_graph, args.storage)
        pc.image_storage = p.image_storage
        pc.coarse_levels = args.coarse_levels
        cpps, cel, _, cst = run_batch(ctx, pc, dev_batch, args.steps, args.warmup, 0)
        cflow = ctx.batch_read(0)
        if args.label_maps:
            ctx.batch_load(p, dev_batch)
            rsw = ctx.batch_run(p, stats=True)[0]
            rflow = ctx.batch_read(0)
        else:
            rflow, rsw = flow0, rst[0]
        gt = label_batch[0]["gt"]
        e_ref, e_cg = flow_error(rflow, gt), flow_error(cflow, gt)
        L = cst[0].n_levels
        coarse = {"coarse_levels": args.coarse_levels, "coarse_sweeps": pc.coarse_sweeps, "coarse_at": pc.coarse_at,
                  "boundary_maps": "device", "pairs_per_s": round(cpps, 3), "ms_per_step": round(cel / args.steps * 1e3, 3),
                  "sweeps_exec_pair0": [cst[0].sweeps_exec[k] for k in range(L)],
                  "sweeps_exec_pair0_reference_rule": [rsw.sweeps_exec[k] for k in range(L)],
                  "pair0_epe_px_vs_truth": round(e_cg[0], 4), "pair0_abs_dw_m_vs_truth": round(e_cg[1], 6),
                  "pair0_epe_px_vs_truth_reference_rule": round(e_ref[0], 4),
                  "pair0_abs_dw_m_vs_truth_reference_rule": round(e_ref[1], 6)}
        log(f"coarse-grid correction {coarse}")

    # the same batch with the systems stored in FP16 (f32 arithmetic)
    f16 = None
    if args.storage == "f32" and not args.no_f16 and args.variant == "tiled":
        p16 = params_for(R, args.variant, not args.no_graph, "f16")
        if not args.no_autotune and not args.load_schedules:
            autotune(B, p16, record=False)
        f16pps, el16, lat16, _ = run_batch(ctx, p16, batch, args.steps, args.warmup, max(5, args.latency_steps // 2))
        ctx.set_relax_timing(True)
        r16 = ctx.batch_run(p16, stats=True)
        ctx.set_relax_timing(False)
        sb16, _ = algorithmic_bytes(r16, B_SWEEP_16)
        a16 = sb16 / (r16[0].relax_ms / 1e3) / 1e9 if r16[0].relax_ms > 0 else 0.0
        f16 = {"pairs_per_s": round(f16pps, 3), "speedup_vs_plain_single": speedup(f16pps),
               "batch_p99_ms": round(float(np.percentile(lat16, 99)), 3),
               "relax_roofline": {"achieved": round(a16, 1), "frac": round(a16 / HBM_PEAK_GBS, 4),
                                  "bytes_per_px_iter": B_SWEEP_16},
               "relax_avg_launch_us": round(r16[0].relax_ms * 1e3 / max(r16[0].relax_launches, 1), 3),
               "stage_ms_per_batch": round(r16[0].device_ms - r16[0].relax_ms, 3)}
        log(f"fp16 storage {f16}")
        # and the feature images (frame 1, frame 2, warped frame 2) in FP16 as well
        p16.image_storage = R.IMAGES_F16
        f16ipps, _, lat16i, _ = run_batch(ctx, p16, batch, args.steps, args.warmup, max(5, args.latency_steps // 2))
        ctx.set_relax_timing(True)
        r16i = ctx.batch_run(p16, stats=True)
        ctx.set_relax_timing(False)
        sb16i, _ = algorithmic_bytes(r16i, B_SWEEP_16)
        a16i = sb16i / (r16i[0].relax_ms / 1e3) / 1e9 if r16i[0].relax_ms > 0 else 0.0
        f16["with_fp16_images"] = {"pairs_per_s": round(f16ipps, 3), "speedup_vs_plain_single": speedup(f16ipps),
                                   "batch_p99_ms": round(float(np.percentile(lat16i, 99)), 3),
                                   "stage": "f32 arithmetic (16-bit feature images)",
                                   "stage_ms_per_batch": round(r16i[0].device_ms - r16i[0].relax_ms, 3),
                                   "relax_roofline": {"achieved": round(a16i, 1), "frac": round(a16i / HBM_PEAK_GBS, 4),
                                                      "bytes_per_px_iter": B_SWEEP_16}}
        log(f"fp16 storage and images {f16['with_fp16_images']}")

    # a smaller batch of the same pairs, whose every pair finishes within a frame interval at
    # 30 fps (batch p99 <= 33 ms): f32, and FP16 records and images
    lat_bounded = None
    if args.latency_batch > 0 and args.variant == "tiled":
        Bl = args.latency_batch
        lat_bounded = {"batch": Bl, "boundary_maps": "device" if not args.label_maps else "generator labels"}
        for name, st, img in (("f32", "f32", R.IMAGES_F64), ("fp16_images", "f16", R.IMAGES_F16)):
            pl = params_for(R, args.variant, not args.no_graph, st)
            pl.image_storage = img
            if not args.no_autotune and not args.load_schedules:
                autotune(Bl, pl, record=False)
            lpps, lel, llat, _ = run_batch(ctx, pl, batch[:Bl], args.steps, args.warmup, max(5, args.latency_steps))
            lat_bounded[name] = {"pairs_per_s": round(lpps, 3), "speedup_vs_plain_single": speedup(lpps),
                                 "batch_p50_ms": round(float(np.percentile(llat, 50)), 3),
                                 "batch_p99_ms": round(float(np.percentile(llat, 99)), 3)}
        log(f"latency-bounded batch {lat_bounded}")
    ctx.close()

    # 64 mixed-size pairs in one batch, f32 and FP16 records and images
    mixed = None
    if args.mixed_pairs > 0 and args.variant == "tiled":
        mixed = mixed_batch(R, args, args.mixed_pairs, label_maps=args.label_maps)
        log(f"mixed batch {mixed}")

    # 16 served streams of four sizes at 30 fps, FP16 records and images, boundary maps made on the
    # device, warm starts of --served-warm-outer and of the configured 15 outer iterations
    served = None
    if args.served_streams > 0:
        per = max(1, args.served_streams // len(MIXED_SIZES))
        spec = [sz for sz in MIXED_SIZES for _ in range(per)]
        frames = serve_frames(R, spec, args.frames, maps=False)
        served, sched = {}, []
        for wo in (args.served_warm_outer, 15):
            r = serve_run(R, spec, frames, args.fps, args.served_seconds, args.served_max_batch, "f16", True, wo,
                          autotune=not args.no_autotune, images="f16", device_hpb=True, schedules=sched or None,
                          tuned_out=sched, compare=args.served_compare if wo != 15 else 0)
            r.update(warm_n_outer=wo)
            served[f"warm_{wo}_outer"] = r
            log(f"served, warm {wo} outer: {r}")
        served.update(streams=",".join(f"{w}x{h}:{per}" for (w, h) in MIXED_SIZES), fps_per_stream=args.fps,
                      max_batch=args.served_max_batch, lanes=2, storage="f16", images="f16",
                      boundary_maps="device")

    if args.save_schedules:
        with open(args.save_schedules, "w") as f:
            json.dump({str(k): v for k, v in raw.items()}, f)

    traffic = None
    if os.path.exists(TRAFFIC_FILE):
        with open(TRAFFIC_FILE) as f:
            tr = json.load(f)
        if tr.get("batch") == B and tr.get("variant") == args.variant and tr.get("storage", "f32") == args.storage:
            traffic = tr.get("bytes_per_launch")

    out = {
        "metric": METRIC,
        "value": round(pairs_per_s, 3),
        "unit": "pairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": ("f32" if args.storage == "f32" else f"f32 compute, {args.storage} system storage") +
                 ("" if args.images == "f64" else f", {args.images} feature images, f32 stage"),
        "data": "synthetic (seeded generator, ground-truth flow)",
        "config": {
            "workload": f"{B} x 640x480 RGB-D pairs per step ({len(pairs)} distinct), 5-level pyramid ratio 0.5, "
                        f"15 outer x 50 inner, HPB regularisation, boundary maps: {maps}",
            "batch": B,
            "precision": "reference mixed: f64 images/derivatives/flow, f32 system/increments",
            "relax": args.variant,
            "batch_p50_ms": round(float(np.percentile(lat, 50)), 3),
            "batch_p99_ms": round(float(np.percentile(lat, 99)), 3),
            "sweeps_exec_pair0": [rst[0].sweeps_exec[k] for k in range(rst[0].n_levels)],
            "outer_exec_pair0": [rst[0].outer_exec[k] for k in range(rst[0].n_levels)],
            "outer_stat_margin_pair0": [round(rst[0].stat_margin[k], 4) for k in range(rst[0].n_levels)],
            "outer_stat_min_margin_batch": round(min(x.stat_margin[k] for x in rst for k in range(x.n_levels)
                                                     if x.stat_margin[k] >= 0), 4),
            "algorithmic_GBps": round((sweep_bytes + stage_bytes) / (np.percentile(lat, 50) / 1e3) / 1e9, 1),
            "relax_launches_per_batch": launches,
            "relax_avg_launch_us": round(avg_launch_us, 3),
            "stage_ms_per_batch": round(rst[0].device_ms - relax_ms, 3),
            "plain_single_pairs_per_s": round(plain_pps, 3) if plain_pps else None,
            "speedup_vs_plain_single": speedup(pairs_per_s),
            "tiled_single_pairs_per_s": round(single_pps, 3) if single_pps else None,
            "tiled_single_p99_ms": round(float(np.percentile(single_lat, 99)), 3) if single_lat is not None else None,
            "autotuned": tuned,
            "two_lanes": two_lanes,
            "fp16_storage": f16,
            "other_boundary_maps": other_maps,
            "coarse_correction": coarse,
            "latency_bounded": lat_bounded,
            "mixed_64": mixed,
            "served_16_mixed": served,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
        },
    }
    if not args.no_cpu_baseline:
        nt = cpu_threads()
        cpu_pair = dict(pairs[0], hpb=None if not args.label_maps else pairs[0]["hpb"])
        v1, n1, el1 = cpu_baseline(cpu_pair, args.cpu_budget / 2, 1)
        vn, nn, eln = cpu_baseline(cpu_pair, args.cpu_budget / 2, nt)
        log(f"cpu baseline: {vn:.3f} pairs/s on {nt} threads, {v1:.3f} on 1")
        out["cpu_baseline"] = {"value": round(vn, 4), "unit": "pairs/s", "cores": nt, "kind": "port",
                               "sample": f"{nn} whole pairs of the same 640x480 workload through the oracle "
                                         f"(reference precision, boundary maps {'from labels' if args.label_maps else 'made per level'}, "
                                         f"level-solve "
                                         f"rows on {nt} threads, {eln:.1f} s); 1 thread: {n1} pairs in {el1:.1f} s",
                               "value_1_thread": round(v1, 4)}
    print(json.dumps(out), flush=True)


def run_batch(ctx, p, batch, steps, warmup, latency_steps):
    """Loads the batch, runs it once with statistics and warmup times, then
    times `steps` batches; then `latency_steps` batches one by one (device
    time from HIP events).  Returns (pairs/s, seconds, latencies in ms,
    statistics of the first run)."""
    ctx.batch_load(p, batch)
    st = ctx.batch_run(p, stats=True)
    for _ in range(warmup):
        ctx.batch_run(p, stats=False)
    el = timed_runs(ctx, p, steps)
    lat = np.array([ctx.batch_run(p, stats=True)[0].device_ms for _ in range(latency_steps)])
    return len(batch) * steps / el, el, lat, st


def mixed_batch(R, args, n, label_maps=False, schedules_in=None, schedules_out=None):
    """n pairs of mixed sizes in one batch, schedules autotuned per level
    shape (f32); at each level, every size of two or more pairs relaxes in a
    step of its own and the other sizes share one step.  f32, then FP16
    records, then FP16 records and images, under the f32 schedules (FP16
    records' own, tuned at each size's finest level, gained 0.7% for 90 s of
    tuning)."""
    pairs, sizes = mixed_workload(R, n, max(1, args.distinct // 2))
    if not label_maps:
        pairs = [dict(pr, hpb=None) for pr in pairs]
    out = {"workload": f"{len(pairs)} pairs, sizes uniform over 320x240/640x480/960x540/1280x720 (intrinsics "
                       "scaled by W/640, 2-8 segments), 5-level pyramids, 15 outer x 50 inner, boundary maps "
                       + ("from generator labels" if label_maps else "made on the device"),
           "sizes": {f"{w}x{h}": sizes.count((w, h)) for (w, h) in MIXED_SIZES}}
    steps = max(1, args.mixed_steps)
    with R.Context(0) as ctx:
        p = R.make_params(640, 480, n_outer=15, n_sor=50, min_width=40, relax_variant=R.RELAX_TILED,
                          graph=not args.no_graph)
        if schedules_in:
            # schedules tuned by an earlier run (e.g. an unprofiled one)
            with open(schedules_in) as f:
                for ch in json.load(f)["mixed"]:
                    ctx.set_relax_schedule(ch["W"], ch["H"], ch)
            log("mixed batch: schedules loaded")
        elif not args.no_autotune:
            tuned = []
            for (w, h) in sorted(set(sizes), key=lambda t: -t[0]):
                for (lw, lh) in R.pyramid_levels(w, h, 0.5, w // 16):
                    tuned.append(dict(ctx.autotune_relax(lw, lh, batch=sizes.count((w, h)), params=p), W=lw, H=lh))
                log(f"mixed batch: f32 schedules of the {w}x{h} levels tuned")
            if schedules_out:
                with open(schedules_out, "w") as f:
                    json.dump({"mixed": tuned}, f)
        npx = sum(w * h for (w, h) in sizes)
        for name, st, img in (("f32", R.STORE_F32, R.IMAGES_F64), ("fp16", R.STORE_F16, R.IMAGES_F64),
                              ("fp16_images", R.STORE_F16, R.IMAGES_F16)):
            p.storage, p.image_storage = st, img
            pps, el, lat, _ = run_batch(ctx, p, pairs, steps, 1, 2)
            log(f"mixed batch, {name}: {pps:.2f} pairs/s")
            out[name] = {"pairs_per_s": round(pps, 3), "ms_per_step": round(el / steps * 1e3, 3),
                         "megapixels_per_s": round(npx * steps / el / 1e6, 2),
                         "batch_p50_ms": round(float(np.percentile(lat, 50)), 3)}
    return out


def mixed_main(R, args):
    """The mixed-size batch alone (--mixed pairs, --steps timed batches)."""
    args.mixed_steps = args.steps
    r = mixed_batch(R, args, args.mixed, label_maps=args.label_maps, schedules_in=args.load_schedules,
                    schedules_out=args.save_schedules)
    out = {"metric": "mixed-size scene-flow pairs/s (one batch)", "value": r["f32"]["pairs_per_s"], "unit": "pairs/s",
           "n_gpus": 1, "steps": args.steps, "warmup": 1, "ms_per_step": r["f32"]["ms_per_step"],
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
           "data": "synthetic (seeded generator)", "config": dict(graph=not args.no_graph, **r)}
    print(json.dumps(out), flush=True)


def relax_main(R, args):
    """The tiled relaxation alone: --batch levels of one size, systems of the
    autotuner's synthetic kind stored as --storage, the autotuned (or loaded)
    schedule; device time per red-black iteration from HIP events around one
    check interval (25 iterations), median of --steps timings, and the
    algorithmic rate at 72 (f32) / 48 (16-bit) bytes per pixel and iteration."""
    w, h = (int(x) for x in args.relax_level.split("x"))
    B = args.batch
    st = {"f32": R.STORE_F32, "f16": R.STORE_F16, "bf16": R.STORE_BF16}[args.storage]
    p = R.make_params(w, h, relax_variant=R.RELAX_TILED, storage=st)
    ctx = R.Context(0)
    if args.load_schedules:
        with open(args.load_schedules) as f:
            ch = json.load(f)["relax"]
    else:
        ch = ctx.autotune_relax(w, h, batch=B, params=p)
        if args.save_schedules:
            with open(args.save_schedules, "w") as f:
                json.dump({"relax": ch}, f)
    log(f"relaxation of {B} x {w}x{h}, schedule {ch}")
    us = sorted(ctx.time_relax_schedule(w, h, ch, batch=B, params=p) for _ in range(max(1, args.steps)))
    med = us[len(us) // 2]
    bpp = 72.0 if args.storage == "f32" else 48.0
    gbps = B * w * h * bpp / (med * 1e-6) / 1e9
    ctx.close()
    out = {"metric": "tiled relaxation, device time per red-black iteration", "value": round(med, 3),
           "unit": "us per iteration", "n_gpus": 1, "steps": args.steps, "higher_is_better": False,
           "dtype": "f32" if args.storage == "f32" else f"f32 compute, {args.storage} system storage",
           "data": "synthetic systems (the autotuner's)",
           "config": {"workload": f"{B} x {w}x{h} levels", "schedule": ch,
                      "roofline": {"bound": "hbm", "achieved": round(gbps, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                   "frac": round(gbps / HBM_PEAK_GBS, 4), "bytes_per_px_iter": bpp}}}
    print(json.dumps(out), flush=True)


def hpb_main(R, args):
    """computeHeuristicPb of one frame on the device (seven kernels), per size:
    wall time of the C-ABI call (host frame in, maps out, PCIe included) over
    --steps calls after --warmup, and the oracle's time for the same frame on
    one host core."""
    import rgbdflow_amd.synth as S
    ctx = R.Context(0)
    rows = {}
    for tok in args.hpb.split(","):
        w, h = (int(x) for x in tok.split("x"))
        cam = S.Camera(w, h, 525.0 * w / 640, 525.0 * w / 640, (w - 1) / 2, (h - 1) / 2)
        pair = S.make_pair(w, h, 21, n_segments=6, cam=cam)
        im1, d1, _, _ = S.to_solver_inputs(pair)
        dm = d1.astype(np.float64)
        for _ in range(args.warmup):
            ctx.heuristic_pb(im1, dm)
        t0 = time.perf_counter()
        for _ in range(args.steps):
            ctx.heuristic_pb(im1, dm)
        el = (time.perf_counter() - t0) / args.steps
        row = {"ms_per_frame": round(el * 1e3, 3), "megapixels_per_s": round(w * h / el / 1e6, 1)}
        if not args.no_cpu_baseline:
            import oracle as O
            t0 = time.perf_counter()
            O.heuristic_pb(im1, dm)
            row["oracle_ms_1_thread"] = round((time.perf_counter() - t0) * 1e3, 1)
        rows[tok] = row
        log(f"hpb {tok}: {row}")
    ctx.close()
    out = {"metric": "computeHeuristicPb frames/s on the device", "unit": "ms per frame", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "higher_is_better": False, "dtype": "f32 (f64 normals)",
           "data": "synthetic (seeded generator)", "value": rows[args.hpb.split(",")[0]]["ms_per_frame"],
           "config": {"workload": "one frame per call, host frame in and maps out", "sizes": rows}}
    print(json.dumps(out), flush=True)


def serve_frames(R, spec, n_frames, maps=True):
    """Per stream size: a synthetic frame sequence (every segment on a smooth
    constant-velocity SE(3) trajectory), with each frame's boundary maps from
    the segment labels (None without maps)."""
    import rgbdflow_amd.synth as S
    frames = {}
    for (w, h) in sorted(set(spec), key=lambda t: -t[0]):
        f = 525.0 * w / 640
        seq = S.make_sequence(w, h, 500 + w, n_frames, n_segments=6, cam=S.Camera(w, h, f, f, (w - 1) / 2, (h - 1) / 2))
        lv = R.pyramid_levels(w, h, 0.5, w // 16)
        frames[(w, h)] = ([(fr["img"], (fr["depth"].astype(np.float64) * .001).astype(np.float32),
                            S.segment_boundary_maps(fr["seg"], [a for a, b in lv], [b for a, b in lv]) if maps
                            else None)
                           for fr in seq["frames"]], seq)
    return frames


def serve_run(R, spec, frames, fps, seconds, max_batch, storage, warm, warm_outer, autotune=True, lanes=2, seed=7,
              score_every=8, warmup=1.0, device_hpb=False, images="f64", fill_holes=False, graph=True,
              save_schedules=None, load_schedules=None, schedules=None, tuned_out=None, compare=0):
    """Streams submit frames at Poisson arrival times (rate fps each); stream k
    walks its size's sequence back and forth from its own starting frame.
    Arrivals run for warmup + seconds; requests that arrive in the first
    `warmup` seconds (while the server's buffers and plans are first made)
    are served but not counted.  Returns the summary of the counted window:
    throughput, request latencies, and the end-point error against the true
    flow of every score_every-th served pair.
    schedules: relaxation schedules to install (else autotuned; the tuned
    ones are appended to tuned_out).  compare > 0 (warm starts): up to that
    many scored warm pairs per size are solved again after the run, as
    sceneFlowWithInit from the estimate the server used, with the served
    outer iterations (the served flow must come out the same) and with the
    configured 15, and coarse to fine at 15 x 50; the served flow's distance
    to the last two is reported."""
    import threading
    import rgbdflow_amd.synth as S
    sizes = sorted(set(spec), key=lambda t: -t[0])
    st = {"f32": R.STORE_F32, "f16": R.STORE_F16, "bf16": R.STORE_BF16}[storage]
    p = R.make_params(640, 480, n_outer=15, n_sor=50, min_width=40, relax_variant=R.RELAX_TILED, graph=graph,
                      storage=st, image_storage={"f64": R.IMAGES_F64, "f16": R.IMAGES_F16, "bf16": R.IMAGES_BF16}[images])
    srv = R.Server(p, max_batch=max_batch, lanes=lanes, warm_start=warm, warm_n_outer=warm_outer,
                   fill_holes=fill_holes)
    if load_schedules:
        with open(load_schedules) as f:
            schedules = json.load(f)["served"]
    if schedules:
        for ch in schedules:
            srv.set_relax_schedule(ch["W"], ch["H"], ch)
    elif autotune:
        tuned = []
        with R.Context(0) as tctx:
            for (w, h) in sizes:
                for (lw, lh) in R.pyramid_levels(w, h, 0.5, w // 16):
                    ch = tctx.autotune_relax(lw, lh, batch=max(1, min(4, max_batch)), params=p)
                    srv.set_relax_schedule(lw, lh, ch)
                    tuned.append(dict(W=lw, H=lh, **ch))
        if tuned_out is not None:
            tuned_out.extend(tuned)
        if save_schedules:
            with open(save_schedules, "w") as f:
                json.dump({"served": tuned}, f)
    cams = [R.make_params(w, h).cam for (w, h) in spec]
    sids = [srv.open(w, h, cams[k], w // 16) for k, (w, h) in enumerate(spec)]
    rng = np.random.default_rng(seed)
    events = []
    for k, (w, h) in enumerate(spec):
        t = rng.exponential(1 / fps)
        n = 0
        while t < warmup + seconds:
            events.append((t, k, n))
            t += rng.exponential(1 / fps)
            n += 1
    events.sort()
    tickets = []

    def index_of(k, n):
        period = 2 * (len(frames[spec[k]][0]) - 1)
        i = (n + 2 * k) % period
        return i if i < period // 2 + 1 else period - i

    def frame_of(k, n):
        return frames[spec[k]][0][index_of(k, n)]

    def submitter():
        t0 = time.perf_counter()
        for (t, k, n) in events:
            dt = t - (time.perf_counter() - t0)
            if dt > 0:
                time.sleep(dt)
            img, dep, hpb = frame_of(k, n)
            tickets.append((k, n, t >= warmup, srv.submit(sids[k], img, dep, None if device_hpb else hpb)))

    t0 = time.perf_counter()
    th = threading.Thread(target=submitter)
    th.start()
    # results are collected while frames still arrive, in submission order
    lat, qms, bsz, scored, again = [], [], [], [], []
    last = {}                    # stream -> its latest result (the next pair's warm-start estimate)
    per_size = {}
    polled = 0
    t_last = t0 + warmup + 1e-9
    while polled < len(tickets) or th.is_alive():
        if polled == len(tickets):
            time.sleep(0.0005)
            continue
        k, n, counted, tk = tickets[polled]
        polled += 1
        info, flow = srv.poll(tk, wait_ms=-1)
        if info["has_flow"] and counted:
            t_last = time.perf_counter()
            lat.append(info["latency_ms"])
            qms.append(info["queue_ms"])
            bsz.append(info["batch_pairs"])
            if len(lat) % score_every == 0:
                scored.append((k, n, flow))
                if compare and warm and n >= 2 and k in last and per_size.get(spec[k], 0) < compare:
                    per_size[spec[k]] = per_size.get(spec[k], 0) + 1
                    again.append((k, n, flow, last[k]))
        if info["has_flow"]:
            last[k] = flow
    th.join()
    wall = t_last - (t0 + warmup)
    srv.close()
    epe = []
    for (k, n, flow) in scored:   # the pair is (frame n-1, frame n) of stream k
        gt = S.sequence_flow(frames[spec[k]][1], index_of(k, n - 1), index_of(k, n))
        epe.append(float(np.median(np.hypot(flow[0] - gt[..., 0], flow[1] - gt[..., 1]))))
    out = dict(pairs_per_s=round(len(lat) / wall, 3), offered_pairs_per_s=round(len(lat) / seconds, 2),
               warmup_s=warmup,
               p50_ms=round(float(np.percentile(lat, 50)), 2), p99_ms=round(float(np.percentile(lat, 99)), 2),
               queue_p50_ms=round(float(np.percentile(qms, 50)), 2), mean_batch_pairs=round(float(np.mean(bsz)), 2),
               requests=len(lat), wall_s=round(wall, 2), boundary_maps="device" if device_hpb else "generator labels",
               epe_vs_truth_median_px=round(float(np.median(epe)), 4) if epe else None, scored_pairs=len(epe))
    if again:
        out["warm_vs_configured"] = resolve_served(R, p, spec, frames, again, index_of, cams, warm_outer or 15,
                                                   device_hpb)
    return out


def flow_distance(flows, refs):
    """EPE (u, v; px) and |dw| (m) over every pixel of the pairs."""
    e = np.concatenate([np.hypot(a[0] - b[0], a[1] - b[1]).ravel() for a, b in zip(flows, refs)])
    dw = np.concatenate([np.abs(a[2] - b[2]).ravel() for a, b in zip(flows, refs)])
    return dict(epe_median_px=float(np.median(e)), epe_mean_px=float(e.mean()), epe_p99_px=float(np.percentile(e, 99)),
                dw_median_m=float(np.median(dw)), dw_p99_m=float(np.percentile(dw, 99)))


def resolve_served(R, p, spec, frames, again, index_of, cams, served_outer, device_hpb):
    """The served warm pairs solved again in one batch each way: as
    sceneFlowWithInit from the estimate the server used, with the served
    outer iterations (bit-equal to the served flow) and with 15
    (sceneFlowWithInit.cpp:74-80 takes nOuterFPIterations), and coarse to
    fine at 15 x 50 from scratch."""
    def inputs(k, n, est):
        w, h = spec[k]
        fr = frames[spec[k]][0]
        a, b = fr[index_of(k, n - 1)], fr[index_of(k, n)]
        d = dict(im1=a[0].astype(np.float64) / 255.0, dm1=a[1].astype(np.float64), im2=b[0].astype(np.float64) / 255.0,
                 dm2=b[1].astype(np.float64), hpb=None if device_hpb else a[2], cam=cams[k], min_width=w // 16)
        if est is not None:
            d["est"] = np.stack(est, axis=-1)
        return d

    def solve(n_outer, warm):
        q = R.make_params(640, 480, n_outer=n_outer, n_sor=50, min_width=40, relax_variant=R.RELAX_TILED,
                          graph=False, storage=p.storage, image_storage=p.image_storage)
        with R.Context(0) as c:
            c.batch_load(q, [inputs(k, n, est if warm else None) for (k, n, _, est) in again])
            c.batch_run(q, stats=False)
            return [c.batch_read(i) for i in range(len(again))]

    served = [f for (_, _, f, _) in again]
    same = solve(served_outer, True)
    w15 = solve(15, True)
    cold = solve(15, False)
    equal = all(all(np.array_equal(x, y) for x, y in zip(a, b)) for a, b in zip(served, same))
    return {"pairs": len(again), "served_equal_resolved": bool(equal),
            "vs_warm_15_outer": flow_distance(served, w15), "vs_cold_15x50": flow_distance(served, cold),
            "warm_15_vs_cold_15x50": flow_distance(w15, cold)}


def serve_main(R, args):
    """Serving: streams submit frames at Poisson arrival times; every frame after
    a stream's first is a pair solved by the server's continuous batching.
    Reports offered and achieved pairs/s and per-request latency."""
    spec = []
    for tok in args.serve.split(","):
        size, cnt = tok.split(":")
        w, h = (int(x) for x in size.split("x"))
        spec += [(w, h)] * int(cnt)
    frames = serve_frames(R, spec, args.frames, maps=not args.device_hpb)
    log(f"serving {len(spec)} streams over {args.seconds} s")
    r = serve_run(R, spec, frames, args.fps, args.seconds, args.max_batch, args.storage, args.warm,
                  args.warm_outer, autotune=not args.no_autotune, device_hpb=args.device_hpb, images=args.images,
                  fill_holes=args.fill_holes, graph=not args.no_graph, save_schedules=args.save_schedules,
                  load_schedules=args.load_schedules, compare=args.served_compare if args.warm else 0)
    out = {
        "metric": "served scene-flow pairs/s and request latency",
        "value": r["pairs_per_s"],
        "unit": "pairs/s",
        "n_gpus": 1,
        "higher_is_better": True,
        "dtype": "f32" if args.storage == "f32" else f"f32 compute, {args.storage} system storage",
        "data": "synthetic (seeded frame sequences, smooth per-segment SE(3) motion)",
        "config": dict(streams=args.serve, fps_per_stream=args.fps, arrival_seconds=args.seconds,
                       max_batch=args.max_batch, lanes=2, warm_start=args.warm, warm_n_outer=args.warm_outer,
                       storage=args.storage, images=args.images, fill_holes=args.fill_holes,
                       graph=not args.no_graph, **r),
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
