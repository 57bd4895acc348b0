"""Precision variants of the scene-flow solve, measured on one GPU.

Usage: python tests/precision_report.py [--out FILE] [--pairs N]

Workload: a batch of mixed-size synthetic pairs (sizes drawn from 320x240,
640x480, 960x540 and 1280x720, intrinsics scaled by W/640, 2-8 moving
segments), 5-level pyramids at ratio 0.5, 15 outer x 50 inner iterations,
boundary-probability regularisation.  Variants: the storage of the per-pixel
systems (f32 as the reference, f16 with a per-pixel power-of-two scale, bf16;
the arithmetic is f32 in every variant), with the feature images (frame 1,
frame 2, warped frame 2) in f64 as the reference or in the same 16-bit
format, times the relaxation's update rule
(the reference's: factor 0.7, a pixel's unknowns from each other's old
values; point SOR: factor 1.8, u then v then w from the newest values;
block SOR: factor 1.8, the pixel's 3x3 block solved as a whole).

Reported per variant: batch pairs/s; flow distance to the f32 reference-rule
result (EPE of u, v in px; |dw| in m); flow distance to the FP64 oracle of
the same rule (every buffer f64, on the host's cores); distance to the
generator's true flow.  Then, on fixed level
systems of each size, the iterations each variant needs to bring the
relative residual ||r - M du|| / ||r|| to 1e-4 (the FP64 residual exit, at
most 20000) and the energy it reaches.

Then the served warm starts (four streams, one per size, FP16 records and
images, warm starts of 4 outer iterations): each warm pair's served flow
against the FP64 oracle of the configured work from the same estimate (15
outer iterations) and coarse to fine (15 x 50), and against the truth.

This script is the checker's view of the variants: it uses the oracle, so
it lives with the tests, not with the product.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
from helpers import level_inputs, load_package, oracle  # noqa: E402

SIZES = [(320, 240), (640, 480), (960, 540), (1280, 720)]


def workload(R, n, seed=300):
    import rgbdflow_amd.synth as S
    O = oracle()
    rng = np.random.default_rng(seed)
    pairs = []
    for i in range(n):
        w, h = SIZES[i % 4] if i < 4 else SIZES[int(rng.integers(0, 4))]
        f = 525.0 * w / 640
        cam = S.Camera(w, h, f, f, (w - 1) / 2, (h - 1) / 2)
        pr = S.make_pair(w, h, seed + i, n_segments=int(rng.integers(2, 9)), cam=cam)
        im1, d1, im2, d2 = S.to_solver_inputs(pr)
        lv = R.pyramid_levels(w, h, 0.5, w // 16)
        hpb = S.segment_boundary_maps(pr["seg1"], [a for a, b in lv], [b for a, b in lv])
        pairs.append(dict(im1=im1, dm1=d1.astype(np.float64), im2=im2, dm2=d2.astype(np.float64), hpb=hpb,
                          sigma1=O.depth_stdevs(d1, 4), cam=R.make_params(w, h, f=f).cam, min_width=w // 16,
                          W=w, H=h, f=f, truth=pr["flow"].astype(np.float64)))
    return pairs


def epe_stats(flows, refs):
    e = np.concatenate([np.sqrt((a[0] - b[0]) ** 2 + (a[1] - b[1]) ** 2).ravel() for a, b in zip(flows, refs)])
    dw = np.concatenate([np.abs(a[2] - b[2]).ravel() for a, b in zip(flows, refs)])
    return dict(epe_median=float(np.median(e)), epe_mean=float(e.mean()), epe_p99=float(np.percentile(e, 99)),
                dw_median=float(np.median(dw)), dw_mean=float(dw.mean()), dw_p99=float(np.percentile(dw, 99)))


def served_vs_fp64(R, O, warm_outer=4, n_frames=5, seed=900):
    """Four served streams, one per size, FP16 records and images, boundary
    maps made on the device, warm starts of warm_outer outer iterations: every
    warm pair's served flow against the FP64 oracle of the same pair solved
    as sceneFlowWithInit from the estimate the server used with the
    configured 15 outer iterations (sceneFlowWithInit.cpp:74-80), and coarse
    to fine at 15 x 50; the truth beside them."""
    import rgbdflow_amd.synth as S
    p = R.make_params(640, 480, n_outer=15, n_sor=50, min_width=40, relax_variant=R.RELAX_TILED, graph=True,
                      storage=R.STORE_F16, image_storage=R.IMAGES_F16)
    streams, flows = [], {}
    with R.Server(p, max_batch=4, lanes=2, warm_start=True, warm_n_outer=warm_outer) as srv:
        for k, (w, h) in enumerate(SIZES):
            f = 525.0 * w / 640
            seq = S.make_sequence(w, h, seed + k, n_frames, n_segments=6,
                                  cam=S.Camera(w, h, f, f, (w - 1) / 2, (h - 1) / 2))
            fr = [(x["img"], (x["depth"].astype(np.float64) * .001).astype(np.float32)) for x in seq["frames"]]
            sid = srv.open(w, h, R.make_params(w, h).cam, w // 16)
            streams.append((sid, w, h, seq, fr))
        tickets = [(k, n, srv.submit(st[0], st[4][n][0], st[4][n][1])) for n in range(n_frames)
                   for k, st in enumerate(streams)]
        for (k, n, tk) in tickets:
            info, flow = srv.poll(tk, wait_ms=120000)
            if info["has_flow"]:
                flows[(k, n)] = flow
    served, w15, cold, truth = [], [], [], []
    for k, (sid, w, h, seq, fr) in enumerate(streams):
        for n in range(2, n_frames):
            a, b = fr[n - 1], fr[n]
            im1, im2 = a[0].astype(np.float64) / 255.0, b[0].astype(np.float64) / 255.0
            d1, d2 = a[1].astype(np.float64), b[1].astype(np.float64)
            sig = O.depth_stdevs(a[1], 4)
            est = np.stack(flows[(k, n - 1)], axis=-1)
            kw = dict(n_outer=15, n_sor=50, min_width=w // 16, precision=O.PREC_F64)
            served.append(flows[(k, n)])
            w15.append(O.with_init(im1, d1, im2, d2, sig, None, est, O.make_params(w, h, **kw))[:3])
            cold.append(O.multiscale(im1, d1, im2, d2, sig, None, O.make_params(w, h, **kw))[:3])
            gt = S.sequence_flow(seq, n - 1, n)
            truth.append((gt[..., 0], gt[..., 1], gt[..., 2]))
        print(f"precision: served {w}x{h} compared", flush=True)
    return dict(workload=f"4 served streams ({', '.join(f'{w}x{h}' for w, h in SIZES)}), {n_frames} frames each, "
                         f"FP16 records and images, device boundary maps, warm starts of {warm_outer} outer "
                         f"iterations; {len(served)} warm pairs",
                vs_fp64_warm_15_outer=epe_stats(served, w15), vs_fp64_cold_15x50=epe_stats(served, cold),
                vs_truth=epe_stats(served, truth), fp64_warm_15_vs_truth=epe_stats(w15, truth),
                fp64_cold_vs_truth=epe_stats(cold, truth))


def coarse_report(R, O, n=4, steps=3):
    """The coarse-grid correction on 640x480 pairs (seeds 21.., 6 segments)
    with boundary maps made from the frames, as sceneFlowMultiscale.cpp:116
    makes them: per rule with and without the correction, batch pairs/s,
    pair 0's sweeps per level, and the flow's distance to the FP64 oracle of
    the reference rule, to the FP64 oracle of the same variant and to the
    truth."""
    import rgbdflow_amd.synth as S
    W, H = 640, 480
    cam = S.Camera(W, H, 525.0, 525.0, 319.5, 239.5)
    pairs = []
    for i in range(n):
        pr = S.make_pair(W, H, 21 + i, n_segments=6, cam=cam)
        im1, d1, im2, d2 = S.to_solver_inputs(pr)
        pairs.append(dict(im1=im1, dm1=d1.astype(np.float64), im2=im2, dm2=d2.astype(np.float64), hpb=None,
                          sigma1=O.depth_stdevs(d1.astype(np.float32), 4), truth=pr["flow"].astype(np.float64)))
    variants = {"reference": (dict(omega=0.7, pixel_update=R.PIXEL_JACOBI), dict(omega=0.7, pixel_gs=0), 0),
                "reference+coarse8": (dict(omega=0.7, pixel_update=R.PIXEL_JACOBI), dict(omega=0.7, pixel_gs=0), 8),
                "block_sor_1.8": (dict(omega=1.8, pixel_update=R.PIXEL_BLOCK), dict(omega=1.8, pixel_gs=2), 0),
                "block_sor_1.8+coarse8": (dict(omega=1.8, pixel_update=R.PIXEL_BLOCK), dict(omega=1.8, pixel_gs=2), 8)}
    base = dict(f=525.0, cx=319.5, cy=239.5, n_outer=15, n_sor=50, min_width=40)
    ctx = R.Context(0)
    for (lw, lh) in R.pyramid_levels(W, H, 0.5, 40):
        ctx.autotune_relax(lw, lh, batch=n)
    truth = [(p["truth"][..., 0], p["truth"][..., 1], p["truth"][..., 2]) for p in pairs]
    out, f64 = [], {}
    for name, (grule, orule, levels) in variants.items():
        p = R.make_params(W, H, relax_variant=R.RELAX_TILED, graph=True, coarse_levels=levels, **grule, **base)
        ctx.batch_load(p, pairs)
        st = ctx.batch_run(p, stats=True)
        ctx.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            ctx.batch_run(p, stats=False)
        ctx.synchronize()
        pps = steps * n / (time.perf_counter() - t0)
        flows = [ctx.batch_read(i) for i in range(n)]
        t0 = time.perf_counter()
        o = []
        for pr in pairs:
            po = O.make_params(W, H, precision=O.PREC_F64, cg_levels=levels, **orule, **base)
            u, v, w, _ = O.multiscale(pr["im1"], pr["dm1"], pr["im2"], pr["dm2"], pr["sigma1"], None, po)
            o.append((u, v, w))
        f64[name] = o
        L = st[0].n_levels
        out.append({"variant": name, "pairs_per_s": round(pps, 2),
                    "sweeps_exec_pair0": [st[0].sweeps_exec[k] for k in range(L)],
                    "vs_fp64_oracle_same_variant": epe_stats(flows, o), "vs_truth": epe_stats(flows, truth)})
        print(f"coarse: {name}: {pps:.1f} pairs/s, FP64 oracle {time.perf_counter() - t0:.1f} s", flush=True)
    for r in out:
        r["vs_fp64_oracle_reference_rule"] = epe_stats([f for f in f64[r["variant"]]], f64["reference"])
    ctx.close()
    return dict(workload=f"{n} x 640x480 pairs (seeds 21..{20 + n}, 6 segments), 5-level pyramids ratio 0.5, "
                         "15 outer x 50 inner, HPB regularisation, boundary maps made per level from the frames",
                note="vs_fp64_oracle_reference_rule compares the variant's FP64 oracle flow with the reference "
                     "rule's FP64 oracle flow (what the variant changes, free of f32 rounding)", variants=out)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--pairs", type=int, default=8)
    ap.add_argument("--out", default="")
    ap.add_argument("--no-served", action="store_true", help="skip the served warm-start comparison")
    ap.add_argument("--served-only", action="store_true", help="only the served warm-start comparison")
    ap.add_argument("--coarse-only", action="store_true",
                    help="only the coarse-grid correction report (boundary maps made from the frames)")
    ap.add_argument("--threads", type=int, default=0, help="oracle threads (0: the cores this process may use)")
    args = ap.parse_args()
    R = load_package()
    O = oracle()
    nt = args.threads or min(len(os.sched_getaffinity(0)), int(os.environ.get("OMP_NUM_THREADS", "64") or 64))
    O.set_threads(max(1, nt))
    if args.served_only or args.coarse_only:
        rep = served_vs_fp64(R, O) if args.served_only else coarse_report(R, O)
        txt = json.dumps(rep, indent=1)
        print(txt)
        if args.out:
            with open(args.out, "w") as f:
                f.write(txt + "\n")
        return
    pairs = workload(R, args.pairs)
    print(f"precision: {len(pairs)} pairs " + ", ".join(f"{p['W']}x{p['H']}" for p in pairs), flush=True)
    base = dict(n_outer=15, n_sor=50, min_width=40, relax_variant=R.RELAX_TILED, graph=True)
    rules = {"reference": dict(omega=0.7, pixel_update=R.PIXEL_JACOBI),
             "point_sor_1.8": dict(omega=1.8, pixel_update=R.PIXEL_GS),
             "block_sor_1.8": dict(omega=1.8, pixel_update=R.PIXEL_BLOCK)}
    stores = {"f32": R.STORE_F32, "f16": R.STORE_F16, "bf16": R.STORE_BF16}
    # (system records, feature images) of the batch variants
    batch_variants = {"f32": (R.STORE_F32, R.IMAGES_F64), "f16": (R.STORE_F16, R.IMAGES_F64),
                      "bf16": (R.STORE_BF16, R.IMAGES_F64), "f16+f16_images": (R.STORE_F16, R.IMAGES_F16),
                      "bf16+bf16_images": (R.STORE_BF16, R.IMAGES_BF16)}
    ctx = R.Context(0)
    for p in pairs[:4]:
        for (lw, lh) in R.pyramid_levels(p["W"], p["H"], 0.5, p["min_width"]):
            ctx.autotune_relax(lw, lh, batch=max(1, len(pairs) // 4))
    results, flows = [], {}
    for rname, rule in rules.items():
        for sname, (st, img) in batch_variants.items():
            if rule["pixel_update"] == R.PIXEL_BLOCK and st != R.STORE_F32:
                continue      # the block update needs f32 records
            p = R.make_params(640, 480, storage=st, image_storage=img, **rule, **base)
            ctx.batch_load(p, pairs)
            ctx.batch_run(p, stats=True)
            ctx.batch_run(p, stats=False)
            ctx.synchronize()
            t0 = time.perf_counter()
            for _ in range(3):
                ctx.batch_run(p, stats=False)
            ctx.synchronize()
            pps = 3 * len(pairs) / (time.perf_counter() - t0)
            flows[(rname, sname)] = [ctx.batch_read(i) for i in range(len(pairs))]
            results.append({"rule": rname, "storage": sname, "pairs_per_s": round(pps, 2)})
            print(f"precision: {rname} {sname}: {pps:.1f} pairs/s", flush=True)
    ref = flows[("reference", "f32")]
    # the FP64 oracle of each rule (reference exits), on the host
    f64 = {}
    for rname, rule in rules.items():
        out = []
        t0 = time.perf_counter()
        for p in pairs:
            po = O.make_params(p["W"], p["H"], f=p["f"], n_outer=15, n_sor=50, min_width=p["min_width"],
                               precision=O.PREC_F64, omega=rule["omega"],
                               pixel_gs=rule["pixel_update"])
            u, v, w, _ = O.multiscale(p["im1"], p["dm1"], p["im2"], p["dm2"], p["sigma1"], p["hpb"], po)
            out.append((u, v, w))
        f64[rname] = out
        print(f"precision: FP64 oracle, {rname}: {time.perf_counter() - t0:.1f} s on {nt} threads", flush=True)
    for r in results:
        fl = flows[(r["rule"], r["storage"])]
        r["vs_f32_reference"] = epe_stats(fl, ref)
        r["vs_fp64_oracle"] = epe_stats(fl, f64[r["rule"]])
        r["vs_truth"] = epe_stats(fl, [(p["truth"][..., 0], p["truth"][..., 1], p["truth"][..., 2]) for p in pairs])
    # iterations to a relative residual of 1e-4 on fixed level systems (the
    # system of a level's first outer iteration at 0.8 x the true flow), per
    # size, storage and rule; the converged energies beside them
    conv = []
    for (w, h) in SIZES:
        li, (u, v, w_), _ = level_inputs(seed=8, W=w, H=h)
        coef = O.level_system(li, O.make_params(w, h), u, v, w_)
        x0 = np.zeros((3, h, w))
        for rname, rule in rules.items():
            for sname, st in stores.items():
                if rule["pixel_update"] == R.PIXEL_BLOCK and st != R.STORE_F32:
                    continue
                sched = R.make_params(w, h, relax_variant=R.RELAX_TILED, storage=st, **rule)
                _, it, rs, e = ctx.relax_residual(coef, 20000, 1e-4, x0, sched)
                conv.append(dict(level=f"{w}x{h}", rule=rname, storage=sname, iterations=it,
                                 reached=bool(0 <= rs <= 1e-4), residual=rs, energy=e))
        print(f"precision: convergence on the {w}x{h} system done", flush=True)
    for c in conv:
        for rn in ("point_sor_1.8", "reference"):
            e0 = [d["energy"] for d in conv if d["level"] == c["level"] and d["rule"] == rn and
                  d["storage"] == "f32"][0]
            c["energy_rel_to_" + ("sor" if rn == "point_sor_1.8" else rn) + "_f32"] = abs(c["energy"] - e0) / abs(e0)
    ctx.close()
    served = None if args.no_served else served_vs_fp64(R, O)
    report = dict(workload=f"{len(pairs)} mixed-size pairs (" + ", ".join(f"{p['W']}x{p['H']}" for p in pairs) +
                  "), 5-level pyramids ratio 0.5, 15 outer x 50 inner, HPB regularisation",
                  oracle_threads=nt, variants=results, convergence_to_residual_1e_4=conv, served_warm_starts=served)
    txt = json.dumps(report, indent=1)
    print(txt)
    if args.out:
        with open(args.out, "w") as f:
            f.write(txt + "\n")


if __name__ == "__main__":
    main()
