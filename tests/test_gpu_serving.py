"""The serving runtime (rgbdflow_server_*): frame streams of different sizes
submitted together, solved in continuously formed batches on two lanes; each
pair's flow equals the same pair solved alone (rgbdflow_multiscale), bit for
bit, however the requests were batched."""
import numpy as np
import pytest

from helpers import load_package, synth

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def R():
    return load_package()


def _stream_frames(S, W, H, seed):
    """Two frames A, B of a synthetic scene, in the submission format (u8 BGR,
    f32 depth), with boundary maps for each as a frame 1."""
    pair = S.make_pair(W, H, seed, 4)
    im1, d1, im2, d2 = S.to_solver_inputs(pair)
    frames = [(pair["img1"], d1), (pair["img2"], d2)]
    return frames, pair


def _maps(R, S, pair, W, H, mw, shift):
    lv = R.pyramid_levels(W, H, 0.5, mw)
    seg = np.roll(pair["seg1"], shift, axis=1)  # a different (still valid) map per frame
    return S.segment_boundary_maps(seg, [a for a, b in lv], [b for a, b in lv])


def _alone(R, ctx, p, W, H, mw, f1, f2, hpb1):
    q = R.make_params(W, H, n_outer=p.n_outer, n_sor=p.n_sor, min_width=mw, relax_variant=p.relax_variant,
                      storage=p.storage, image_storage=p.image_storage, coarse_levels=p.coarse_levels,
                      coarse_sweeps=p.coarse_sweeps, coarse_at=p.coarse_at)
    im1 = f1[0].astype(np.float64) / 255.0
    im2 = f2[0].astype(np.float64) / 255.0
    u, v, w, _ = ctx.multiscale(q, im1, f1[1].astype(np.float64), im2, f2[1].astype(np.float64), None, hpb1)
    return u, v, w


@pytest.mark.parametrize("max_batch,lanes", [(8, 2), (2, 2), (1, 1)])
@pytest.mark.parametrize("mode", ["maps", "device_maps", "fp16", "coarse"])
def test_serving_results_equal_alone(R, max_batch, lanes, mode):
    """mode: boundary maps submitted with the frames; made on the device
    (frames submitted without maps; the pair alone is loaded without them
    too); FP16 records and feature images with device maps; or the
    coarse-grid correction with device maps (batches of several sizes share
    its launches)."""
    S = synth()
    sizes = [(128, 96, 16), (160, 120, 20), (96, 72, 12)]
    extra = dict(storage=R.STORE_F16, image_storage=R.IMAGES_F16) if mode == "fp16" else {}
    if mode == "coarse":
        extra = dict(coarse_levels=8, coarse_sweeps=2, coarse_at=4)
    p = R.make_params(128, 96, n_outer=3, n_sor=25, min_width=16, relax_variant=R.RELAX_TILED, **extra)
    streams = []
    with R.Server(p, max_batch=max_batch, lanes=lanes) as srv:
        for k, (W, H, mw) in enumerate(sizes):
            frames, pair = _stream_frames(S, W, H, 200 + k)
            maps = [_maps(R, S, pair, W, H, mw, sh) if mode == "maps" else None for sh in (0, 3)]
            sid = srv.open(W, H, R.make_params(W, H).cam, mw)
            streams.append((sid, W, H, mw, frames, maps))
        # frames A, B, A per stream, interleaved across streams
        tickets = []
        for t, fi in enumerate([0, 1, 0]):
            for (sid, W, H, mw, frames, maps) in streams:
                tickets.append((sid, t, fi, srv.submit(sid, frames[fi][0], frames[fi][1], maps[fi])))
        results = {}
        for (sid, t, fi, tk) in tickets:
            print(f"poll stream {sid} frame {t}", flush=True)
            info, flow = srv.poll(tk, wait_ms=60000)
            assert info["status"] == 0
            assert info["has_flow"] == (t > 0)
            results[(sid, t)] = flow
    with R.Context(0) as ctx:
        for (sid, W, H, mw, frames, maps) in streams:
            for t, (a, b) in [(1, (0, 1)), (2, (1, 0))]:
                u, v, w = _alone(R, ctx, p, W, H, mw, frames[a], frames[b], maps[a])
                ub, vb, wb = results[(sid, t)]
                assert np.array_equal(u, ub) and np.array_equal(v, vb) and np.array_equal(w, wb), (sid, t)


def test_serving_budget_and_errors(R):
    S = synth()
    p = R.make_params(96, 72, n_outer=2, n_sor=25, min_width=12, reg=R.REG_LETOUZEY)
    # a budget that holds the stream state and about one pair in flight at a time
    with R.Server(p, max_batch=8, lanes=2, hbm_budget=8 << 20) as srv:
        frames, pair = _stream_frames(S, 96, 72, 300)
        sids = [srv.open(96, 72, None, 12) for _ in range(3)]
        tk = [srv.submit(s, frames[i % 2][0], frames[i % 2][1]) for i in range(2) for s in sids]
        for t in tk:
            info, flow = srv.poll(t, wait_ms=60000)
            assert info["status"] == 0
        with pytest.raises(R.RgbdFlowError):
            srv.submit(12345, frames[0][0], frames[0][1])   # no such stream
        with pytest.raises(R.RgbdFlowError):
            srv.open(96, 71, None, 12)                      # camera size differs (params.cam is 96x72)
        for s in sids:
            srv.close_stream(s)


@pytest.mark.parametrize("max_batch,lanes", [(8, 2), (1, 1)])
@pytest.mark.parametrize("mode", ["maps", "device_maps"])
def test_serving_warm_start_equals_with_init(R, max_batch, lanes, mode):
    """With warm starts a stream's first pair is solved coarse to fine and
    every later pair as sceneFlowWithInit from the stream's previous result
    (here with 2 outer iterations): each result equals rgbdflow_with_init of
    that pair alone with that estimate, bit for bit.  Batches repeat their
    composition, so later ones replay a captured HIP graph of the plan."""
    S = synth()
    sizes = [(128, 96, 16), (96, 72, 12)]
    p = R.make_params(128, 96, n_outer=3, n_sor=25, min_width=16, relax_variant=R.RELAX_TILED, graph=True)
    order = [0, 1, 0, 1, 0, 1]
    streams = []
    with R.Server(p, max_batch=max_batch, lanes=lanes, warm_start=True, warm_n_outer=2) as srv:
        for k, (W, H, mw) in enumerate(sizes):
            frames, pair = _stream_frames(S, W, H, 400 + k)
            maps = [_maps(R, S, pair, W, H, mw, sh) if mode == "maps" else None for sh in (0, 3)]
            sid = srv.open(W, H, R.make_params(W, H).cam, mw)
            streams.append((sid, W, H, mw, frames, maps))
        tickets = []
        for t, fi in enumerate(order):
            for (sid, W, H, mw, frames, maps) in streams:
                tickets.append((sid, t, srv.submit(sid, frames[fi][0], frames[fi][1], maps[fi])))
        results = {}
        for (sid, t, tk) in tickets:
            print(f"poll stream {sid} frame {t}", flush=True)
            info, flow = srv.poll(tk, wait_ms=60000)
            assert info["status"] == 0 and info["has_flow"] == (t > 0)
            results[(sid, t)] = flow
    with R.Context(0) as ctx:
        for (sid, W, H, mw, frames, maps) in streams:
            u, v, w = _alone(R, ctx, p, W, H, mw, frames[order[0]], frames[order[1]], maps[order[0]])
            ub, vb, wb = results[(sid, 1)]
            assert np.array_equal(u, ub) and np.array_equal(v, vb) and np.array_equal(w, wb), (sid, 1)
            q = R.make_params(W, H, n_outer=2, n_sor=25, min_width=mw, relax_variant=R.RELAX_TILED)
            for t in range(2, len(order)):
                a, b = order[t - 1], order[t]
                est = np.stack(results[(sid, t - 1)], axis=-1)
                im1 = frames[a][0].astype(np.float64) / 255.0
                im2 = frames[b][0].astype(np.float64) / 255.0
                u, v, w, st = ctx.with_init(q, im1, frames[a][1].astype(np.float64), im2,
                                            frames[b][1].astype(np.float64), est, None,
                                            maps[a][: W * H * 2] if maps[a] is not None else None)
                ub, vb, wb = results[(sid, t)]
                assert np.array_equal(u, ub) and np.array_equal(v, vb) and np.array_equal(w, wb), (sid, t)


@pytest.mark.parametrize("max_batch,lanes", [(8, 2), (1, 1)])
@pytest.mark.parametrize("mode", ["maps", "device_maps"])
def test_serving_fill_holes_equals_alone(R, max_batch, lanes, mode):
    """fill_holes: frames whose depth has holes; each result equals the pair
    solved alone with both depth maps filled by rgbdflow_fill_depth_holes,
    frame 1's depth stdevs computed from its depth as submitted (holes
    included) and -- boundary maps made on the device -- those maps made from
    frame 1's depth with its small holes filled by
    rgbdflow_fill_small_holes_colour, bit for bit."""
    S = synth()
    sizes = [(128, 96, 16), (96, 72, 12)]
    p = R.make_params(128, 96, n_outer=3, n_sor=25, min_width=16, relax_variant=R.RELAX_TILED)
    streams = []
    with R.Server(p, max_batch=max_batch, lanes=lanes, fill_holes=True) as srv:
        for k, (W, H, mw) in enumerate(sizes):
            frames, pair = _stream_frames(S, W, H, 500 + k)
            rng = np.random.default_rng(k)
            holed = []
            for img, d in frames:
                d = d.copy()
                d[rng.random(d.shape) < 0.03] = 0.0
                d[H // 3:H // 3 + 5, W // 4:W // 4 + 7] = 0.0
                d[H // 2:H // 2 + 9, W // 2:W // 2 + 9] = 0.0      # above 60 pixels: no colour fill
                holed.append((img, d))
            maps = [_maps(R, S, pair, W, H, mw, sh) if mode == "maps" else None for sh in (0, 3)]
            sid = srv.open(W, H, R.make_params(W, H).cam, mw)
            streams.append((sid, W, H, mw, holed, maps))
        tickets = []
        for t, fi in enumerate([0, 1, 0]):
            for (sid, W, H, mw, frames, maps) in streams:
                tickets.append((sid, t, srv.submit(sid, frames[fi][0], frames[fi][1], maps[fi])))
        results = {}
        for (sid, t, tk) in tickets:
            print(f"poll stream {sid} frame {t}", flush=True)
            info, flow = srv.poll(tk, wait_ms=60000)
            assert info["status"] == 0 and info["has_flow"] == (t > 0)
            results[(sid, t)] = flow
    with R.Context(0) as ctx:
        for (sid, W, H, mw, frames, maps) in streams:
            q = R.make_params(W, H, n_outer=p.n_outer, n_sor=p.n_sor, min_width=mw, relax_variant=p.relax_variant)
            for t, (a, b) in [(1, (0, 1)), (2, (1, 0))]:
                fa, la = ctx.fill_depth_holes(frames[a][1].astype(np.float64))
                fb, lb = ctx.fill_depth_holes(frames[b][1].astype(np.float64))
                assert la == 0 and lb == 0
                sa = ctx.depth_stdevs(frames[a][1])
                im_a = frames[a][0].astype(np.float64) / 255.0
                pr = dict(im1=im_a, dm1=fa, im2=frames[b][0].astype(np.float64) / 255.0, dm2=fb, sigma1=sa,
                          hpb=maps[a], cam=R.make_params(W, H).cam, min_width=mw)
                if maps[a] is None:
                    hd, _ = ctx.fill_small_holes_colour(frames[a][1].astype(np.float64), im_a)
                    assert (hd[H // 2 + 4, W // 2 + 4] <= 0) and (hd[H // 3 + 2, W // 4 + 3] > 0)
                    pr["hpb_dm1"] = hd
                ctx.batch_load(q, [pr])
                ctx.batch_run(q, stats=False)
                u, v, w = ctx.batch_read(0)
                ub, vb, wb = results[(sid, t)]
                assert np.array_equal(u, ub) and np.array_equal(v, vb) and np.array_equal(w, wb), (sid, t)


def test_serving_warm_fill_holes_maps_from_the_raw_depth(R):
    """Warm starts with fill_holes and boundary maps made on the device: a
    warm pair is sceneFlowWithInit, whose maps come from frame 1's depth as
    measured (computeHeuristicPb(Im1, dm1), sceneFlowWithInit.cpp:47), while
    everything else uses the hole-filled depths (:19-21).  Each warm result
    equals the pair solved alone that way, bit for bit."""
    S = synth()
    W, H, mw = 128, 96, 16
    p = R.make_params(W, H, n_outer=3, n_sor=25, min_width=mw, relax_variant=R.RELAX_TILED, graph=True)
    frames, pair = _stream_frames(S, W, H, 610)
    rng = np.random.default_rng(6)
    holed = []
    for img, d in frames:
        d = d.copy()
        d[rng.random(d.shape) < 0.03] = 0.0
        d[H // 3:H // 3 + 5, W // 4:W // 4 + 7] = 0.0
        holed.append((img, d))
    order = [0, 1, 0, 1]
    with R.Server(p, max_batch=4, lanes=2, warm_start=True, warm_n_outer=2, fill_holes=True) as srv:
        sid = srv.open(W, H, R.make_params(W, H).cam, mw)
        tickets = [(t, srv.submit(sid, holed[fi][0], holed[fi][1], None)) for t, fi in enumerate(order)]
        results = {}
        for (t, tk) in tickets:
            print(f"poll frame {t}", flush=True)
            info, flow = srv.poll(tk, wait_ms=60000)
            assert info["status"] == 0 and info["has_flow"] == (t > 0)
            results[t] = flow
    q = R.make_params(W, H, n_outer=2, n_sor=25, min_width=mw, relax_variant=R.RELAX_TILED)
    with R.Context(0) as ctx:
        for t in range(2, len(order)):
            a, b = order[t - 1], order[t]
            fa, _ = ctx.fill_depth_holes(holed[a][1].astype(np.float64))
            fb, _ = ctx.fill_depth_holes(holed[b][1].astype(np.float64))
            pr = dict(im1=holed[a][0].astype(np.float64) / 255.0, dm1=fa, im2=holed[b][0].astype(np.float64) / 255.0,
                      dm2=fb, sigma1=ctx.depth_stdevs(holed[a][1]), hpb=None, cam=R.make_params(W, H).cam,
                      min_width=mw, est=np.stack(results[t - 1], axis=-1),
                      hpb_dm1=holed[a][1].astype(np.float64))
            ctx.batch_load(q, [pr])
            ctx.batch_run(q, stats=False)
            u, v, w = ctx.batch_read(0)
            ub, vb, wb = results[t]
            assert np.array_equal(u, ub) and np.array_equal(v, vb) and np.array_equal(w, wb), t
