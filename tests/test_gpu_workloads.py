"""GPU parity at the sizes of the benchmarked workloads: one pair each of
320x240, 640x480, 960x540 and 1280x720 (minimum width W/16: levels such as
120x67, 60x33 and 80x45; 2-8 moving segments; 15 outer x 50 inner), solved

- as one mixed batch (one launch per step for all four, tiled relaxation,
  the schedule replayed as a HIP graph), in the reference's precision and in
  the FP16 and BF16 variants (16-bit system records and feature images):
  every pair equals the oracle run the same way, bit for bit, with the same
  iterations per level;
- as four served streams, one per size (boundary maps made on the device,
  FP16 records and images, warm starts of 4 outer iterations): every served
  result equals that pair solved alone, the first coarse to fine and the
  later ones as sceneFlowWithInit from the stream's previous result.
"""
import numpy as np
import pytest

from helpers import hpb_for, load_package, make_case, oracle

pytestmark = pytest.mark.gpu

SIZES = [(320, 240, 2), (640, 480, 5), (960, 540, 8), (1280, 720, 3)]   # (W, H, moving segments)
N_OUTER, N_SOR = 15, 50


@pytest.fixture(scope="module")
def R():
    return load_package()


@pytest.fixture(scope="module")
def cases():
    O = oracle()
    O.set_threads(16)
    out = []
    for k, (W, H, seg) in enumerate(SIZES):
        print(f"generating {W}x{H} with {seg} segments", flush=True)
        pair, im1, d1, im2, d2, sigma = make_case(W, H, 700 + k, seg)
        hpb, lv = hpb_for(pair, W, H, 0.5, W // 16)
        out.append(dict(W=W, H=H, pair=pair, im1=im1, dm1=d1, im2=im2, dm2=d2, sigma1=sigma, hpb=hpb,
                        min_width=W // 16, levels=lv))
    return out


VARIANTS = {"f32": (0, 0), "fp16": (1, 1), "bf16": (2, 2)}   # (record storage, feature-image storage)


@pytest.mark.parametrize("variant", list(VARIANTS))
def test_mixed_batch_matches_oracle(R, cases, variant):
    O = oracle()
    rec, img = VARIANTS[variant]
    p = R.make_params(640, 480, n_outer=N_OUTER, n_sor=N_SOR, min_width=40, relax_variant=R.RELAX_TILED,
                      graph=True, storage=rec, image_storage=img)
    pairs = [dict(im1=c["im1"], dm1=c["dm1"], im2=c["im2"], dm2=c["dm2"], sigma1=c["sigma1"], hpb=c["hpb"],
                  cam=R.make_params(c["W"], c["H"]).cam, min_width=c["min_width"]) for c in cases]
    with R.Context(0) as ctx:
        print(f"{variant}: batch of {len(pairs)} mixed sizes", flush=True)
        ctx.batch_load(p, pairs)
        stats = ctx.batch_run(p, stats=True)
        first = [ctx.batch_read(i) for i in range(len(cases))]
        ctx.batch_run(p, stats=False)       # captures the schedule as a HIP graph
        ctx.batch_run(p, stats=False)       # replays it
        ctx.synchronize()
        replay = [ctx.batch_read(i) for i in range(len(cases))]
    for i, c in enumerate(cases):
        W, H = c["W"], c["H"]
        po = O.make_params(W, H, n_outer=N_OUTER, n_sor=N_SOR, min_width=c["min_width"], record_storage=rec,
                           image_storage=img)
        ua, va, wa, sa = O.multiscale(c["im1"], c["dm1"], c["im2"], c["dm2"], c["sigma1"], c["hpb"], po)
        sb = stats[i]
        L = sa.n_levels
        lv = [(sb.width[k], sb.height[k]) for k in range(sb.n_levels)]
        print(f"{variant} {W}x{H}: levels {lv}, sweeps {[sb.sweeps_exec[k] for k in range(L)]}, "
              f"outer {[sb.outer_exec[k] for k in range(L)]}", flush=True)
        assert lv == c["levels"]
        assert [sb.outer_exec[k] for k in range(L)] == [sa.outer_exec[k] for k in range(L)]
        assert [sb.sweeps_exec[k] for k in range(L)] == [sa.sweeps_exec[k] for k in range(L)]
        for got in (first[i], replay[i]):
            assert np.array_equal(ua, got[0]) and np.array_equal(va, got[1]) and np.array_equal(wa, got[2]), (i, W)
        if variant == "f32":
            # the task's bound (implied by the equality above) and the distance to the FP64 oracle
            pf = O.make_params(W, H, n_outer=N_OUTER, n_sor=N_SOR, min_width=c["min_width"], precision=O.PREC_F64)
            uf, vf, wf, _ = O.multiscale(c["im1"], c["dm1"], c["im2"], c["dm2"], c["sigma1"], c["hpb"], pf)
            epe = np.hypot(ua - uf, va - vf)
            print(f"  vs FP64 oracle: EPE median {np.median(epe):.3g} px, |dw| median "
                  f"{np.median(np.abs(wa - wf)):.3g} m", flush=True)


def test_served_streams_of_every_size_equal_alone(R, cases):
    p = R.make_params(640, 480, n_outer=N_OUTER, n_sor=N_SOR, min_width=40, relax_variant=R.RELAX_TILED,
                      graph=True, storage=R.STORE_F16, image_storage=R.IMAGES_F16)
    order = [0, 1, 0]
    streams = []
    with R.Server(p, max_batch=4, lanes=2, warm_start=True, warm_n_outer=4) as srv:
        for c in cases:
            W, H = c["W"], c["H"]
            frames = [(c["pair"]["img1"], c["dm1"].astype(np.float32)), (c["pair"]["img2"], c["dm2"].astype(np.float32))]
            sid = srv.open(W, H, R.make_params(W, H).cam, c["min_width"])
            streams.append((sid, c, frames))
        tickets = []
        for t, fi in enumerate(order):
            for (sid, c, frames) in streams:
                tickets.append((sid, t, srv.submit(sid, frames[fi][0], frames[fi][1], None)))
        results = {}
        for (sid, t, tk) in tickets:
            print(f"poll stream {sid} frame {t}", flush=True)
            info, flow = srv.poll(tk, wait_ms=120000)
            assert info["status"] == 0 and info["has_flow"] == (t > 0)
            results[(sid, t)] = flow
    with R.Context(0) as ctx:
        for (sid, c, frames) in streams:
            W, H, mw = c["W"], c["H"], c["min_width"]
            q = R.make_params(W, H, n_outer=N_OUTER, n_sor=N_SOR, min_width=mw, relax_variant=R.RELAX_TILED,
                              storage=R.STORE_F16, image_storage=R.IMAGES_F16)
            ims = [f[0].astype(np.float64) / 255.0 for f in frames]
            dms = [f[1].astype(np.float64) for f in frames]
            u, v, w, _ = ctx.multiscale(q, ims[0], dms[0], ims[1], dms[1], None, None)
            ub, vb, wb = results[(sid, 1)]
            print(f"stream {sid} {W}x{H}: first pair", flush=True)
            assert np.array_equal(u, ub) and np.array_equal(v, vb) and np.array_equal(w, wb), (sid, 1)
            qw = R.make_params(W, H, n_outer=4, n_sor=N_SOR, min_width=mw, relax_variant=R.RELAX_TILED,
                               storage=R.STORE_F16, image_storage=R.IMAGES_F16)
            est = np.stack(results[(sid, 1)], axis=-1)
            u, v, w, st = ctx.with_init(qw, ims[1], dms[1], ims[0], dms[0], est, None, None)
            ub, vb, wb = results[(sid, 2)]
            print(f"stream {sid} {W}x{H}: warm pair, {st.outer_exec[0]} outer, {st.sweeps_exec[0]} sweeps", flush=True)
            assert np.array_equal(u, ub) and np.array_equal(v, vb) and np.array_equal(w, wb), (sid, 2)
