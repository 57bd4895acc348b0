"""GPU parity of the batched solve: every pair of a batch, solved together
with the others (one launch per step for the whole batch, uneven sizes and
level counts allowed), equals the oracle's solve of that pair alone, bit for
bit, with the same per-level iteration counts."""
import numpy as np
import pytest

from helpers import hpb_for, load_package, make_case, oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def R():
    return load_package()


@pytest.fixture(scope="module")
def ctx(R):
    c = R.Context(0)
    yield c
    c.close()


def _case(W, H, seed, min_width, segments=4):
    pair, im1, d1, im2, d2, sigma = make_case(W, H, seed, segments)
    hpb, lv = hpb_for(pair, W, H, 0.5, min_width)
    return dict(W=W, H=H, im1=im1, dm1=d1, im2=im2, dm2=d2, sigma1=sigma, hpb=hpb, min_width=min_width)


def _oracle_solve(c, n_outer, n_sor, exits=True):
    O = oracle()
    po = O.make_params(c["W"], c["H"], n_outer=n_outer, n_sor=n_sor, min_width=c["min_width"],
                       disable_inner_exit=not exits, disable_outer_exit=not exits)
    return O.multiscale(c["im1"], c["dm1"], c["im2"], c["dm2"], c["sigma1"], c["hpb"], po)


def _check_batch(ctx, R, cases, relax_variant, n_outer=4, n_sor=30, exits=True, drop_sigma=()):
    W0, H0 = cases[0]["W"], cases[0]["H"]
    p = R.make_params(W0, H0, n_outer=n_outer, n_sor=n_sor, min_width=cases[0]["min_width"],
                      relax_variant=relax_variant, inner_exit=exits, outer_exit=exits)
    pairs = []
    for i, c in enumerate(cases):
        q = dict(c)
        q["cam"] = R.make_params(c["W"], c["H"]).cam
        if i in drop_sigma:
            q["sigma1"] = None   # computed on the device
        pairs.append(q)
    ctx.batch_load(p, pairs)
    stats = ctx.batch_run(p, stats=True)
    assert len(stats) == len(cases)
    for i, c in enumerate(cases):
        print(f"pair {i}: {c['W']}x{c['H']} min_width {c['min_width']}", flush=True)
        ua, va, wa, sa = _oracle_solve(c, n_outer, n_sor, exits)
        ub, vb, wb = ctx.batch_read(i)
        sb = stats[i]
        L = sa.n_levels
        assert sb.n_levels == L
        assert [sb.outer_exec[k] for k in range(L)] == [sa.outer_exec[k] for k in range(L)]
        assert [sb.sweeps_exec[k] for k in range(L)] == [sa.sweeps_exec[k] for k in range(L)]
        assert np.array_equal(ua, ub) and np.array_equal(va, vb) and np.array_equal(wa, wb), i
    return p


@pytest.mark.parametrize("variant", ["plain", "tiled"])
def test_batch_same_size(ctx, R, variant):
    cases = [_case(160, 120, seed, 20) for seed in (31, 32, 33)]
    _check_batch(ctx, R, cases, R.RELAX_TILED if variant == "tiled" else R.RELAX_PLAIN)


@pytest.mark.parametrize("variant", ["plain", "tiled"])
def test_batch_uneven_sizes(ctx, R, variant):
    """Different sizes and level counts in one batch (pyramids aligned on the
    finest level; the largest item of a step fixes the tiled launch shape)."""
    cases = [_case(200, 150, 41, 20), _case(160, 120, 42, 20), _case(96, 72, 43, 20), _case(128, 96, 44, 40),
             _case(64, 48, 45, 12)]
    _check_batch(ctx, R, cases, R.RELAX_TILED if variant == "tiled" else R.RELAX_PLAIN)


def test_batch_no_exits_and_device_sigma(ctx, R):
    cases = [_case(128, 96, 51, 16), _case(128, 96, 52, 16)]
    O = oracle()
    # the device stdevs equal the oracle's to 1 f32 ulp; the oracle's are used
    # for pair 0, the device's for pair 1, whose solve is checked at tolerance
    W, H = 128, 96
    p = R.make_params(W, H, n_outer=3, n_sor=30, min_width=16, relax_variant=R.RELAX_TILED, inner_exit=False,
                      outer_exit=False)
    pairs = [dict(cases[0]), dict(cases[1], sigma1=None)]
    ctx.batch_load(p, pairs)
    st = ctx.batch_run(p, stats=True)
    ua, va, wa, sa = _oracle_solve(cases[0], 3, 30, exits=False)
    ub, vb, wb = ctx.batch_read(0)
    assert np.array_equal(ua, ub) and np.array_equal(va, vb) and np.array_equal(wa, wb)
    ua, va, wa, sa = _oracle_solve(cases[1], 3, 30, exits=False)
    ub, vb, wb = ctx.batch_read(1)
    assert np.max(np.abs(ua - ub)) <= 1e-3 and np.max(np.abs(va - vb)) <= 1e-3 and np.max(np.abs(wa - wb)) <= 1e-5
    assert st[1].sweeps_exec[0] == 3 * 30


def test_batch_rerun_and_new_params(ctx, R):
    """Descriptors are reused across runs and rebuilt when the params change."""
    cases = [_case(160, 120, 61, 20), _case(160, 120, 62, 20)]
    p = _check_batch(ctx, R, cases, R.RELAX_TILED, n_outer=3, n_sor=25)
    ctx.batch_run(p, stats=False)
    ctx.synchronize()
    ua, va, wa, _ = _oracle_solve(cases[1], 3, 25)
    ub, vb, wb = ctx.batch_read(1)
    assert np.array_equal(ua, ub) and np.array_equal(va, vb) and np.array_equal(wa, wb)
    # new iteration counts on the loaded batch
    p.n_outer, p.n_sor = 5, 40
    st = ctx.batch_run(p, stats=True)
    for i, c in enumerate(cases):
        ua, va, wa, sa = _oracle_solve(c, 5, 40)
        ub, vb, wb = ctx.batch_read(i)
        assert [st[i].sweeps_exec[k] for k in range(sa.n_levels)] == [sa.sweeps_exec[k] for k in range(sa.n_levels)]
        assert np.array_equal(ua, ub) and np.array_equal(va, vb) and np.array_equal(wa, wb)


def test_batch_autotuned(R):
    """The autotuned default schedule, timed on batches, leaves results unchanged."""
    cases = [_case(200, 150, 71, 25), _case(200, 150, 72, 25), _case(200, 150, 73, 25)]
    O = oracle()
    with R.Context(0) as c:
        for (w, h) in O.pyramid_levels(200, 150, 0.5, 25):
            ch = c.autotune_relax(w, h, batch=3)
            assert ch["candidates"] > 0 and ch["us_per_iter"] > 0
        _check_batch(c, R, cases, R.RELAX_TILED)


def test_batch_errors(R):
    with R.Context(0) as c:
        cases = [_case(96, 72, 81, 12), _case(96, 72, 82, 12)]
        p = R.make_params(96, 72, n_outer=2, n_sor=25, min_width=12)
        with pytest.raises(R.RgbdFlowError) as e:
            c.batch_run(p)       # nothing loaded
        assert e.value.code == R.ERR_STATE
        c.batch_load(p, cases)
        with pytest.raises(R.RgbdFlowError) as e:
            c.pair_run(p)        # a batch of two is loaded
        assert e.value.code == R.ERR_STATE
        c.batch_run(p, stats=False)
        with pytest.raises(R.RgbdFlowError) as e:
            c._batch_shapes.append((72, 96))
            c.batch_read(2)
        assert e.value.code == R.ERR_ARG
        c._batch_shapes.pop()
        with pytest.raises(R.RgbdFlowError) as e:
            # maps made on the device for a pair loaded without them, but not for a
            # pair loaded under other params
            c.batch_load(R.make_params(96, 72, n_outer=2, n_sor=25, min_width=12, reg=R.REG_ISOTROPIC),
                         [dict(cases[0], hpb=None)])
            c.batch_run(p)
        assert e.value.code == R.ERR_ARG


@pytest.mark.parametrize("variant", ["plain", "tiled"])
def test_batch_graph_replays(ctx, R, variant):
    """params.graph: the second run captures the schedule as a HIP graph and
    later runs replay it; every replay (with other device work and
    synchronisations in between) equals the oracle bit for bit."""
    O = oracle()
    cases = [_case(160, 120, 91, 20), _case(96, 72, 92, 12)]
    p = R.make_params(160, 120, n_outer=4, n_sor=30, min_width=20, graph=True,
                      relax_variant=R.RELAX_TILED if variant == "tiled" else R.RELAX_PLAIN)
    pairs = [dict(c, cam=R.make_params(c["W"], c["H"]).cam) for c in cases]
    ctx.batch_load(p, pairs)
    ref = [_oracle_solve(c, 4, 30) for c in cases]
    coef = O.level_system(*_small_system(O))
    for run in range(4):
        st = ctx.batch_run(p, stats=True)
        for i, c in enumerate(cases):
            ua, va, wa, sa = ref[i]
            ub, vb, wb = ctx.batch_read(i)
            assert [st[i].sweeps_exec[k] for k in range(sa.n_levels)] == \
                [sa.sweeps_exec[k] for k in range(sa.n_levels)], (run, i)
            assert np.array_equal(ua, ub) and np.array_equal(va, vb) and np.array_equal(wa, wb), (run, i)
        # eager work on the same stream between replays
        ctx.relax(coef, 7, np.zeros((3, 24, 32)))
        ctx.synchronize()


def _small_system(O):
    from helpers import level_inputs
    li, (u, v, w), _ = level_inputs(seed=5, W=32, H=24)
    return li, O.make_params(32, 24), u, v, w


@pytest.mark.parametrize("relax_variant", [0, 1])
def test_batch_with_init_pairs(ctx, R, relax_variant):
    """Pairs given an estimate are solved as sceneFlowWithInit (finest level
    only, from the estimate, no outer early exit) inside a batch that also
    holds ordinary coarse-to-fine pairs of other sizes: each equals the
    oracle's with_init / multiscale of that pair alone, bit for bit."""
    O = oracle()
    a = _case(128, 96, 41, 16)
    b = _case(160, 120, 42, 20)
    c = _case(96, 72, 43, 12)
    for k, case in enumerate((a, c)):
        pair, *_ = make_case(case["W"], case["H"], 41 if k == 0 else 43, 4)
        case["est"] = pair["flow"].astype(np.float64) * (0.9 if k == 0 else 0.5)
    for case in (a, b, c):
        case["cam"] = R.make_params(case["W"], case["H"]).cam
    p = R.make_params(128, 96, n_outer=3, n_sor=40, min_width=16, relax_variant=relax_variant)
    ctx.batch_load(p, [a, b, c])
    sts = ctx.batch_run(p, stats=True)
    for i, case in enumerate((a, b, c)):
        ub, vb, wb = ctx.batch_read(i)
        po = O.make_params(case["W"], case["H"], n_outer=3, n_sor=40, min_width=case["min_width"])
        if "est" in case:
            hpb0 = case["hpb"][: case["W"] * case["H"] * 2]
            ua, va, wa, sa = O.with_init(case["im1"], case["dm1"], case["im2"], case["dm2"], case["sigma1"], hpb0,
                                         case["est"], po)
            assert sts[i].n_levels == 1 and sts[i].outer_exec[0] == 3
            # the same as the single-pair entry point, bit for bit
            pw = R.make_params(case["W"], case["H"], n_outer=3, n_sor=40, relax_variant=relax_variant)
            with R.Context(0) as c2:
                uc, vc, wc, sc = c2.with_init(pw, case["im1"], case["dm1"], case["im2"], case["dm2"], case["est"],
                                              case["sigma1"], hpb0)
            assert np.array_equal(uc, ub) and np.array_equal(vc, vb) and np.array_equal(wc, wb)
            assert sc.sweeps_exec[0] == sts[i].sweeps_exec[0]
        else:
            ua, va, wa, sa = O.multiscale(case["im1"], case["dm1"], case["im2"], case["dm2"], case["sigma1"],
                                          case["hpb"], po)
        L = sa.n_levels
        assert [sts[i].sweeps_exec[k] for k in range(L)] == [sa.sweeps_exec[k] for k in range(L)]
        du, dv, dw = (float(np.abs(x - y).max()) for x, y in ((ua, ub), (va, vb), (wa, wb)))
        assert du <= 1e-3 and dv <= 1e-3 and dw <= 1e-5, (i, du, dv, dw)
        if "est" not in case:
            assert np.array_equal(ua, ub) and np.array_equal(va, vb) and np.array_equal(wa, wb)


@pytest.mark.parametrize("relax_variant", [0, 1])
def test_batch_boundary_maps_on_device(ctx, R, relax_variant):
    """Pairs loaded without boundary maps under HPB regularisation get them
    made on the device from their colour and depth-map pyramids
    (sceneFlowMultiscale.cpp:116): each pair equals the oracle run the same
    way, bit for bit; a with-init pair gets its level-0 map the same way."""
    O = oracle()
    a = _case(128, 96, 51, 16)
    b = _case(160, 120, 52, 20)
    c = _case(96, 72, 53, 12)
    pair, *_ = make_case(96, 72, 53, 4)
    c["est"] = pair["flow"].astype(np.float64) * 0.7
    for case in (a, b, c):
        case["cam"] = R.make_params(case["W"], case["H"]).cam
        case["hpb"] = None
    p = R.make_params(128, 96, n_outer=3, n_sor=30, min_width=16, relax_variant=relax_variant, graph=True)
    print(f"device maps, relax variant {relax_variant}", flush=True)
    ctx.batch_load(p, [a, b, c])
    for run in range(3):    # eager, capture, replay
        sts = ctx.batch_run(p, stats=True)
        for i, case in enumerate((a, b, c)):
            ub, vb, wb = ctx.batch_read(i)
            po = O.make_params(case["W"], case["H"], n_outer=3, n_sor=30, min_width=case["min_width"])
            if "est" in case:
                ua, va, wa, sa = O.with_init(case["im1"], case["dm1"], case["im2"], case["dm2"], case["sigma1"], None,
                                             case["est"], po)
                with R.Context(0) as c2:
                    pw = R.make_params(case["W"], case["H"], n_outer=3, n_sor=30, relax_variant=relax_variant)
                    uc, vc, wc, _ = c2.with_init(pw, case["im1"], case["dm1"], case["im2"], case["dm2"], case["est"],
                                                 case["sigma1"], None)
                assert np.array_equal(uc, ub) and np.array_equal(vc, vb) and np.array_equal(wc, wb)
                du, dv, dw = (float(np.abs(x - y).max()) for x, y in ((ua, ub), (va, vb), (wa, wb)))
                assert du <= 1e-3 and dv <= 1e-3 and dw <= 1e-5, (run, i, du, dv, dw)
            else:
                ua, va, wa, sa = O.multiscale(case["im1"], case["dm1"], case["im2"], case["dm2"], case["sigma1"],
                                              None, po)
                L = sa.n_levels
                assert [sts[i].sweeps_exec[k] for k in range(L)] == [sa.sweeps_exec[k] for k in range(L)]
                assert np.array_equal(ua, ub) and np.array_equal(va, vb) and np.array_equal(wa, wb), (run, i)


def test_batch_boundary_maps_two_groups(ctx, R):
    """More pair levels than one launch of the boundary-map kernels takes
    (24): 7 pairs x 4 levels run as two groups one after the other, sharing
    their scratch; every pair equals the oracle run with maps made the same
    way, bit for bit."""
    O = oracle()
    cases = [_case(64 + 8 * (i % 3), 48 + 6 * (i % 3), 60 + i, 8) for i in range(7)]
    for case in cases:
        case["cam"] = R.make_params(case["W"], case["H"]).cam
        case["hpb"] = None
    p = R.make_params(64, 48, n_outer=3, n_sor=20, min_width=8, relax_variant=1, graph=False)
    ctx.batch_load(p, cases)
    sts = ctx.batch_run(p, stats=True)
    for i, case in enumerate(cases):
        print(f"pair {i}: {case['W']}x{case['H']}", flush=True)
        po = O.make_params(case["W"], case["H"], n_outer=3, n_sor=20, min_width=8)
        ua, va, wa, sa = O.multiscale(case["im1"], case["dm1"], case["im2"], case["dm2"], case["sigma1"], None, po)
        assert sa.n_levels == 4
        ub, vb, wb = ctx.batch_read(i)
        assert [sts[i].sweeps_exec[k] for k in range(4)] == [sa.sweeps_exec[k] for k in range(4)]
        assert np.array_equal(ua, ub) and np.array_equal(va, vb) and np.array_equal(wa, wb), i


@pytest.mark.parametrize("schedule", ["streamed", "planned"])
def test_batch_sizes_in_steps_of_their_own(ctx, R, schedule):
    """Two or more pairs of one size relax in a step of their own at each
    level (under a streamed schedule, or the planned one), a size of one pair
    shares a step with the other such sizes; with params.graph the steps are
    captured and replayed.  Every pair equals the oracle's solve of it alone,
    every run."""
    cases = [_case(160, 120, 61 + i, 20) for i in range(4)] + [_case(96, 72, 71, 12), _case(200, 150, 72, 25),
                                                               _case(96, 72, 73, 12)]
    kw = {}
    if schedule == "streamed":
        kw["tile_config"] = max((sh["streamed"], i) for i, sh in enumerate(R.relax_shapes()) if sh["streamed"] > 0)[1]
    p = R.make_params(160, 120, n_outer=3, n_sor=30, min_width=20, graph=True, relax_variant=R.RELAX_TILED, **kw)
    pairs = [dict(c, cam=R.make_params(c["W"], c["H"]).cam) for c in cases]
    ctx.batch_load(p, pairs)
    ref = [_oracle_solve(c, 3, 30) for c in cases]
    for run in range(3):
        print(f"{schedule} run {run}", flush=True)
        st = ctx.batch_run(p, stats=True)
        for i, c in enumerate(cases):
            ua, va, wa, sa = ref[i]
            ub, vb, wb = ctx.batch_read(i)
            L = sa.n_levels
            assert [st[i].outer_exec[k] for k in range(L)] == [sa.outer_exec[k] for k in range(L)], (run, i)
            assert [st[i].sweeps_exec[k] for k in range(L)] == [sa.sweeps_exec[k] for k in range(L)], (run, i)
            assert np.array_equal(ua, ub) and np.array_equal(va, vb) and np.array_equal(wa, wb), (run, i)
