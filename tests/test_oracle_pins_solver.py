"""Pins of the oracle's level-solve pieces that the GPU is bit-identical to
and that no other check fixes independently: the robust weights phi', psi',
pi' (sceneFlowSingleScaleJacobiCUDA.cu:340-371), the assembly's quirks
(q^2 on the It products only, .cu:846-847; the rhs Laplacian normalised but
the couplings not, .cu:274-282 vs :498-540; edge ownership and s_o vs s_p,
.cu:331,504-534), the relaxation's red-first colour order and old in-pixel
values (.cu:406,431-461,910-912), sceneFlowWithInit's single-level solve
(sceneFlowWithInit.cpp:7-81), the coarse-to-fine wiring with u, v scaled by
1/ratio and w not (sceneFlowMultiscale.cpp:95-201), and the depth-map
pyramid's 5 mm x stereoErrorRatio threshold (depthMapPyramid.cpp:66-93).

Each expectation is computed from the mathematics (Charbonnier derivatives
by finite differences, matrix algebra with an independently assembled sparse
operator, invariances under scalings) or from synthetic scenes with a known
answer, and each is chosen so that a plausible mistake -- a swapped epsilon,
a dropped q or delta, the wrong owner of an edge weight, the black colour
first, new values inside a pixel, w scaled on upsampling, a flat 5 mm
threshold -- fails it.
"""
import numpy as np
import scipy.sparse as sp

from helpers import make_case, oracle, synth
from test_oracle_pins import _sparse_system

O = oracle()
EPS_D, EPS_Z, EPS_S = 1e-6, 1e-10, 1e-6          # .cu:749


def _charbonnier_slope(x, eps):
    """d/dx sqrt(x + eps) by a central difference (relative step 1e-4)."""
    h = 1e-4 * (x + eps)
    return (np.sqrt(x + h + eps) - np.sqrt(x - h + eps)) / (2 * h)


def _fwd(a):
    """Forward differences with zero on the last column / row (.cu:156-182)."""
    gx = np.zeros_like(a)
    gy = np.zeros_like(a)
    gx[:, :-1] = a[:, 1:] - a[:, :-1]
    gy[:-1, :] = a[1:, :] - a[:-1, :]
    return gx, gy


def _weights_scene(rng, W=10, H=8):
    """Inputs whose residuals span the regimes of each penalty: colour and
    depth residuals far below and far above sqrt(eps), depth residuals past
    the +-0.04 clamp, flat flow patches (phi at its maximum), occluded and
    visible pixels, with both sides of the occlusion margin at two pixels."""
    yy, xx = np.mgrid[0:H, 0:W].astype(float)
    I1 = np.stack([np.sin(0.4 * xx + k) * np.cos(0.3 * yy - k) + 0.1 * k for k in range(5)], -1)
    It = rng.choice([1e-5, 3e-4, 2e-2, 1e-1], size=(H, W, 5)) * rng.choice([-1, 1], size=(H, W, 5))
    Iw = I1 + It
    D1 = 1.5 + 0.01 * xx - 0.02 * yy
    Dt = rng.choice([1e-7, 3e-6, 1e-2, 0.1], size=(H, W)) * rng.choice([-1, 1], size=(H, W))
    Dw = D1 + Dt
    u, v, w = (0.3 * rng.normal(size=(H, W)) for _ in range(3))
    for a in (u, v, w):
        a[2:5, 3:7] = 0.25                    # flat patch: phi = 1/2 / sqrt(eps_S) inside
    margin = D1 + w - 0.02 * D1 ** 2          # occluded (q = 0.01) iff D2 <= margin
    D2 = np.where(rng.random((H, W)) < 0.5, margin + 0.05, margin - 0.05)
    D2[0, 0] = margin[0, 0] + 1e-6            # just visible
    D2[0, 1] = margin[0, 1] - 1e-6            # just occluded
    q = np.where(D2 > margin, 1.0, 0.01)
    s = rng.uniform(0.5, 3.0, (H, W))
    delta = rng.uniform(0.5, 1.5, (H, W))
    m = rng.uniform(1, 2, (H, W))
    r = O.reg_weights(None, D1, O.REG_ISOTROPIC)
    li = O.LevelInputs(I1, I1, Iw, D1, D2, Dw, delta, s, m, r, np.float32(1.0))
    return li, (u, v, w), dict(It=It, Dt=np.clip(Dt, -0.04, 0.04), q=q, s=s, delta=delta)


def test_robust_weights_are_the_charbonnier_slopes():
    """phi, psi_c and pi equal d/d(rho^2) of sqrt(rho^2 + eps) at the
    residuals of the increment du = 0 (nInner = 1, du reset each outer
    iteration, .cu:776):
      phi:   rho^2 = ux^2+uy^2+vx^2+vy^2 + s (wx^2+wy^2), eps_S = 1e-6
      psi_c: rho = q It_c, eps_D = 1e-6, times the colour data weight
      pi:    rho = Dt (clamped to +-0.04), eps_Z = 1e-10, times delta.
    A swapped epsilon, a dropped q, s, delta or colour weight, a backward
    difference, or q from the wrong side of its margin changes them by far
    more than the tolerance."""
    rng = np.random.default_rng(31)
    li, (u, v, w), k = _weights_scene(rng)
    cw = 0.7
    ux, uy = _fwd(u)
    vx, vy = _fwd(v)
    wx, wy = _fwd(w)
    rho_s = ux ** 2 + uy ** 2 + vx ** 2 + vy ** 2 + k["s"] * (wx ** 2 + wy ** 2)
    phi_want = _charbonnier_slope(rho_s, EPS_S)
    psi_want = cw * _charbonnier_slope((k["q"][..., None] * k["It"]) ** 2, EPS_D)
    pi_want = k["delta"] * _charbonnier_slope(k["Dt"] ** 2, EPS_Z)
    assert (phi_want[2:4, 3:6] == _charbonnier_slope(0.0, EPS_S)).all()      # the flat patch
    assert k["q"][0, 0] == 1 and k["q"][0, 1] == 0.01
    for prec, tol in ((O.PREC_F64, 1e-7), (O.PREC_REF, 2e-6)):
        p = O.make_params(li.W, li.H, precision=prec, cw=cw)
        phi, psi, pi = O.level_weights(li, p, u, v, w)
        for got, want in ((phi, phi_want), (psi, psi_want), (pi, pi_want)):
            assert np.max(np.abs(got / want - 1)) < tol, np.max(np.abs(got / want - 1))


def _sum_edges(e_right, e_down):
    """For every pixel, the sum of the stored edge weights of its existing
    neighbours: its own right/down edges and its left/up neighbours'."""
    S = e_right + e_down
    S[:, 1:] += e_right[:, :-1]
    S[1:, :] += e_down[:-1, :]
    return S


def _system(li, alpha=0.5, gamma=0.0, cw=1.0, prec=None):
    p = O.make_params(li.W, li.H, alpha=alpha, gamma=gamma, cw=cw,
                      precision=O.PREC_F64 if prec is None else prec, reg=O.REG_HPB)
    return p


def _with(li, **kw):
    a = dict(I1=li.I1, I2=li.I2, warpI2=li.warpI2, D1=li.D1, D2=li.D2, warpD2=li.warpD2, delta=li.delta, s=li.s,
             m=li.m, r=li.r, focal_level=li.focal_level)
    a.update(kw)
    return O.LevelInputs(**a)


def test_assembly_normalises_the_rhs_laplacian_but_not_the_couplings():
    """A uniform regularisation weight c on every edge scales the couplings
    (and their sum on the diagonal) by c but leaves the rhs unchanged: the
    rhs Laplacians are divided by the mean weight of the existing edges
    (.cu:274-282), the coefficients of A are not (.cu:498-540)."""
    rng = np.random.default_rng(32)
    li, (u, v, w), _ = _weights_scene(rng)
    r1 = O.reg_weights(None, li.D1, O.REG_ISOTROPIC)
    c = 0.37
    A = O.level_system(_with(li, r=r1), _system(li, gamma=0.02), u, v, w)
    B = O.level_system(_with(li, r=c * r1), _system(li, gamma=0.02), u, v, w)
    assert np.abs(A[0:4]).max() > 1e-3 and np.abs(A[10:13]).max() > 1e-3
    assert np.allclose(B[0:4], c * A[0:4], rtol=1e-13, atol=0)                      # couplings x c
    assert np.allclose(B[10:13], A[10:13], rtol=1e-12, atol=1e-15)                  # rhs unchanged
    for d, (er, ed) in ((4, (0, 2)), (5, (0, 2)), (6, (1, 3))):                     # Buu, Bvv: xy; Bww: z
        S = _sum_edges(A[er], A[ed])
        assert np.allclose(B[d] - A[d], -(c - 1) * S, rtol=1e-9, atol=1e-14)
    assert np.array_equal(B[7:10], A[7:10])                                         # off-diagonal block


def test_q_squared_multiplies_only_the_it_products():
    """Occlusion q = 0.01 everywhere versus q = 1, with colour residuals so
    small that psi is the same in both (|q It|^2 << eps_D): the It products
    Itx, Ity (the rhs) shrink by q^2 = 1e-4, Ixx, Iyy, Ixy (the block) do
    not (.cu:842-847).  alpha = gamma = delta = 0 leaves only these terms."""
    rng = np.random.default_rng(33)
    li, (u, v, w), _ = _weights_scene(rng)
    It = 1e-7 * rng.uniform(0.5, 1.0, li.I1.shape) * rng.choice([-1, 1], li.I1.shape)
    base = dict(warpI2=li.I1 + It, delta=np.zeros((li.H, li.W)))
    margin = li.D1 + w - 0.02 * li.D1 ** 2
    vis = O.level_system(_with(li, D2=margin + 1.0, **base), _system(li, alpha=0.0), u, v, w)
    occ = O.level_system(_with(li, D2=margin - 1.0, **base), _system(li, alpha=0.0), u, v, w)
    assert np.abs(vis[4]).min() > 1e-3 and np.abs(vis[10]).min() > 1e-12
    for k in (4, 5, 7):                                          # -Ixx, -Iyy, -Ixy
        assert np.max(np.abs(occ[k] / vis[k] - 1)) < 1e-5
    for k in (10, 11):                                           # Itx, Ity
        assert np.max(np.abs(occ[k] / vis[k] / 1e-4 - 1)) < 1e-5
    assert not np.abs(occ[8:10]).any() and not np.abs(occ[12]).any()   # pi = 0: no depth terms


def test_edge_weights_belong_to_the_left_and_up_pixel():
    """The coupling of the edge between p and its right (down) neighbour is
    alpha phi_p r_p (edge owner p, .cu:504-534): eRxy = alpha phi r with the
    owner's phi, and the diagonal's smoothness part is the sum of the four
    stored couplings around the pixel -- so the left neighbour's coupling on
    the diagonal is the left pixel's phi r, not the centre's.  Checked on a
    scene with a step in the boundary maps and in the flow, where phi and r
    differ across the step."""
    rng = np.random.default_rng(34)
    li, (u, v, w), _ = _weights_scene(rng)
    hpb = np.zeros((li.H, li.W, 2), np.float32)
    hpb[:, 4, 0] = 0.9                                  # a boundary between columns 4 and 5
    hpb[3, :, 1] = 0.6                                  # and between rows 3 and 4
    r = O.reg_weights(hpb, li.D1, O.REG_HPB)
    u = u.copy()
    u[:, 5:] += 1.0                                     # the flow steps across it too
    li = _with(li, r=r)
    alpha = 0.5
    A = O.level_system(li, _system(li, alpha=alpha), u, v, w)
    A0 = O.level_system(li, _system(li, alpha=0.0), u, v, w)
    phi, _, _ = O.level_weights(li, _system(li, alpha=alpha), u, v, w)
    own_r = alpha * phi * r[..., 0]
    own_d = alpha * phi * r[..., 1]
    assert np.allclose(A[0][:, :-1], own_r[:, :-1], rtol=1e-14, atol=0)
    assert np.allclose(A[2][:-1, :], own_d[:-1, :], rtol=1e-14, atol=0)
    assert np.allclose(A[1][:, :-1], own_r[:, :-1] * li.s[:, :-1], rtol=1e-14, atol=0)   # z: the owner's s
    assert np.allclose(A[3][:-1, :], own_d[:-1, :] * li.s[:-1, :], rtol=1e-14, atol=0)
    assert np.ptp(phi[:, 4:6]) > 0.1 * phi[:, 4:6].max()      # phi really differs across the step
    for d, (er, ed) in ((4, (0, 2)), (5, (0, 2)), (6, (1, 3))):
        assert np.allclose(A[d] - A0[d], -_sum_edges(A[er], A[ed]), rtol=1e-10, atol=1e-14)


def test_w_laplacian_takes_the_centre_s_and_z_couplings_the_owners():
    """Doubling the depth-smoothness weight s at one pixel p0 whose flow has
    zero forward differences (so phi does not change anywhere): the w rhs
    there doubles (the w Laplacian is scaled by the centre's s, .cu:331), the
    z couplings p0 owns (right, down) double and those its left and up
    neighbours own do not (.cu:504-534), and the z diagonal changes by the
    couplings p0 owns only."""
    rng = np.random.default_rng(35)
    li, (u, v, w), _ = _weights_scene(rng)
    i0, j0 = 6, 7                             # outside the flat patch
    w = w.copy()
    w[i0, j0 + 1] = w[i0, j0]
    w[i0 + 1, j0] = w[i0, j0]
    s2 = li.s.copy()
    s2[i0, j0] *= 2
    zero = np.zeros((li.H, li.W))
    a = O.level_system(_with(li, delta=zero), _system(li), u, v, w)
    b = O.level_system(_with(li, delta=zero, s=s2), _system(li), u, v, w)
    assert abs(a[12][i0, j0]) > 1e-6
    assert np.isclose(b[12][i0, j0], 2 * a[12][i0, j0], rtol=1e-13, atol=0)
    assert b[1][i0, j0] == 2 * a[1][i0, j0] and b[3][i0, j0] == 2 * a[3][i0, j0]
    assert b[1][i0, j0 - 1] == a[1][i0, j0 - 1] and b[3][i0 - 1, j0] == a[3][i0 - 1, j0]
    assert np.array_equal(b[0], a[0]) and np.array_equal(b[2], a[2])        # xy couplings: phi unchanged
    assert np.isclose(b[6][i0, j0] - a[6][i0, j0], -(a[1][i0, j0] + a[3][i0, j0]), rtol=1e-12, atol=0)
    assert np.isclose(b[6][i0, j0 + 1] - a[6][i0, j0 + 1], -a[1][i0, j0], rtol=1e-12, atol=0)
    assert b[6][i0, j0 - 1] == a[6][i0, j0 - 1] and b[12][i0, j0 - 1] == a[12][i0, j0 - 1]


# ------------------------------------------------------------- relaxation order
def _level_system_f64(seed, W, H, scale=1e-2):
    from helpers import level_inputs
    li, (u, v, w), _ = level_inputs(seed=seed, W=W, H=H)
    coef = O.level_system(li, O.make_params(W, H, precision=O.PREC_F64), u, v, w)
    coef[10:13] *= scale                     # keep every step inside the 0.3 clamp
    return coef


def test_red_pixels_first_then_black_from_the_new_red():
    """One iteration from du = 0: the red pixels ((i + j) even, .cu:406) are
    relaxed first, from zero neighbours -- 0.7 rhs / B_kk -- and the black
    pixels then see the new red values: 0.7 (rhs - A x_red) / B_kk, with A
    assembled independently (scipy) from the couplings."""
    W, H = 10, 8
    coef = _level_system_f64(41, W, H)
    M, b = _sparse_system(coef)
    N = W * H
    D = M.diagonal()
    red = (np.add.outer(np.arange(H), np.arange(W)) % 2 == 0).ravel()
    red3 = np.tile(red, 3)
    om = float(np.float32(0.7))               # `const float omega = .7` (.cu:461)
    x_red = np.where(red3, om * b / D, 0.0)
    x_black = om * (b - M @ x_red) / D
    want = np.where(red3, x_red, x_black)
    got = O.relax(coef, 1, np.zeros((3, H, W)), precision=O.PREC_F64).reshape(-1)
    assert np.abs(want).max() < 0.3
    assert np.allclose(got, want, rtol=1e-12, atol=1e-18)
    # black first would give the black pixels 0.7 rhs / B_kk
    assert not np.allclose(got[~red3], (om * b / D)[~red3], rtol=1e-6)
    assert N == red.size


def _blocks(coef):
    Buu, Bvv, Bww, Buv, Buw, Bvw = (coef[k].reshape(-1) for k in (4, 5, 6, 7, 8, 9))
    B = np.stack([np.stack([Buu, Buv, Buw], -1), np.stack([Buv, Bvv, Bvw], -1), np.stack([Buw, Bvw, Bww], -1)], -2)
    return B, coef[10:13].reshape(3, -1).T


def test_pixel_update_uses_the_old_values_and_the_factor():
    """With the couplings removed every pixel is its own 3x3 system B x = r.
    One iteration of the reference's update (each unknown from the others'
    old values, factor 0.7, .cu:431-461) is the damped Jacobi step
    x + 0.7 D^-1 (r - B x); the point-SOR order (u, then v from the new u,
    then w) is the SOR step (D + w L)^-1 (w r - (w U + (w - 1) D) x) --
    both by matrix algebra on every pixel's block."""
    W, H = 8, 6
    coef = _level_system_f64(42, W, H, scale=0.1)
    coef[0:4] = 0.0
    rng = np.random.default_rng(42)
    x0 = rng.normal(scale=0.005, size=(3, H, W))
    B, r = _blocks(coef)
    X = x0.reshape(3, -1).T
    Dg = np.einsum("pkk->pk", B)
    f07 = float(np.float32(0.7))             # `const float omega = .7` (.cu:461)
    jac = X + f07 * (r - np.einsum("pkl,pl->pk", B, X)) / Dg
    assert np.abs(jac - X).max() < 0.3
    got = O.relax(coef, 1, x0, precision=O.PREC_F64).reshape(3, -1).T
    assert np.allclose(got, jac, rtol=1e-11, atol=1e-16)
    for om in (float(np.float32(0.7)), float(np.float32(1.8))):
        L = np.tril(B, -1)
        U = np.triu(B, 1)
        Dm = np.einsum("pk,kl->pkl", Dg, np.eye(3))
        rhs = om * r - np.einsum("pkl,pl->pk", om * U + (om - 1) * Dm, X)
        sor = np.linalg.solve(Dm + om * L, rhs[..., None])[..., 0]
        assert np.abs(sor - X).max() < 0.3
        gs, it, _, _ = O.relax_residual(coef, 1, 0.0, x0, omega=om, precision=O.PREC_F64, pixel_gs=True)
        assert it == 1 and np.allclose(gs.reshape(3, -1).T, sor, rtol=1e-11, atol=1e-16)
        jac_om = X + om * (r - np.einsum("pkl,pl->pk", B, X)) / Dg
        ref, it, _, _ = O.relax_residual(coef, 1, 0.0, x0, omega=om, precision=O.PREC_F64, pixel_gs=False)
        assert it == 1 and np.allclose(ref.reshape(3, -1).T, jac_om, rtol=1e-11, atol=1e-16)
    assert not np.allclose(got, sor, rtol=1e-6)              # the two orders really differ here


def test_block_update_solves_each_pixel_s_3x3_system():
    """The block update (pixel mode 2, not in the reference): with the
    couplings removed one iteration is x + w (B^-1 r - x) on every pixel, B^-1 r
    by numpy's LU solve -- in f64 to rounding, and in the f32 working
    precision (the inverse from the adjugate) to f32 accuracy."""
    W, H = 8, 6
    coef = _level_system_f64(43, W, H, scale=0.1)
    coef[0:4] = 0.0
    rng = np.random.default_rng(43)
    x0 = rng.normal(scale=0.005, size=(3, H, W))
    B, r = _blocks(coef)
    X = x0.reshape(3, -1).T
    xh = np.linalg.solve(B, r[..., None])[..., 0]
    for om in (1.0, float(np.float32(0.7)), float(np.float32(1.8))):
        want = om * xh + (1 - om) * X
        assert np.abs(want - X).max() < 0.3
        got, it, _, _ = O.relax_residual(coef, 1, 0.0, x0, omega=om, precision=O.PREC_F64, pixel_gs=2)
        assert it == 1 and np.allclose(got.reshape(3, -1).T, want, rtol=1e-10, atol=1e-15)
        g32, it, _, _ = O.relax_residual(coef, 1, 0.0, x0, omega=om, precision=O.PREC_REF, pixel_gs=2)
        assert it == 1 and np.abs(g32.reshape(3, -1).T - want).max() < 2e-5 * np.abs(want).max()
    # the point-SOR and reference orders stop short of the block's solution in one step
    gs, _, _, _ = O.relax_residual(coef, 1, 0.0, x0, omega=1.0, precision=O.PREC_F64, pixel_gs=1)
    assert not np.allclose(gs.reshape(3, -1).T, xh, rtol=1e-6)


def test_block_update_keeps_a_pixel_whose_block_is_not_definite():
    """A pixel whose stored block is not negative definite (here Buu made
    positive, or the block made singular) keeps its values in the block
    update (x^ = its old values, so with factor 1 it is unchanged); its
    neighbours still update from it."""
    W, H = 8, 6
    coef = _level_system_f64(45, W, H, scale=0.1)
    rng = np.random.default_rng(45)
    x0 = rng.normal(scale=0.005, size=(3, H, W))
    bad = (2, 3)
    c = coef.copy()
    c[4][bad] = abs(c[4][bad])                       # Buu > 0
    c2 = coef.copy()
    c2[5][bad] = c2[7][bad] ** 2 / c2[4][bad]        # Buu Bvv = Buv^2: the 2x2 minor vanishes
    for cc in (c, c2):
        for prec in (O.PREC_F64, O.PREC_REF):
            got, _, _, _ = O.relax_residual(cc, 1, 0.0, x0, omega=1.0, precision=prec, pixel_gs=2)
            xo = x0[:, bad[0], bad[1]].astype(np.float32 if prec == O.PREC_REF else np.float64)
            assert np.array_equal(got[:, bad[0], bad[1]], xo.astype(np.float64))
            ref, _, _, _ = O.relax_residual(coef, 1, 0.0, x0, omega=1.0, precision=prec, pixel_gs=2)
            assert not np.array_equal(ref[:, bad[0], bad[1]], got[:, bad[0], bad[1]])
            far = np.ones((H, W), bool)
            far[max(bad[0] - 2, 0):bad[0] + 3, max(bad[1] - 2, 0):bad[1] + 3] = False
            assert np.array_equal(got[:, far], ref[:, far])  # pixels two away see no difference


def test_block_update_reads_the_neighbours_red_then_black():
    """One block iteration with the couplings: red pixels solve
    B_p x_p = r_p - (E x0)_p (E: the couplings, assembled by scipy), black
    pixels the same from the new red values; each then relaxed by w."""
    W, H = 10, 8
    coef = _level_system_f64(44, W, H)
    M, b = _sparse_system(coef)
    N = W * H
    B, r = _blocks(coef)
    # E = M minus its pixel blocks
    Bd = sp.block_diag(list(B)).tocsr()
    perm = np.arange(3 * N).reshape(3, N).T.reshape(-1)        # (k, p) -> (p, k) order
    P = sp.csr_matrix((np.ones(3 * N), (np.arange(3 * N), perm)), shape=(3 * N, 3 * N))
    E = M - P.T @ Bd @ P
    rng = np.random.default_rng(44)
    x0 = rng.normal(scale=0.002, size=(3, H, W))
    red = (np.add.outer(np.arange(H), np.arange(W)) % 2 == 0).ravel()
    for om in (1.0, float(np.float32(1.8))):
        x = x0.reshape(-1).copy()
        for mask in (red, ~red):
            a = (b - E @ x).reshape(3, N).T
            xh = np.linalg.solve(B, a[..., None])[..., 0]
            xo = x.reshape(3, N).T
            xn = np.where(mask[:, None], om * xh + (1 - om) * xo, xo)
            assert np.abs(xn - xo).max() < 0.3
            x = xn.T.reshape(-1)
        got, it, _, _ = O.relax_residual(coef, 1, 0.0, x0, omega=om, precision=O.PREC_F64, pixel_gs=2)
        assert it == 1 and np.allclose(got.reshape(-1), x, rtol=1e-10, atol=1e-15)


def test_block_sor_converges_to_the_direct_solve():
    """Block SOR at w = 1.8 reaches M x = b (scipy's direct solve) in f64, and
    in f32 its converged energy x.b - x.Mx/2 is the direct solution's b.x*/2
    within 1e-4 -- the energy the reference's own update reaches too."""
    import scipy.sparse.linalg as spla
    from helpers import level_inputs
    W, H = 14, 12
    li, (u, v, w), _ = level_inputs(seed=5, W=W, H=H)
    coef = O.level_system(li, O.make_params(W, H, precision=O.PREC_F64), u, v, w)
    coef[10:13] *= 1e-2
    M, b = _sparse_system(coef)
    xs = spla.spsolve(M.tocsc(), b)
    e_star = 0.5 * b @ xs
    x, it, res, en = O.relax_residual(coef, 20000, 1e-10, np.zeros((3, H, W)), omega=float(np.float32(1.8)),
                                      precision=O.PREC_F64, pixel_gs=2)
    assert 0 < it < 20000 and res < 1e-10
    assert np.max(np.abs(x.reshape(-1) - xs)) < 1e-7 * np.abs(xs).max()
    assert abs(en - e_star) < 1e-9 * abs(e_star)
    x32, it32, res32, en32 = O.relax_residual(coef, 20000, 1e-4, np.zeros((3, H, W)), omega=float(np.float32(1.8)),
                                              precision=O.PREC_REF, pixel_gs=2)
    assert 0 < it32 < 20000 and res32 < 1e-4
    assert abs(en32 - e_star) < 1e-4 * abs(e_star)
    xr, itr, _, enr = O.relax_residual(coef, 20000, 1e-4, np.zeros((3, H, W)), precision=O.PREC_REF)
    assert abs(enr - e_star) < 1e-4 * abs(e_star)
    # block SOR at 1.8 needs fewer iterations than the reference's damped update
    assert it32 < itr


# ------------------------------------------------------------------ with_init
def _shifted_pair(W, H, dx, dy, seed):
    """Frame 2 is frame 1 moved by the integer (dx, dy): I2(x + dx, y + dy) =
    I1(x, y).  The colour is flat within 8 pixels of the border, so the
    5-tap derivatives (clamped at the border) of both frames' features are
    shifted copies of each other too."""
    pair, im1, d1, im2, d2, sigma = make_case(W, H, seed, 4)
    band = np.ones((H, W), bool)
    band[8:H - 8, 8:W - 8] = False
    im1 = im1.copy()
    im1[band] = 0.5
    rng = np.random.default_rng(seed)
    im2 = np.full(im1.shape, 0.5)
    d2 = rng.uniform(1, 3, d1.shape)
    im2[dy:, dx:] = im1[:H - dy, :W - dx]
    d2[dy:, dx:] = d1[:H - dy, :W - dx]
    return im1, d1, im2, d2, sigma


def test_with_init_keeps_an_exact_estimate_and_never_exits_early():
    """sceneFlowWithInit solves the full-resolution level only, from the
    estimate, with the outer early exit off (sceneFlowWithInit.cpp:73-80).
    An exact integer-translation estimate gives warped images equal to frame
    1 (It = Dt = 0) and a constant flow (zero Laplacian), so with gamma = 0
    the rhs is zero and the estimate comes back unchanged; the inner exit
    stops each relaxation at its first check, every outer iteration runs.
    Identical frames with a zero estimate give zero flow."""
    W, H, dx, dy = 48, 36, 2, 1
    im1, d1, im2, d2, sigma = _shifted_pair(W, H, dx, dy, 51)
    p = O.make_params(W, H, gamma=0.0, n_outer=4, n_sor=60, min_width=12, reg=O.REG_ISOTROPIC)
    est = np.zeros((H, W, 3))
    est[..., 0], est[..., 1] = dx, dy
    u, v, w, st = O.with_init(im1, d1, im2, d2, sigma, None, est, p)
    assert np.array_equal(u, est[..., 0]) and np.array_equal(v, est[..., 1]) and np.array_equal(w, est[..., 2])
    assert st.n_levels == 1 and st.width[0] == W and st.height[0] == H
    assert st.outer_exec[0] == 4 and st.sweeps_exec[0] == 4 * 25
    z = np.zeros((H, W, 3))
    u, v, w, st = O.with_init(im1, d1, im1, d1, sigma, None, z, O.make_params(W, H, n_outer=3, n_sor=30,
                                                                             min_width=12))
    assert not (u.any() or v.any() or w.any()) and st.outer_exec[0] == 3


def test_with_init_is_the_full_resolution_level_solve():
    """sceneFlowWithInit (sceneFlowWithInit.cpp:7-81) composed from the
    pinned stage functions: features of both frames, depth holes filled,
    the CPU warp by the estimate, delta from the stdevs, s = f^2, the
    magnitude and regularisation weights of the full-resolution depth, one
    level solve without the outer exit."""
    W, H = 48, 36
    pair, im1, d1, im2, d2, sigma = make_case(W, H, 52, 4)
    S = synth()
    hpb0 = S.segment_boundary_maps(pair["seg1"], [W], [H], 0.5).reshape(H, W, 2)
    gt = pair["flow"].astype(np.float64)
    est = 0.9 * gt
    d1h, d2h = d1.copy(), d2.copy()
    d1h[5:7, 9:12] = 0.0                               # holes the solve fills first
    p = O.make_params(W, H, n_outer=5, n_sor=40, min_width=12)
    u, v, w, st = O.with_init(im1, d1h, im2, d2h, sigma, hpb0, est, p)
    f = float(np.float32(p.cam.focalLength))
    D1, D2 = O.median_fill(d1h), O.median_fill(d2h)
    I1, I2 = O.im2feature(im1), O.im2feature(im2)
    Iw, Dw = O.warp_cpu(I1, D1, I2, D2, est[..., 0], est[..., 1], est[..., 2])
    li = O.LevelInputs(I1, I2, Iw, D1, D2, Dw, O.depth_data_weights(sigma), np.full((H, W), f * f),
                       O.depth_mag_weights(D1, W, H, f), O.reg_weights(hpb0, D1, O.REG_HPB), np.float32(f))
    u2, v2, w2, st2 = O.level_solve(li, p, False, est[..., 0], est[..., 1], est[..., 2])
    assert np.array_equal(u, u2) and np.array_equal(v, v2) and np.array_equal(w, w2)
    assert st.outer_exec[0] == st2.outer_exec == 5 and st.sweeps_exec[0] == st2.sweeps_exec


# ------------------------------------------------------- coarse-to-fine wiring
def test_multiscale_is_the_reference_level_loop():
    """sceneFlowMultiscale (sceneFlowMultiscale.cpp:30-201) composed from the
    pinned stage functions: the six Gaussian pyramids (colour x2, filled
    depth x2, delta, s), coarsest level from zero flow without a warp, each
    finer level entered with the flow resized to it, u and v multiplied by
    1/ratio and w not (:130-134), the CPU warp, m from the level's depth
    under the full-resolution camera (quirk A1), r from the level's map, and
    the outer exit on every level but the coarsest (:175)."""
    W, H = 64, 48
    pair, im1, d1, im2, d2, sigma = make_case(W, H, 53, 4)
    S = synth()
    ratio, mw = 0.5, 12
    lv = O.pyramid_levels(W, H, ratio, mw)
    hpb = S.segment_boundary_maps(pair["seg1"], [a for a, b in lv], [b for a, b in lv], ratio)
    p = O.make_params(W, H, n_outer=6, n_sor=30, min_width=mw)
    u, v, w, st = O.multiscale(im1, d1, im2, d2, sigma, hpb, p)
    f = float(np.float32(p.cam.focalLength))
    D1f, D2f = O.median_fill(d1), O.median_fill(d2)
    P1, P2 = O.gaussian_pyramid(im1, ratio, mw), O.gaussian_pyramid(im2, ratio, mw)
    PD1, PD2 = O.gaussian_pyramid(D1f, ratio, mw), O.gaussian_pyramid(D2f, ratio, mw)
    Pdel = O.gaussian_pyramid(O.depth_data_weights(sigma), ratio, mw)
    Ps = O.gaussian_pyramid(np.full((H, W), f * f), ratio, mw)
    offs = np.cumsum([0] + [a * b * 2 for a, b in lv])
    L = len(lv)
    cu = cv = cw_ = None
    for k in range(L - 1, -1, -1):
        Wk, Hk = lv[k]
        I1, I2 = O.im2feature(P1[k]), O.im2feature(P2[k])
        D1, D2 = PD1[k], PD2[k]
        if k == L - 1:
            cu = cv = cw_ = np.zeros((Hk, Wk))
            Iw, Dw = I2, D2
        else:
            cu = O.resize_dims(cu, Wk, Hk) * (1 / ratio)
            cv = O.resize_dims(cv, Wk, Hk) * (1 / ratio)
            cw_ = O.resize_dims(cw_, Wk, Hk)
            Iw, Dw = O.warp_cpu(I1, D1, I2, D2, cu, cv, cw_)
        fk = float(np.float32(np.float32(p.cam.focalLength) * (Wk / W)))
        hk = hpb[offs[k]:offs[k + 1]].reshape(Hk, Wk, 2)
        li = O.LevelInputs(I1, I2, Iw, D1, D2, Dw, Pdel[k], Ps[k], O.depth_mag_weights(D1, W, H, f),
                           O.reg_weights(hk, D1, O.REG_HPB), np.float32(fk))
        cu, cv, cw_, ls = O.level_solve(li, p, k < L - 1, cu, cv, cw_)
        assert (ls.outer_exec, ls.sweeps_exec) == (st.outer_exec[k], st.sweeps_exec[k])
    assert np.array_equal(u, cu) and np.array_equal(v, cv) and np.array_equal(w, cw_)


def test_fronto_parallel_z_motion_is_carried_across_levels_unscaled():
    """A textured plane facing the camera that moves 10 cm toward it, solved
    with alpha = 0 (each pixel on its own) and one outer iteration per level
    of a 4-level pyramid: the depth residual is clamped to +-0.04
    (.cu:768), so a level can move w by at most 4 cm, and the estimate
    reaches -0.1 m only by carrying each level's w to the next unchanged
    (sceneFlowMultiscale.cpp:130-134).  Scaling w by 1/ratio on each
    upsampling, as u and v are, would enter the two finest levels at -0.2
    and -0.32 m and end near -0.28 m."""
    W, H, z0, dz = 64, 48, 2.0, -0.1
    rng = np.random.default_rng(54)
    import scipy.ndimage as ndi
    tex = ndi.gaussian_filter(rng.uniform(0, 1, (3 * H, 3 * W, 3)), (2, 2, 0))
    f = 525.0 * W / 640
    cx, cy = (W - 1) / 2, (H - 1) / 2
    yy, xx = np.mgrid[0:H, 0:W].astype(float)

    def view(z):
        # a plane at depth z: pixel (x, y) sees texture point ((x - cx) z / f, (y - cy) z / f)
        X = (xx - cx) * z / f * 20 + 1.5 * W
        Y = (yy - cy) * z / f * 20 + 1.5 * H
        img = np.stack([ndi.map_coordinates(tex[..., c], [Y, X], order=1) for c in range(3)], -1)
        return img, np.full((H, W), z)

    im1, d1 = view(z0)
    im2, d2 = view(z0 + dz)
    sigma = O.depth_stdevs(d1.astype(np.float32), 4)
    p = O.make_params(W, H, alpha=0.0, n_outer=1, n_sor=50, min_width=8, reg=O.REG_ISOTROPIC)
    assert len(O.pyramid_levels(W, H, 0.5, 8)) == 4
    u, v, w, st = O.multiscale(im1, d1, im2, d2, sigma, None, p)
    inner = (slice(6, H - 6), slice(6, W - 6))
    assert abs(np.median(w[inner]) - dz) < 0.005, np.median(w[inner])


# ------------------------------------------------------- depth-map pyramid
def _blocks_level1(level0):
    return O.depth_map_pyramid(level0, 0.5, 2)[1]


def test_depth_map_pyramid_threshold_is_5mm_times_the_stereo_error_ratio():
    """Level 1 at ratio 0.5 samples each 2x2 block at its top-left pixel and
    averages it (in float) with the block's pixels within 5 mm x
    stereoErrorRatio(z) of it (depthMapPyramid.cpp:66-93): at 1 m a 4 mm
    step joins the average and a 6 mm step does not; at 2 m the threshold is
    5 mm x 3.989 = 19.9 mm, so 18 mm joins and 22 mm does not.  A centre
    with no neighbour within it takes the average of the opposite corner's
    component, else its own value; an invalid centre gives -1."""
    def block(c, right, down, diag):
        a = np.array([[c, right], [down, diag]], np.float64)
        return np.tile(a, (2, 2))

    def avg(*z):
        return float(np.float32(np.sum(np.float32(z), dtype=np.float32) / np.float32(len(z))))

    ratio2 = (4 / (1 / .00285 + 2)) / (1 / (1 / .00285 + 1))
    assert abs(0.005 * ratio2 - 0.01994) < 1e-5
    cases = [
        (block(1.0, 1.004, 1.0, 1.0), avg(1.0, 1.004, 1.0, 1.0)),
        (block(1.0, 1.006, 1.0, 1.0), avg(1.0, 1.0, 1.0)),
        (block(1.0, 1.0, 0.995, 1.0), avg(1.0, 1.0, 1.0)),           # |dz| = 5 mm: not strictly below
        (block(2.0, 2.018, 2.0, 2.0), avg(2.0, 2.018, 2.0, 2.0)),
        (block(2.0, 2.022, 2.0, 2.0), avg(2.0, 2.0, 2.0)),
        (block(1.0, 1.5, 1.5, 1.501), avg(1.501, 1.5, 1.5)),          # the opposite corner's component
        (block(1.0, 1.5, 1.7, 1.9), 1.0),                             # nobody near: the centre
        (block(-1.0, 1.0, 1.0, 1.0), -1.0),                           # invalid centre
    ]
    for lvl0, want in cases:
        got = _blocks_level1(lvl0)
        assert got.shape == (2, 2) and np.all(got == want), (lvl0[:2, :2], got, want)
