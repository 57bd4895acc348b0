"""Pins of the oracle's coarse-grid correction (an opt-in of this build, not
in the reference): the coarse operator, the restriction and the
prolongation against their definitions as sparse matrices (scipy), and the
V-cycle's effect on convergence.

The relaxed operator M of a compact system (13 planes: eRxy, eRz, eDxy, eDz,
Buu, Bvv, Bww, Buv, Buw, Bvw, ru, rv, rw) is (M x)_k(p) = sum over the
neighbours n of e_k(p, n) x_k(n) + sum_k' B_kk'(p) x_k'(p), e the xy weight
for u and v and the z weight for w.  P copies a coarse pixel's unknowns to
the (up to) 2x2 fine pixels of its aggregate.  The coarse operator is
P^T M P with its edges halved and its diagonal raised by half of P^T M P's
four edges around the pixel; the coarse rhs is P^T (r - M x)."""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from helpers import oracle


def _system(W, H, seed, data=1.0, ez=1.0, gamma=1e-5):
    """A random compact system of the solver's form: positive edge weights,
    negative definite blocks (a negative semidefinite data term, minus the
    edge weights around the pixel, minus gamma)."""
    rng = np.random.default_rng(seed)
    c = np.zeros((13, H, W))
    c[0, :, :-1] = rng.uniform(0.01, 1.0, (H, W - 1))
    c[1, :, :-1] = ez * rng.uniform(0.01, 1.0, (H, W - 1))
    c[2, :-1, :] = rng.uniform(0.01, 1.0, (H - 1, W))
    c[3, :-1, :] = ez * rng.uniform(0.01, 1.0, (H - 1, W))
    s_xy = np.zeros((H, W))
    s_z = np.zeros((H, W))
    for k, s in ((0, s_xy), (1, s_z)):
        s[:, :-1] += c[k, :, :-1]
        s[:, 1:] += c[k, :, :-1]
        s[:-1, :] += c[2 + k, :-1, :]
        s[1:, :] += c[2 + k, :-1, :]
    g = rng.normal(size=(H, W, 3, 2)) * np.sqrt(data)
    D = -np.einsum("hwim,hwjm->hwij", g, g)  # negative semidefinite, rank 2
    c[4] = D[..., 0, 0] - s_xy - gamma
    c[5] = D[..., 1, 1] - s_xy - gamma
    c[6] = D[..., 2, 2] - s_z - gamma
    c[7], c[8], c[9] = D[..., 0, 1], D[..., 0, 2], D[..., 1, 2]
    c[10:13] = rng.normal(size=(3, H, W))
    return c


def _matrix(c):
    """M as a sparse (3N x 3N) matrix, unknown (k, pixel) at k * N + pixel."""
    _, H, W = c.shape
    N = W * H
    rows, cols, vals = [], [], []
    idx = np.arange(N).reshape(H, W)
    B = [[4, 7, 8], [7, 5, 9], [8, 9, 6]]
    for k in range(3):
        for m in range(3):
            rows.append(k * N + idx.ravel()); cols.append(m * N + idx.ravel()); vals.append(c[B[k][m]].ravel())
        er, ed = (0, 2) if k < 2 else (1, 3)
        a, b = idx[:, :-1].ravel(), idx[:, 1:].ravel()
        e = c[er, :, :-1].ravel()
        rows += [k * N + a, k * N + b]; cols += [k * N + b, k * N + a]; vals += [e, e]
        a, b = idx[:-1, :].ravel(), idx[1:, :].ravel()
        e = c[ed, :-1, :].ravel()
        rows += [k * N + a, k * N + b]; cols += [k * N + b, k * N + a]; vals += [e, e]
    return sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(3 * N, 3 * N))


def _prolongation(W, H):
    Wc, Hc = (W + 1) // 2, (H + 1) // 2
    N, Nc = W * H, Wc * Hc
    i, j = np.divmod(np.arange(N), W)
    parent = (i // 2) * Wc + j // 2
    rows = np.concatenate([k * N + np.arange(N) for k in range(3)])
    cols = np.concatenate([k * Nc + parent for k in range(3)])
    return sp.csr_matrix((np.ones(3 * N), (rows, cols)), shape=(3 * N, 3 * Nc))


def _expected_operator(c):
    """The coarse planes from P^T M P (f64)."""
    _, H, W = c.shape
    Wc, Hc = (W + 1) // 2, (H + 1) // 2
    Nc = Wc * Hc
    G = (_prolongation(W, H).T @ _matrix(c) @ _prolongation(W, H)).toarray()
    idx = np.arange(Nc).reshape(Hc, Wc)
    out = np.zeros((13, Hc, Wc))
    gR = np.zeros((2, Hc, Wc))
    gD = np.zeros((2, Hc, Wc))
    for z, k in ((0, 0), (1, 2)):
        gR[z, :, :-1] = G[k * Nc + idx[:, :-1], k * Nc + idx[:, 1:]]
        gD[z, :-1, :] = G[k * Nc + idx[:-1, :], k * Nc + idx[1:, :]]
    out[0], out[1], out[2], out[3] = gR[0] / 2, gR[1] / 2, gD[0] / 2, gD[1] / 2
    around = np.zeros((2, Hc, Wc))
    around += gR
    around += gD
    around[:, :, 1:] += gR[:, :, :-1]
    around[:, 1:, :] += gD[:, :-1, :]
    for plane, (k, m) in zip(range(4, 10), [(0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2)]):
        out[plane] = G[k * Nc + idx, m * Nc + idx]
    out[4] += around[0] / 2
    out[5] += around[0] / 2
    out[6] += around[1] / 2
    return out, G


@pytest.mark.parametrize("W,H", [(8, 6), (9, 7), (5, 4), (12, 5)])
def test_coarse_operator_is_halved_galerkin(W, H):
    O = oracle()
    c = _system(W, H, seed=W * 31 + H)
    got = O.cg_operator(c, O.PREC_F64)
    want, G = _expected_operator(c)
    assert np.allclose(got, want, rtol=1e-12, atol=1e-12)
    # P^T M P of 2x2 aggregates is again a 5-point system with 3x3 blocks
    _, Hc, Wc = got.shape
    Nc = Wc * Hc
    mask = np.zeros_like(G, dtype=bool)
    idx = np.arange(Nc).reshape(Hc, Wc)
    for k in range(3):
        for m in range(3):
            mask[k * Nc + idx, m * Nc + idx] = True
        for a, b in ((idx[:, :-1], idx[:, 1:]), (idx[:-1, :], idx[1:, :])):
            mask[k * Nc + a, k * Nc + b] = mask[k * Nc + b, k * Nc + a] = True
    assert np.all(G[~mask] == 0)


def test_coarse_operator_keeps_data_term_and_laplacian():
    """A pure weighted Laplacian (no data, gamma 0) coarsens to a pure
    Laplacian whose edges are the means of the crossing fine edges (the
    weight a smooth field's coarse Laplacian has); a pure data term (no
    edges) coarsens to the sum of the children's blocks."""
    O = oracle()
    W, H = 10, 8
    c = _system(W, H, seed=5, data=0.0, gamma=0.0)
    co = O.cg_operator(c, O.PREC_F64)
    for k, er, ed in ((4, 0, 2), (6, 1, 3)):
        s = co[er].copy() + co[ed]
        s[:, 1:] += co[er][:, :-1]
        s[1:, :] += co[ed][:-1, :]
        assert np.allclose(co[k], -s, rtol=1e-12, atol=1e-12)
    assert np.allclose(co[0][:, :-1], (c[0][0::2, 1:-1:2] + c[0][1::2, 1:-1:2]) / 2)
    d = _system(W, H, seed=6)
    d[0:4] = 0
    d[4:10] = _system(W, H, seed=6, data=1.0)[4:10] - _system(W, H, seed=6, data=0.0)[4:10]
    cd = O.cg_operator(d, O.PREC_F64)
    for k in range(4, 10):
        want = d[k].reshape(H // 2, 2, W // 2, 2).sum(axis=(1, 3))
        assert np.allclose(cd[k], want, rtol=1e-12, atol=1e-12)
    assert np.all(cd[0:4] == 0)


@pytest.mark.parametrize("W,H", [(8, 6), (9, 7)])
def test_restrict_is_aggregated_residual(W, H):
    O = oracle()
    c = _system(W, H, seed=77)
    rng = np.random.default_rng(78)
    x = rng.normal(size=(3, H, W))
    got = O.cg_restrict(c, x, O.PREC_F64)
    N = W * H
    r = c[10:13].reshape(3 * N) - _matrix(c) @ x.reshape(3 * N)
    want = (_prolongation(W, H).T @ r).reshape(got.shape)
    assert np.allclose(got, want, rtol=1e-11, atol=1e-11)


@pytest.mark.parametrize("W,H", [(8, 6), (9, 7)])
def test_prolong_adds_coarse_unknowns(W, H):
    O = oracle()
    rng = np.random.default_rng(3)
    x = rng.normal(size=(3, H, W))
    xc = rng.normal(size=(3, (H + 1) // 2, (W + 1) // 2))
    got = O.cg_prolong(x, xc, O.PREC_F64)
    want = x.reshape(-1) + _prolongation(W, H) @ xc.reshape(-1)
    assert np.array_equal(got.reshape(-1), want)


def test_reference_precision_rounds_the_f64_operator():
    O = oracle()
    c = _system(9, 7, seed=12).astype(np.float32).astype(np.float64)
    a = O.cg_operator(c, O.PREC_REF)
    b = O.cg_operator(c, O.PREC_F64)
    assert np.all(a.astype(np.float32) == a)
    assert np.allclose(a, b, rtol=1e-5, atol=1e-6)


def _stiff(W=64, H=48, seed=9):
    """A system whose w rows are a stiff weighted Laplacian with a weak data
    term (as the depth-flow rows of the level systems are)."""
    return _system(W, H, seed, data=1e-3, ez=100.0)


def test_vcycle_speeds_up_convergence():
    O = oracle()
    c = _stiff()
    _, H, W = c.shape
    x0 = np.zeros((3, H, W))
    n = 50
    off = O.make_params(W, H, precision=O.PREC_F64)
    on = O.make_params(W, H, precision=O.PREC_F64, cg_levels=8, cg_sweeps=2, cg_at=9)
    _, _, r_off = O.relax_cg(c, n, x0, off)
    _, _, r_on = O.relax_cg(c, n, x0, on)
    assert r_on[-1] < r_off[-1] / 2
    # the V-cycle after iteration 9 is what changes the iteration
    assert np.array_equal(r_on[:9], r_off[:9]) and r_on[9] != r_off[9]


def test_solution_is_a_fixed_point_of_the_vcycle():
    O = oracle()
    c = _stiff(32, 24, seed=4)
    _, H, W = c.shape
    N = W * H
    xs = spla.spsolve(_matrix(c).tocsc(), c[10:13].reshape(3 * N)).reshape(3, H, W)
    on = O.make_params(W, H, precision=O.PREC_F64, cg_levels=8, cg_sweeps=2, cg_at=0)
    x, md, _ = O.relax_cg(c, 3, xs, on)
    assert np.max(np.abs(x - xs)) < 1e-9 * max(1.0, np.max(np.abs(xs)))


def test_planted_mistake_fails_the_galerkin_pin():
    """A coarse operator that keeps P^T M P's edges unhalved (a plausible
    slip) differs from the oracle's: the pin above would catch it."""
    O = oracle()
    c = _system(8, 6, seed=1)
    want, G = _expected_operator(c)
    wrong = want.copy()
    wrong[0:4] *= 2
    assert not np.allclose(O.cg_operator(c, O.PREC_F64), wrong, rtol=1e-6)
