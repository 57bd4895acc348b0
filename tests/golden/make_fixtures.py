"""Writes the small regression fixtures of the scene-flow path.

The reference (C++/CUDA with Boost, OpenCV, PCL and thrust) cannot be built
here and ships no tests, fixtures or golden vectors, so these fixtures come
from the oracle -- the CPU restatement whose functions the oracle pin tests
check against closed forms, scipy and synthetic ground truth.  They freeze
the oracle's and the GPU path's output on a few seeded inputs, so a change
of either shows up as a fixture mismatch.

Usage: python tests/golden/make_fixtures.py   (rewrites tests/golden/*.npz)
Inputs are regenerated from their seeds by the synthetic generator; the
files hold the generator's outputs too, so a generator change is caught.
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
from helpers import hpb_for, make_case, oracle  # noqa: E402

# (name, W, H, seed, segments, min_width, n_outer, n_sor, extra oracle params)
CASES = [
    ("pair_64x48_seed7", 64, 48, 7, 3, 16, 4, 30, {}),
    ("pair_64x48_seed9_sor", 64, 48, 9, 4, 16, 3, 40, {"omega": 1.8, "pixel_gs": True}),
    ("pair_48x36_seed5_residual", 48, 36, 5, 3, 12, 3, 100,
     {"omega": 1.8, "pixel_gs": True, "residual_exit": True, "residual_tol": 1e-3}),
]


def inputs(case):
    """The case's solver inputs, from its seed."""
    name, W, H, seed, segs, mw, n_outer, n_sor, extra = case
    pair, im1, d1, im2, d2, sigma = make_case(W, H, seed, segs)
    hpb, _ = hpb_for(pair, W, H, 0.5, mw)
    return pair, dict(im1=im1, dm1=d1, im2=im2, dm2=d2, sigma=sigma, hpb=hpb)


def from_fixture(z):
    """Solver inputs rebuilt from a fixture's stored frames, exactly as the
    synthetic pipeline converts them (u8 / 255; uint16 mm * .001 -> f32 -> f64)."""
    return dict(im1=z["img1"].astype(np.float64) / 255.0, im2=z["img2"].astype(np.float64) / 255.0,
                dm1=(z["depth1"].astype(np.float64) * .001).astype(np.float32).astype(np.float64),
                dm2=(z["depth2"].astype(np.float64) * .001).astype(np.float32).astype(np.float64),
                sigma=z["sigma"], hpb=z["hpb"].astype(np.float32))


def params(case, precision):
    O = oracle()
    name, W, H, seed, segs, mw, n_outer, n_sor, extra = case
    return O.make_params(W, H, n_outer=n_outer, n_sor=n_sor, min_width=mw, precision=precision, **extra)


def solve(case, x, precision):
    O = oracle()
    u, v, w, st = O.multiscale(x["im1"], x["dm1"], x["im2"], x["dm2"], x["sigma"], x["hpb"], params(case, precision))
    L = st.n_levels
    return dict(u=u, v=v, w=w, sweeps=np.array([st.sweeps_exec[k] for k in range(L)]),
                outer=np.array([st.outer_exec[k] for k in range(L)]))


def main():
    O = oracle()
    O.build()
    for case in CASES:
        pair, x = inputs(case)
        ref = solve(case, x, O.PREC_REF)
        f64 = solve(case, x, O.PREC_F64)
        out = os.path.join(HERE, case[0] + ".npz")
        np.savez_compressed(out, img1=pair["img1"], img2=pair["img2"], depth1=pair["depth1"], depth2=pair["depth2"],
                            sigma=x["sigma"], hpb=x["hpb"].astype(np.uint8), u=ref["u"], v=ref["v"], w=ref["w"],
                            sweeps=ref["sweeps"], outer=ref["outer"], u64=f64["u"], v64=f64["v"], w64=f64["w"])
        print(out, os.path.getsize(out), "bytes")


if __name__ == "__main__":
    main()
