"""CPU checks of the C ABI: the library loads without a GPU, exports every
function include/rgbdflow/rgbdflow.h declares, and its pure host functions
agree with the oracle."""
import ctypes as C
import os
import re

import pytest

from helpers import ROOT, load_package, oracle

HEADER = os.path.join(ROOT, "include", "rgbdflow", "rgbdflow.h")


def declared_functions(path):
    text = open(path).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(rgbdflow_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_binding_list():
    R = load_package()
    assert declared_functions(HEADER) == sorted(R.EXPORTS)


def test_library_exports_every_declared_symbol():
    R = load_package()
    lib = C.CDLL(R.LIB_PATH)
    missing = [name for name in declared_functions(HEADER) if not hasattr(lib, name)]
    assert not missing, missing


def test_default_params_match_reference_constructor():
    """OpticalFlow::sceneFlowParams() defaults (OpticalFlow.h:107-108)."""
    R = load_package()
    p = R.Params()
    assert R.lib().rgbdflow_default_params(C.byref(p)) == 0
    assert (p.variant, p.regularization_type) == (R.VARIANT_FULL, R.REG_HPB)
    assert (p.color_data_weight, p.depth_vs_color_data_weight, p.alpha, p.gamma) == (1, 1, 0, 0)
    assert (p.pyramid_ratio, p.min_width, p.n_outer, p.n_inner, p.n_sor) == (.75, 20, 7, 1, 30)


@pytest.mark.parametrize("W,H,r,mw", [(640, 480, .5, 40), (960, 540, .5, 60), (1280, 720, .5, 80),
                                      (320, 240, .75, 20), (101, 77, .6, 9), (640, 480, .5, 60)])
def test_pyramid_levels_match_oracle(W, H, r, mw):
    R = load_package()
    assert R.pyramid_levels(W, H, r, mw) == oracle().pyramid_levels(W, H, r, mw)


def test_null_and_bad_arguments_without_gpu():
    R = load_package()
    lib = R.lib()
    assert lib.rgbdflow_default_params(None) == R.ERR_ARG
    assert lib.rgbdflow_destroy(None) == R.ERR_ARG
    assert lib.rgbdflow_pyramid_levels(1, 1, C.c_double(.5), 10, None, None) == R.ERR_ARG
    assert lib.rgbdflow_last_error(None) == b"null context"
