"""GPU tests of the cubic B-spline resampling (ResizeProcessor, BSpline3) and
the downsampled TGV-k deshade (tgvk_deshade_downsampled/): the resampling
kernels perform the oracle's FP64 operations in the same order, so their
outputs are compared bit for bit; the deshades through the oracle's
compositions."""
import os
import sys

import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu

TAU = float(np.float32(1 / np.sqrt(20)))
KPRM = dict(lam=0.7, tau=TAU, sigma=TAU, theta=1.0)


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _f(shape, seed):
    from ambient_image_processor_amd import phantoms
    return phantoms.deshade_phantom(shape, seed=seed)[0]


def _mask(shape, seed):
    rng = np.random.default_rng(seed)
    d, h, w = shape
    zz, yy, xx = np.mgrid[0:d, 0:h, 0:w]
    m = ((zz - d / 2) ** 2 / (0.6 * d + 1) ** 2 + (yy - h / 2) ** 2 / (0.45 * h) ** 2
         + (xx - w / 2) ** 2 / (0.45 * w) ** 2) < 1
    return np.where(m, 1.0, rng.random(shape) * 0.5).astype(np.float32)


RESIZES = [((9, 30, 41), (5, 15, 21)), ((5, 15, 21), (9, 30, 41)), ((8, 25, 33), (3, 8, 10)),
           ((6, 7, 64), (2, 2, 16)), ((2, 2, 16), (6, 7, 64)), ((1, 40, 52), (1, 20, 26)),
           ((1, 20, 26), (1, 40, 52)), ((1, 1, 3), (1, 1, 30)), ((40, 48, 64), (20, 24, 32)),
           ((20, 24, 32), (40, 48, 64)), ((7, 9, 11), (7, 9, 11)),
           # x lines across the 32-sample chunks of the x prefilter: a last chunk of 1
           # sample, 3 chunks + 1, the 18-term start at 19 samples, the exact start
           # at 18 and at 2
           ((4, 5, 33), (2, 3, 17)), ((2, 3, 97), (5, 7, 200)), ((2, 70, 19), (1, 35, 10)),
           ((1, 3, 18), (1, 6, 36)), ((3, 4, 2), (6, 8, 4)), ((3, 66, 130), (2, 33, 65))]


@pytest.mark.parametrize("src,dst", RESIZES)
def test_resize_bitwise_against_oracle(ctx, orc, src, dst):
    _say("resize", src, "->", dst)
    a = (np.random.default_rng(3).standard_normal(src) + 1.0).astype(np.float32)
    got = ctx.resize_bspline3(a, dst)
    np.testing.assert_array_equal(got, orc.resize_spline3(a, dst))


def test_resize_extents_match(orc):
    from ambient_image_processor_amd import native
    for shape, factor in (((256, 256, 256), 0.5), ((1, 100, 101), 0.3), ((3, 10, 10), 0.01),
                          ((17, 33, 65), 0.25)):
        assert native.resize_extents(shape, factor) == orc.resize_extents(shape, factor)


@pytest.mark.parametrize("factor,order,with_mask", [(0.5, 3, True), (0.5, 3, False),
                                                     (0.6, 4, True), (1.0, 3, True),
                                                     (0.5, 2, True)])
def test_downsampled_masked_processor(ctx, orc, factor, order, with_mask):
    """TGVKDeshadeDownsampledProcessor (the widget's and CLI's form)."""
    from ambient_image_processor_amd import native
    _say("downsampled masked", factor, order, with_mask)
    shape = (12, 20, 28)
    f = _f(shape, 4)
    m = _mask(shape, 5) if with_mask else None
    alpha = tuple(float(i + 1) for i in range(order))
    p = native.default_params(lam=KPRM["lam"], tau=TAU, sigma=TAU, theta=1.0, iterations=30,
                              flags=native.FLAG_MASKED_ITERATION | native.FLAG_SET_NEGATIVE_ZERO
                              | native.FLAG_ADD_BACKGROUND_BACK)
    u, l, r = ctx.tgvk_deshade_downsampled(f, p, factor, alpha, m)
    uo, lo, ro = orc.deshade_k_downsampled(f, factor, 30, alpha=alpha, mask=m,
                                           add_background_back=True, masked=True, **KPRM)
    np.testing.assert_array_equal(u, uo)
    assert rel_err(l, lo) < 1e-5 and rel_err(r, ro) < 1e-5
    assert np.all(r >= 0)


@pytest.mark.parametrize("with_mask", [False, True])
def test_downsampled_not_masked_processor(ctx, orc, with_mask):
    from ambient_image_processor_amd import native
    shape = (10, 22, 30)
    f = _f(shape, 6)
    m = _mask(shape, 7) if with_mask else None
    alpha = (1.0, 2.0, 3.0)
    p = native.default_params(lam=KPRM["lam"], tau=TAU, sigma=TAU, theta=1.0, iterations=30,
                              flags=native.FLAG_SET_NEGATIVE_ZERO | native.FLAG_ADD_BACKGROUND_BACK)
    u, l, r = ctx.tgvk_deshade_downsampled(f, p, 0.5, alpha, m)
    uo, lo, ro = orc.deshade_k_downsampled(f, 0.5, 30, alpha=alpha, mask=m,
                                           add_background_back=True, masked=False, **KPRM)
    np.testing.assert_array_equal(u, uo)
    assert rel_err(l, lo) < 1e-5 and rel_err(r, ro) < 1e-5


def test_downsampled_depth_1(ctx, orc):
    from ambient_image_processor_amd import native
    shape = (1, 40, 52)
    f = _f(shape, 8)
    m = _mask(shape, 9)
    p = native.default_params(lam=KPRM["lam"], tau=TAU, sigma=TAU, theta=1.0, iterations=30,
                              flags=native.FLAG_MASKED_ITERATION | native.FLAG_SET_NEGATIVE_ZERO)
    u, l, r = ctx.tgvk_deshade_downsampled(f, p, 0.5, (1.0, 2.0, 3.0), m)
    uo, lo, ro = orc.deshade_k_downsampled(f, 0.5, 30, alpha=(1.0, 2.0, 3.0), mask=m,
                                           masked=True, **KPRM)
    np.testing.assert_array_equal(u, uo)
    assert rel_err(l, lo) < 3e-5 and rel_err(r, ro) < 3e-5


def test_processor_mirrors(orc):
    from ambient_image_processor_amd.processor import (ResizeProcessor,
                                                       TGVKDeshadeDownsampledNotMaskedProcessor,
                                                       TGVKDeshadeDownsampledProcessor)
    shape = (8, 18, 26)
    f = _f(shape, 10)
    m = _mask(shape, 11)
    a = ResizeProcessor.process(f, 0.5)
    np.testing.assert_array_equal(a, orc.resize_spline3(f, orc.resize_extents(shape, 0.5)))
    b = ResizeProcessor.process(a, 26, 18, 8)
    np.testing.assert_array_equal(b, orc.resize_spline3(a, shape))
    seen = []
    u, l, r, div_v = TGVKDeshadeDownsampledProcessor.processTGVKL1Cuda(
        f, 0.5, KPRM["lam"], 3, (1.0, 2.0, 3.0), 20, mask=m, add_background_back=True,
        paint_iteration_interval=10, iteration_finished_callback=lambda it, n, uu, ll, rr:
        seen.append((it, uu.shape)) or False, calculate_div_v=True)
    uo, lo, ro = orc.deshade_k_downsampled(f, 0.5, 20, alpha=(1.0, 2.0, 3.0), mask=m,
                                           add_background_back=True, **KPRM)
    np.testing.assert_array_equal(u, uo)
    assert rel_err(l, lo) < 1e-5 and rel_err(r, ro) < 1e-5
    assert seen == [(10, orc.resize_extents(shape, 0.5))]
    assert div_v.shape == shape
    u2, l2, r2, _ = TGVKDeshadeDownsampledNotMaskedProcessor.processTGVKL1Cuda(
        f, 0.5, KPRM["lam"], 3, (1.0, 2.0, 3.0), 20, mask=m)
    uo2, lo2, ro2 = orc.deshade_k_downsampled(f, 0.5, 20, alpha=(1.0, 2.0, 3.0), mask=m,
                                              masked=False, **KPRM)
    np.testing.assert_array_equal(u2, uo2)
    assert rel_err(r2, ro2) < 1e-5


def test_downsampled_cli(orc, tmp_path):
    """tgvk_deshade_downsampled_application with a mask: alpha[i] = i + 1, negatives
    to zero, background added back; <prefix>_shading.mha and _deshaded.mha."""
    import subprocess
    from ambient_image_processor_amd import _build, metaimage
    shape = (10, 16, 24)
    f = _f(shape, 12)
    m = _mask(shape, 13)
    src, msk = str(tmp_path / "in.mha"), str(tmp_path / "mask.mha")
    metaimage.write(src, f)
    metaimage.write(msk, m)
    prefix = str(tmp_path / "out")
    _say("running", _build.CLI_DOWNSAMPLED)
    r = subprocess.run([_build.CLI_DOWNSAMPLED, src, "0.5", "3", str(KPRM["lam"]), "25", msk,
                        prefix], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    l = metaimage.read(prefix + "_shading.mha")[0]
    rr = metaimage.read(prefix + "_deshaded.mha")[0]
    uo, lo, ro = orc.deshade_k_downsampled(f, 0.5, 25, alpha=(1.0, 2.0, 3.0), mask=m,
                                           add_background_back=True, **KPRM)
    assert rel_err(l, lo) < 1e-5 and rel_err(rr, ro) < 1e-5
    bad = subprocess.run([_build.CLI_DOWNSAMPLED, src], capture_output=True, text=True,
                         timeout=60)
    assert bad.returncode == 1 and "parameter:" in bad.stdout
    assert os.path.exists(prefix + "_deshaded.mha")
