"""CPU tests that pin the oracle's cubic B-spline resampling (the ITK 4.8.1
ResampleImageFilter + BSplineInterpolateImageFunction path of
resize/ResizeProcessor.cpp, restated in oracle/tgv_oracle.cpp) against
scipy.ndimage's cubic B-spline with mirror boundaries (the same spline: whole-
sample symmetric extension, Unser's recursive prefilter), the interpolation
property, reproduction of constants and of linear ramps away from the
boundaries (which pins the resampling geometry, including ResizeProcessor's
+2 upsampling quirk), and the outside-sample rule."""
import numpy as np
import pytest
import scipy.ndimage as ndi


def _coords(n_out, n_in, pad):
    """Continuous input indices of the kept output voxels along one axis."""
    sf = np.float32(n_out + pad) / np.float32(n_in)
    return np.arange(n_out) * (1.0 / np.float64(sf))


def _scipy_resize(a, out_shape):
    up = any(o > i for o, i in zip(out_shape, a.shape))
    pad = 2 if up else 0
    axes = [_coords(o, i, pad) for o, i in zip(out_shape, a.shape)]
    zz, yy, xx = np.meshgrid(*axes, indexing="ij")
    return ndi.map_coordinates(a.astype(np.float64), [zz, yy, xx], order=3, mode="mirror",
                               prefilter=True)


@pytest.mark.parametrize("shape", [(7, 30, 41), (1, 12, 9), (3, 5, 80), (2, 2, 2)])
def test_coefficients_match_scipy_mirror_prefilter(orc, shape):
    a = np.random.default_rng(1).standard_normal(shape).astype(np.float32)
    c = orc.spline3_coefficients(a)
    axes = [ax for ax in range(3) if shape[ax] > 1]
    ref = a.astype(np.float64)
    for ax in axes:
        ref = ndi.spline_filter1d(ref, order=3, axis=ax, mode="mirror")
    # lines longer than 18 samples start from the 18-term truncated sum (tolerance 1e-10)
    assert np.abs(c - ref).max() <= 1e-9 * max(1.0, np.abs(ref).max())


@pytest.mark.parametrize("shape,factor", [((9, 30, 41), 0.5), ((8, 25, 33), 0.3),
                                          ((1, 40, 52), 0.5), ((6, 7, 64), 0.25)])
def test_down_and_up_resize_match_scipy(orc, shape, factor):
    a = np.random.default_rng(2).standard_normal(shape).astype(np.float32)
    small = orc.resize_extents(shape, factor)
    down = orc.resize_spline3(a, small)
    assert down.shape == small
    ref = _scipy_resize(a, small)
    np.testing.assert_allclose(down, ref, rtol=0, atol=3e-7 * np.abs(ref).max())
    back = orc.resize_spline3(down, shape)
    ref_back = _scipy_resize(down, shape)
    np.testing.assert_allclose(back, ref_back, rtol=0, atol=3e-7 * max(np.abs(ref_back).max(), 1))


def test_resize_extents_follow_float_ceil(orc):
    assert orc.resize_extents((256, 256, 256), 0.5) == (128, 128, 128)
    # float(100) * 0.3f rounds to 30.000002: the reference's float ceil gives 31
    assert orc.resize_extents((1, 100, 101), 0.3) == (1, 31, 31)
    assert orc.resize_extents((3, 10, 10), 0.01) == (1, 1, 1)


def test_halving_samples_the_even_voxels(orc):
    """Factor 0.5 from even extents: output j samples input 2j exactly (integer
    positions), where the spline passes through the data."""
    a = np.random.default_rng(3).standard_normal((8, 20, 30)).astype(np.float32)
    down = orc.resize_spline3(a, (4, 10, 15))
    np.testing.assert_allclose(down, a[::2, ::2, ::2], rtol=0, atol=2e-6)


def test_constant_is_reproduced(orc):
    a = np.full((5, 9, 17), 0.625, np.float32)
    for shape in ((3, 5, 9), (5, 9, 17), (10, 18, 34)):
        np.testing.assert_allclose(orc.resize_spline3(a, shape), 0.625, rtol=0, atol=1e-7)


def test_upsampling_geometry_quirk_with_a_ramp(orc):
    """Upsampling 40 -> 96 along x resamples with 98 voxels (ResizeProcessor's +2)
    and keeps the first 96: away from the mirror boundaries a cubic spline
    reproduces a linear ramp, so output j reads j * 40 / 98, not j * 40 / 96."""
    x = np.arange(40, dtype=np.float32)
    a = np.broadcast_to(x, (3, 4, 40)).copy()
    up = orc.resize_spline3(a, (3, 4, 96))
    j = np.arange(30, 70)
    np.testing.assert_allclose(up[1, 2, j], j * (1.0 / np.float64(np.float32(98) / np.float32(40))),
                               rtol=0, atol=1e-5)
    assert np.abs(up[1, 2, j] - j * 40 / 96).max() > 0.1


def test_samples_outside_the_input_are_zero(orc):
    """A 3-voxel line up to 30: the resampling geometry has 32 voxels, output j
    reads j * 3 / 32, and j * 3/32 >= 2.5 (j >= 27) lies outside -> 0."""
    a = np.array([[[1.0, 2.0, 4.0]]], np.float32)
    up = orc.resize_spline3(a, (1, 1, 30))
    x = np.arange(30) * (1.0 / np.float64(np.float32(32) / np.float32(3)))
    outside = x >= 2.5
    assert outside.any() and not outside.all()
    assert np.all(up[0, 0, outside] == 0) and np.all(up[0, 0, ~outside] > 0)
