"""GPU tests of the 64-bit-offset kernel forms, which volumes of 2^30 voxels and
more run (every field then spans 4 GiB or more).  The plan override index64 = 1
(amb_context_set_plan_override) forces those forms on small volumes: they perform the same per-voxel arithmetic, so their
results equal the 32-bit forms' bit for bit (TGV2 fused and two-sweep, L1 and
L2, FP32 and BF16 fields, masked, 2-D; TGV-k through the generic kernels).
Real 2049 x 1024 x 512 and 1025 x 1024 x 1024 volumes (whole-row bands and x
tiles of the fused sweep) are checked in a block at the far end of the volume, where byte offsets exceed 2^32, against the FP32 oracle run
on a crop around it (after N iterations a voxel depends only on voxels within
2N of it, so the crop's own edges do not reach the compared block)."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TAU = float(np.float32(1 / np.sqrt(20)))


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _f(shape, seed):
    from ambient_image_processor_amd import phantoms
    return phantoms.deshade_phantom(shape, seed=seed)[0]


def _p(native, iterations, data_term, **kw):
    return native.default_params(lam=0.7, alpha0=2.0, alpha1=1.0, tau=TAU, sigma=TAU, theta=1.0,
                                 iterations=iterations, data_term=data_term, **kw)


def _both(ctx, run):
    a = run()
    ctx.set_override("index64", 1)
    b = run()
    ctx.set_override("index64", -1)
    return a, b


@pytest.mark.parametrize("shape", [(8, 12, 64), (7, 9, 30), (5, 6, 41)])
@pytest.mark.parametrize("data_term,storage", [(0, 0), (1, 0), (0, 1)])
@pytest.mark.parametrize("fused", ["1", "0"])
def test_tgv2_index64_equals_index32(ctx, monkeypatch, shape, data_term, storage, fused):
    from ambient_image_processor_amd import native
    _say("index64 tgv2", shape, data_term, storage, fused)
    ctx.set_override("fused", int(fused))
    f = _f(shape, 40)
    p = _p(native, 12, data_term, storage=storage)
    a, b = _both(ctx, lambda: ctx.tgv2_deshade(f, p))
    for x, y in zip(a, b):
        np.testing.assert_array_equal(x, y)


@pytest.mark.parametrize("shape,tile", [((6, 9, 96), "24"), ((5, 7, 72), "40")])
def test_tgv2_index64_x_tiles(ctx, orc, monkeypatch, shape, tile):
    """x tiles of the four-voxel 64-bit form (rows too wide for whole-row
    bands at 2^30 voxels and more), forced on small volumes, against the
    oracle bit for bit (FP32 fields, L1)."""
    from ambient_image_processor_amd import native
    ctx.set_override("index64", 1)
    ctx.set_override("fused_tx", int(tile))
    f = _f(shape, 46)
    ctx.profile(True)
    ctx.profile_reset()
    u, vx, vy, vz = ctx.tgv2_deshade(f, _p(native, 12, 0))
    prof = ctx.profile_read()
    ctx.profile(False)
    assert prof["fused_sweep"][1] == 11, prof
    st = orc.tgv2(f, 12, lam=0.7, alpha0=2.0, alpha1=1.0, tau=TAU, sigma=TAU, data_term="l1",
                  dtype=np.float32)
    np.testing.assert_array_equal(u, st.u)
    for k, v in enumerate((vx, vy, vz)):
        np.testing.assert_array_equal(v, st.v[k])


@pytest.mark.parametrize("data_term,storage,masked", [(1, 0, False), (0, 1, False), (0, 0, True),
                                                       (1, 1, True)])
def test_tgv2_index64_x_tiles_equal_two_sweeps(ctx, monkeypatch, data_term, storage, masked):
    """The same 64-bit x tiles with L2, BF16 fields and the masked iteration
    equal the two-sweep loop of the 64-bit kernels bit for bit."""
    from ambient_image_processor_amd import native
    _say("index64 x tiles", data_term, storage, masked)
    ctx.set_override("index64", 1)
    f = _f((5, 9, 88), 47)
    act = np.random.default_rng(48).random(f.shape) > 0.3
    p = _p(native, 12, data_term, storage=storage)

    def run():
        ctx.profile(True)
        ctx.profile_reset()
        out = ctx.tgv2_deshade_masked(f, p, act) if masked else ctx.tgv2_deshade(f, p)
        prof = ctx.profile_read()
        ctx.profile(False)
        return out, prof["fused_sweep"][1]

    ctx.set_override("fused_tx", 32)
    a, n_fused = run()
    assert n_fused == 11
    ctx.set_override("fused", 0)
    b, n_two = run()
    assert n_two == 0
    for x, y in zip(a, b):
        np.testing.assert_array_equal(x, y)


def test_tgv2_index64_masked_and_planar(ctx, monkeypatch):
    from ambient_image_processor_amd import native
    f = _f((6, 10, 48), 41)
    act = np.random.default_rng(42).random(f.shape) > 0.3
    p = _p(native, 12, 0)
    a, b = _both(ctx, lambda: ctx.tgv2_deshade_masked(f, p, act))
    for x, y in zip(a, b):
        np.testing.assert_array_equal(x, y)
    f1 = _f((1, 20, 36), 43)
    p1 = _p(native, 12, 0, flags=native.FLAG_PLANAR)
    a, b = _both(ctx, lambda: ctx.tgv2_deshade(f1, p1))
    for x, y in zip(a, b):
        np.testing.assert_array_equal(x, y)


@pytest.mark.parametrize("alpha", [(1.0, 2.0, 0.6), (1.0, 2.0, 0.8, 0.5)])
def test_tgvk_index64_equals_index32(ctx, monkeypatch, alpha):
    """k = 3 runs the fused sweep with 32-bit offsets and the generic kernels
    with 64-bit ones; k = 4 the fixed-level and the generic kernels."""
    from ambient_image_processor_amd import native
    f = _f((7, 11, 40), 44)
    p = native.default_params(lam=0.7, tau=TAU, sigma=TAU, theta=1.0, iterations=15,
                              data_term=native.DATA_L1)
    a, b = _both(ctx, lambda: ctx.tgvk_deshade(f, p, alpha))
    for x, y in zip(a, b):
        np.testing.assert_array_equal(x, y)


@pytest.mark.parametrize("dims", [(2049, 1024, 512), (1025, 1024, 1024)])
def test_tgv2_over_2_30_voxels_against_cropped_oracle(ctx, orc, dims):
    """2049 x 1024 x 512 and 1025 x 1024 x 1024 (4.3 GB per field), three
    TGV2-L1 iterations (dual sweep, two fused sweeps in the four-voxel form -
    whole rows of 512, x tiles for rows of 1024 - and the primal sweep, all
    with 64-bit offsets); the last 24 planes of a 24 x 24 column window at the
    x end equal the FP32 oracle run on the crop."""
    from ambient_image_processor_amd import native
    d, h, w = dims
    iters, m = 3, 8
    _say("generating the input", dims)
    f = np.random.default_rng(45).random((d, h, w), dtype=np.float32)
    _say("GPU: 3 iterations at", dims)
    ctx.profile(True)
    ctx.profile_reset()
    u, vx, vy, vz = ctx.tgv2_deshade(f, _p(native, iters, 0))
    prof = ctx.profile_read()
    ctx.profile(False)
    assert prof["fused_sweep"][1] == iters - 1, prof
    z0, y0, x0, n = d - 24, 500, w - 24, 24
    crop = f[z0 - m:, y0 - m:y0 + n + m, x0 - m:]
    _say("oracle on the crop", crop.shape)
    st = orc.tgv2(crop, iters, lam=0.7, alpha0=2.0, alpha1=1.0, tau=TAU, sigma=TAU,
                  data_term="l1", dtype=np.float32)
    sl = (slice(z0, d), slice(y0, y0 + n), slice(x0, w))
    cs = (slice(m, None), slice(m, m + n), slice(m, None))
    np.testing.assert_array_equal(u[sl], st.u[cs])
    np.testing.assert_array_equal(vx[sl], st.v[0][cs])
    np.testing.assert_array_equal(vz[sl], st.v[2][cs])
