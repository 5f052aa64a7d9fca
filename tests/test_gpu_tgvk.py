"""GPU tests of the TGV-k (k >= 3) deshade (reference
tgvk_deshade/tgvk_l1_deshade.cu, its 2-D form tgvk_l1_deshade_2d.cu and the
processor TGVKDeshadeProcessor.cpp:76-174).

The two TGV-k sweeps perform the oracle's per-voxel operations in the same
order, so u and v are compared bit for bit with the FP32 oracle after enough
iterations for every level to feed back into them (odd and even levels, the
last-level rule and q~ = r_{i-1} all reach v through w_0)."""
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


def _params(native, iterations, **kw):
    return native.default_params(lam=KPRM["lam"], tau=TAU, sigma=TAU, theta=1.0,
                                 iterations=iterations, data_term=native.DATA_L1, **kw)


def _f(shape, seed):
    from ambient_image_processor_amd import phantoms
    return phantoms.deshade_phantom(shape, seed=seed)[0]


SHAPES = [(9, 13, 21), (8, 12, 64), (5, 70, 33), (17, 4, 130), (3, 3, 3), (1, 20, 28)]
ALPHAS = [(1.0, 2.0, 0.6), (1.0, 2.0, 0.8, 0.5), (0.8, 1.5, 0.7, 0.4, 0.3)]


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("alpha", ALPHAS)
def test_tgvk_bitwise_against_oracle(ctx, orc, shape, alpha):
    from ambient_image_processor_amd import native
    _say("tgvk", shape, "order", len(alpha))
    f = _f(shape, 3)
    u, vx, vy, vz = ctx.tgvk_deshade(f, _params(native, 25), alpha)
    st = orc.tgvk(f, 25, alpha=alpha, dtype=np.float32, **KPRM)
    np.testing.assert_array_equal(u, st.u)
    np.testing.assert_array_equal(vx, st.v[0])
    np.testing.assert_array_equal(vy, st.v[1])
    np.testing.assert_array_equal(vz, st.v[2])


@pytest.mark.parametrize("zchunk", ["1", "3", "7"])
@pytest.mark.parametrize("alpha", ALPHAS + [(1.0, 2.0, 0.9, 0.8, 0.7, 0.6)])
def test_tgvk_z_chunks(ctx, orc, monkeypatch, zchunk, alpha):
    """Chunk edges and the -z register carries of the fixed-level kernels (1 to
    3 levels) and the generic kernels (4 levels)."""
    from ambient_image_processor_amd import native
    _say("tgvk zchunk", zchunk, "order", len(alpha))
    ctx.set_override("tgvk_zchunk", int(zchunk))
    f = _f((17, 9, 70), 17)
    u, vx, vy, vz = ctx.tgvk_deshade(f, _params(native, 15), alpha)
    st = orc.tgvk(f, 15, alpha=alpha, dtype=np.float32, **KPRM)
    np.testing.assert_array_equal(u, st.u)
    np.testing.assert_array_equal(vz, st.v[2])


def test_tgvk_many_levels(ctx, orc):
    from ambient_image_processor_amd import native
    alpha = (1.0, 2.0, 0.9, 0.8, 0.7, 0.6, 0.5, 0.4)
    f = _f((6, 10, 40), 4)
    u, vx, vy, vz = ctx.tgvk_deshade(f, _params(native, 30), alpha)
    st = orc.tgvk(f, 30, alpha=alpha, dtype=np.float32, **KPRM)
    np.testing.assert_array_equal(u, st.u)
    np.testing.assert_array_equal(vz, st.v[2])


@pytest.mark.parametrize("alpha", ALPHAS[:2])
def test_tgvk_planar_bitwise(ctx, orc, alpha):
    from ambient_image_processor_amd import native
    f = _f((1, 24, 36), 5)
    u, vx, vy, vz = ctx.tgvk_deshade(f, _params(native, 30, flags=native.FLAG_PLANAR), alpha)
    st = orc.tgvk(f, 30, alpha=alpha, dtype=np.float32, planar=True, **KPRM)
    np.testing.assert_array_equal(u, st.u)
    np.testing.assert_array_equal(vx, st.v[0])
    np.testing.assert_array_equal(vy, st.v[1])
    assert not np.any(vz)


def test_tgvk_callback_and_stop(ctx, orc):
    from ambient_image_processor_amd import native
    f = _f((7, 9, 20), 6)
    alpha = (1.0, 2.0, 0.6, 0.4)
    seen = []

    def cb(it, count, u, vx, vy, vz):
        seen.append(it)
        return it >= 20

    u, vx, vy, vz = ctx.tgvk_deshade(f, _params(native, 100, callback_interval=10), alpha, cb)
    assert seen == [10, 20]
    st = orc.tgvk(f, 21, alpha=alpha, dtype=np.float32, **KPRM)
    np.testing.assert_array_equal(u, st.u)
    np.testing.assert_array_equal(vy, st.v[1])


def test_tgvk_launch_counts(ctx):
    from ambient_image_processor_amd import native
    f = _f((5, 8, 16), 7)
    ctx.profile(True)
    ctx.profile_reset()
    ctx.tgvk_deshade(f, _params(native, 13), (1.0, 2.0, 0.5, 0.5, 0.5))
    prof = ctx.profile_read()
    ctx.profile(False)
    assert prof["tgvk_dual"][1] == 13 and prof["tgvk_primal"][1] == 13
    assert prof["dual_sweep"][1] == 0 and prof["fused_sweep"][1] == 0
    assert prof["tgvk_fused"][1] == 0


# ---------------------------------------------------------------------------
# fused sweep (k = 3: primal half of iteration k + dual half of iteration k + 1)
# ---------------------------------------------------------------------------
FUSED_SHAPES = [(9, 13, 21), (12, 10, 64), (5, 70, 33), (17, 4, 130), (3, 3, 3), (6, 9, 256),
                (2, 5, 341)]


@pytest.mark.parametrize("shape", FUSED_SHAPES)
@pytest.mark.parametrize("tile,band,zchunk", [("0", "0", "0"), ("1000", "1", "1"),
                                              ("1000", "2", "3"), ("8", "1", "1"),
                                              ("16", "2", "3"), ("64", "5", "7"),
                                              ("30", "13", "2")])
def test_tgvk_fused_bitwise_against_oracle(ctx, orc, monkeypatch, shape, tile, band, zchunk):
    """Tile, band and chunk edges of the fused sweep: the halo columns, rows
    and planes, the +z loads one plane ahead and the -z carries (tile 1000:
    whole rows).  A tile that does not fit (more rows than the volume, or more
    than 1024 threads) leaves the two launches."""
    from ambient_image_processor_amd import native
    _say("fused tgvk", shape, "tile", tile, "band", band, "zchunk", zchunk)
    d, h, w = shape
    tx = min(int(tile), w)
    txh = tx + 2 if tx < w else w
    fits = tile == "0" or (int(band) <= h and (int(band) + 2) * txh <= 1024)
    if tile != "0":
        ctx.set_override("tgvk_fused_tx", int(tile))
        ctx.set_override("tgvk_fused_ty", int(band))
        ctx.set_override("tgvk_fused_zchunk", int(zchunk))
    f = _f(shape, 30)
    alpha = (1.0, 2.0, 0.6)
    ctx.profile(True)
    ctx.profile_reset()
    u, vx, vy, vz = ctx.tgvk_deshade(f, _params(native, 18), alpha)
    prof = ctx.profile_read()
    ctx.profile(False)
    assert prof["tgvk_fused"][1] == (17 if fits else 0), prof
    st = orc.tgvk(f, 18, alpha=alpha, dtype=np.float32, **KPRM)
    np.testing.assert_array_equal(u, st.u)
    np.testing.assert_array_equal(vx, st.v[0])
    np.testing.assert_array_equal(vy, st.v[1])
    np.testing.assert_array_equal(vz, st.v[2])


@pytest.mark.parametrize("shape", [(7, 11, 48), (1, 24, 36)])
def test_tgvk_fused_equals_two_launches(ctx, monkeypatch, shape):
    """u and v after the fused loop equal the two-launch loop's, with the
    projection of r binding (alpha[2] = 0.05), also in the 2-D form."""
    from ambient_image_processor_amd import native
    f = _f(shape, 31)
    alpha = (1.0, 2.0, 0.05)   # a binding projection of r
    p = _params(native, 40, flags=native.FLAG_PLANAR if shape[0] == 1 else 0)
    fused = ctx.tgvk_deshade(f, p, alpha)
    ctx.set_override("tgvk_fused", 0)
    two = ctx.tgvk_deshade(f, p, alpha)
    for a, b in zip(fused, two):
        np.testing.assert_array_equal(a, b)


def test_tgvk_fused_callback_and_stop(ctx, orc):
    """Observed iterations end with the primal launch alone; the state the
    callback sees and an early stop match the oracle."""
    from ambient_image_processor_amd import native
    f = _f((7, 9, 20), 32)
    alpha = (1.0, 2.0, 0.6)
    seen = []

    def cb(it, count, u, vx, vy, vz):
        st = orc.tgvk(f, it + 1, alpha=alpha, dtype=np.float32, **KPRM)
        seen.append((it, np.array_equal(u, st.u) and np.array_equal(vz, st.v[2])))
        return it >= 14

    ctx.profile(True)
    ctx.profile_reset()
    u, vx, vy, vz = ctx.tgvk_deshade(f, _params(native, 100, callback_interval=7), alpha, cb)
    prof = ctx.profile_read()
    ctx.profile(False)
    assert seen == [(7, True), (14, True)]
    # iterations 0..14: dual launches at 0, 8 (after the observed 7), primal
    # launches at 7 and 14, fused steps for the rest
    assert prof["tgvk_dual"][1] == 2 and prof["tgvk_primal"][1] == 2
    assert prof["tgvk_fused"][1] == 13
    st = orc.tgvk(f, 15, alpha=alpha, dtype=np.float32, **KPRM)
    np.testing.assert_array_equal(u, st.u)
    np.testing.assert_array_equal(vy, st.v[1])


def test_tgvk_fused_wide_rows_and_fallback(ctx, orc, monkeypatch):
    """Rows wider than 341 voxels run in x tiles; whole rows of 400 voxels do
    not fit a 3-row band of 1024 threads, so that forced tile leaves the two
    launches."""
    from ambient_image_processor_amd import native
    f = _f((3, 5, 400), 33)
    alpha = (1.0, 2.0, 0.6)
    st = orc.tgvk(f, 10, alpha=alpha, dtype=np.float32, **KPRM)
    ctx.profile(True)
    ctx.profile_reset()
    u, vx, vy, vz = ctx.tgvk_deshade(f, _params(native, 10), alpha)
    prof = ctx.profile_read()
    assert prof["tgvk_fused"][1] == 9 and prof["tgvk_dual"][1] == 1
    np.testing.assert_array_equal(u, st.u)
    np.testing.assert_array_equal(vx, st.v[0])
    ctx.set_override("tgvk_fused_tx", 400)
    ctx.profile_reset()
    u, vx, vy, vz = ctx.tgvk_deshade(f, _params(native, 10), alpha)
    prof = ctx.profile_read()
    ctx.profile(False)
    assert prof["tgvk_fused"][1] == 0 and prof["tgvk_dual"][1] == 10
    np.testing.assert_array_equal(u, st.u)
    np.testing.assert_array_equal(vx, st.v[0])


@pytest.mark.parametrize("with_mask", [False, True])
def test_tgvk_pipeline_against_oracle(ctx, orc, with_mask):
    """processTGVKL1Cuda: the TGV-k loop, the Poisson integration, the mask on
    l and r, the clamp and the add-back."""
    from ambient_image_processor_amd import native
    f = _f((10, 14, 22), 8)
    alpha = (1.0, 2.0, 0.6)
    m = None
    flags = native.FLAG_SET_NEGATIVE_ZERO
    if with_mask:
        m = (np.random.default_rng(9).random(f.shape) > 0.3).astype(np.float32)
        flags |= native.FLAG_ADD_BACKGROUND_BACK
    u, l, r = ctx.tgvk_deshade_pipeline(f, _params(native, 40, flags=flags), alpha, m)
    uo, lo, ro, _ = orc.deshade_k(f, 40, alpha=alpha, mask=m, set_negative_zero=True,
                                  add_background_back=with_mask, **KPRM)
    np.testing.assert_array_equal(u, uo)
    assert rel_err(l, lo) < 1e-5 and rel_err(r, ro) < 1e-5


def test_tgvk_pipeline_order_2_and_depth_1(ctx, orc):
    from ambient_image_processor_amd import native
    f = _f((6, 10, 16), 10)
    u, l, r = ctx.tgvk_deshade_pipeline(f, _params(native, 30), (1.0, 2.0))
    uo, lo, ro, _ = orc.deshade_k(f, 30, alpha=(1.0, 2.0), **KPRM)
    np.testing.assert_array_equal(u, uo)
    assert rel_err(l, lo) < 1e-5
    f1 = _f((1, 18, 26), 11)
    for alpha in ((1.0, 2.0), (1.0, 2.0, 0.5)):
        u, l, r = ctx.tgvk_deshade_pipeline(f1, _params(native, 30), alpha)
        uo, lo, ro, _ = orc.deshade_k(f1, 30, alpha=alpha, **KPRM)
        np.testing.assert_array_equal(u, uo)
        assert rel_err(l, lo) < 3e-5 and rel_err(r, ro) < 3e-5


def test_tgvk_argument_errors(ctx):
    from ambient_image_processor_amd import native
    f = _f((3, 4, 8), 1)
    with pytest.raises(native.AmbError):
        ctx.tgvk_deshade(f, _params(native, 2), (1.0, 2.0))
    with pytest.raises(native.AmbError):
        ctx.tgvk_deshade(f, _params(native, 2), (1.0, 2.0, 0.0))
    with pytest.raises(native.AmbError):
        p = _params(native, 2)
        p.data_term = native.DATA_L2
        ctx.tgvk_deshade(f, p, (1.0, 2.0, 0.5))
    with pytest.raises(native.AmbError):
        ctx.tgvk_deshade(f, _params(native, 2, storage=native.STORAGE_BF16), (1.0, 2.0, 0.5))
    with pytest.raises(native.AmbError):
        ctx.tgvk_deshade(f, _params(native, 2, flags=native.FLAG_PLANAR), (1.0, 2.0, 0.5))
    # the context still works afterwards
    u, *_ = ctx.tgvk_deshade(f, _params(native, 0), (1.0, 2.0, 0.5))
    np.testing.assert_array_equal(u, f)


def test_processor_mirror_with_callback_and_div_v(orc):
    """TGVKDeshadeProcessor.processTGVKL1Cuda: outputs, the callback's
    (u, l, r) and div v."""
    from ambient_image_processor_amd.processor import TGVKDeshadeProcessor
    f = _f((8, 12, 20), 12)
    m = (np.random.default_rng(13).random(f.shape) > 0.25).astype(np.float32)
    alpha = (1.0, 2.0, 0.6)
    got = {}

    def cb(it, count, u, l, r):
        got[it] = (u.copy(), l.copy(), r.copy())
        return False

    u, l, r, div_v = TGVKDeshadeProcessor.processTGVKL1Cuda(
        f, KPRM["lam"], 3, alpha, 30, mask=m, add_background_back=True,
        paint_iteration_interval=15, iteration_finished_callback=cb, calculate_div_v=True)
    uo, lo, ro, st = orc.deshade_k(f, 30, alpha=alpha, mask=m, add_background_back=True, **KPRM)
    np.testing.assert_array_equal(u, uo)
    assert rel_err(l, lo) < 1e-5 and rel_err(r, ro) < 1e-5
    np.testing.assert_array_equal(div_v, orc.divergence(*st.v, dtype=np.float32))
    assert sorted(got) == [15]
    u16, l16, r16, _ = orc.deshade_k(f, 16, alpha=alpha, mask=m, add_background_back=True,
                                     **KPRM)
    np.testing.assert_array_equal(got[15][0], u16)
    assert rel_err(got[15][1], l16) < 1e-5 and rel_err(got[15][2], r16) < 1e-5
    seen = []
    TGVKDeshadeProcessor.updateAlpha(4, seen.append)
    assert seen == [1, 4, 7, 10]


def test_reference_tgvk_launchers(orc):
    from ambient_image_processor_amd import processor
    f = _f((5, 9, 24), 14)
    d, h, w = f.shape
    alpha = (1.0, 2.0, 0.7, 0.4)
    u, vx, vy, vz = processor.tgvk_l1_deshade_launch(f.ravel(), w, h, d, KPRM["lam"], 20, 0,
                                                     None, 4, alpha)
    st = orc.tgvk(f, 20, alpha=alpha, dtype=np.float32, **KPRM)
    np.testing.assert_array_equal(u, st.u)
    np.testing.assert_array_equal(vz, st.v[2])
    f1 = _f((1, 16, 22), 15)
    u, vx, vy = processor.tgvk_l1_deshade_launch_2d(f1.ravel(), 22, 16, KPRM["lam"], 20, 0, None,
                                                    4, alpha)
    st = orc.tgvk(f1, 20, alpha=alpha, dtype=np.float32, planar=True, **KPRM)
    np.testing.assert_array_equal(u, st.u)
    np.testing.assert_array_equal(vy, st.v[1])


@pytest.mark.parametrize("two_d", [False, True])
def test_reference_tgvk_launcher_caller(orc, tmp_path, two_d):
    """tests/compat/processor_demo.cpp --tgvk: tgvk_l1_deshade_launch(_2d) of the
    compatibility header, then the Poisson integration and r = u - l, clamp."""
    import os
    import subprocess
    from ambient_image_processor_amd import _build
    assert os.path.exists(_build.COMPAT_DEMO), "run __graft_entry__.build() first"
    shape = (1, 18, 30) if two_d else (6, 10, 24)
    f = _f(shape, 16)
    d, h, w = shape
    f.tofile(tmp_path / "f.raw")
    out = [str(tmp_path / n) for n in ("u.raw", "l.raw", "r.raw")]
    alpha = (1.0, 2.0, 0.6, 0.4)
    args = [_build.COMPAT_DEMO] + (["--2d"] if two_d else []) + [
        "--tgvk", ",".join(str(a) for a in alpha), str(tmp_path / "f.raw"), str(w), str(h),
        str(d), str(KPRM["lam"]), "0", "0", "25"] + out
    _say("running", " ".join(args[:4]))
    r = subprocess.run(args, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    u, l, rr = (np.fromfile(p, dtype=np.float32).reshape(shape) for p in out)
    uo, lo, ro, _ = orc.deshade_k(f, 25, alpha=alpha, **KPRM)
    np.testing.assert_array_equal(u, uo)
    np.testing.assert_array_equal(l, lo)
    np.testing.assert_array_equal(rr, ro)


# ---------------------------------------------------------------------------
# masked TGV-k (tgvk_deshade_masked/)
# ---------------------------------------------------------------------------
def _mask(shape, seed):
    rng = np.random.default_rng(seed)
    return np.where(rng.random(shape) > 0.35, 1.0, rng.random(shape) * 0.999).astype(np.float32)


@pytest.mark.parametrize("alpha", ALPHAS + [(1.0, 2.0, 0.9, 0.8, 0.7, 0.6)])
@pytest.mark.parametrize("zchunk", ["0", "4"])
def test_masked_tgvk_bitwise(ctx, orc, monkeypatch, alpha, zchunk):
    from ambient_image_processor_amd import native
    _say("masked tgvk order", len(alpha), "zchunk", zchunk)
    if zchunk != "0":
        ctx.set_override("tgvk_zchunk", int(zchunk))
    shape = (11, 13, 68)
    f = _f(shape, 18)
    m = _mask(shape, 19)
    u, vx, vy, vz = ctx.tgvk_deshade_masked(f, _params(native, 20), alpha, m >= 1)
    st = orc.tgvk(f, 20, alpha=alpha, dtype=np.float32, mask=m, **KPRM)
    np.testing.assert_array_equal(u, st.u)
    np.testing.assert_array_equal(vx, st.v[0])
    np.testing.assert_array_equal(vz, st.v[2])
    out = m < 1
    np.testing.assert_array_equal(u[out], f[out])


@pytest.mark.parametrize("order,add_back", [(3, False), (3, True), (4, True), (2, False)])
def test_masked_tgvk_processor(orc, order, add_back):
    """TGVKDeshadeMaskedProcessor.processTGVKL1Cuda against the oracle's
    composition (order 2: the masked TGV2 processor with swapped weights)."""
    from ambient_image_processor_amd.processor import TGVKDeshadeMaskedProcessor
    shape = (9, 14, 24)
    f = _f(shape, 20)
    m = _mask(shape, 21)
    alpha = (1.0, 2.0, 0.6, 0.4)[:order]
    u, l, r, div_v = TGVKDeshadeMaskedProcessor.processTGVKL1Cuda(
        f, KPRM["lam"], order, alpha, 25, mask=m, add_background_back=add_back,
        calculate_div_v=True)
    uo, lo, ro, st = orc.deshade_k_masked(f, 25, alpha=alpha, mask=m,
                                          add_background_back=add_back, **KPRM)
    np.testing.assert_array_equal(u, uo)
    assert rel_err(l, lo) < 1e-5 and rel_err(r, ro) < 1e-5
    np.testing.assert_array_equal(div_v, orc.divergence(*st.v, dtype=np.float32))


def test_masked_tgvk_depth1_and_launchers(orc):
    from ambient_image_processor_amd import processor
    from ambient_image_processor_amd.processor import (TGVDeshadeMaskedProcessor,
                                                       TGVKDeshadeMaskedProcessor)
    shape = (1, 26, 30)
    f = _f(shape, 22)
    m = _mask(shape, 23)
    alpha = (1.0, 2.0, 0.6)
    u, l, r, _ = TGVKDeshadeMaskedProcessor.processTGVKL1Cuda(f, KPRM["lam"], 3, alpha, 25,
                                                              mask=m)
    uo, lo, ro, _ = orc.deshade_k_masked(f, 25, alpha=alpha, mask=m, **KPRM)
    np.testing.assert_array_equal(u, uo)
    assert rel_err(l, lo) < 3e-5 and rel_err(r, ro) < 3e-5
    lists = TGVDeshadeMaskedProcessor.buildMaskIndices2D(m)
    u2, vx2, vy2 = processor.tgvk_l1_deshade_masked_2d_launch(
        f.ravel(), 30, 26, KPRM["lam"], 25, 0, -1, None, 3, alpha, *lists)
    st = orc.tgvk(f, 25, alpha=alpha, dtype=np.float32, planar=True, mask=m, **KPRM)
    np.testing.assert_array_equal(u2, st.u)
    np.testing.assert_array_equal(vy2, st.v[1])
    f3 = _f((5, 8, 20), 24)
    m3 = _mask(f3.shape, 25)
    lists3 = TGVDeshadeMaskedProcessor.buildMaskIndices(m3)
    u3, vx3, vy3, vz3 = processor.tgvk_l1_deshade_masked_launch(
        f3.ravel(), 20, 8, 5, KPRM["lam"], 25, 0, -1, None, 3, alpha, *lists3)
    st3 = orc.tgvk(f3, 25, alpha=alpha, dtype=np.float32, mask=m3, **KPRM)
    np.testing.assert_array_equal(u3, st3.u)
    np.testing.assert_array_equal(vz3, st3.v[2])
