"""GPU tests of the fused sweep (primal half of iteration k + dual half of
iteration k + 1 in one launch, u/v/p/q ping-ponged between two copies).

The fused schedule performs the same per-voxel operations in the same order
as the two-sweep loop, so every result is compared bit for bit: with the two
sweeps (plan override fused = 0) and with the FP32 oracle.  Cases cover band and
z-chunk edges (voxels per thread, band heights, chunk lengths and prefetch
forced through the overrides fused_vx / _ty / _zchunk / _pf, including
1024-thread blocks),
partial bands, odd widths (one voxel per thread, whole rows and x tiles),
depth 1, L1/L2, BF16 storage, the masked loop, callbacks (which split the
fused schedule) and continuing from a state the fused loop left in the
second buffer copy."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TAU = float(np.float32(1 / np.sqrt(20)))
PRM = dict(lam=0.5, alpha0=2.0, alpha1=1.0, tau=TAU, sigma=TAU, theta=1.0)


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _params(native, iterations, data_term="l1", **kw):
    return native.default_params(lam=PRM["lam"], alpha0=PRM["alpha0"], alpha1=PRM["alpha1"],
                                 tau=TAU, sigma=TAU, theta=1.0, iterations=iterations,
                                 data_term=native.DATA_L1 if data_term == "l1" else native.DATA_L2,
                                 **kw)


def _f(shape, seed):
    from ambient_image_processor_amd import phantoms
    return phantoms.deshade_phantom(shape, seed=seed)[0]


def _profile_counts(ctx, fn):
    ctx.profile(True)
    ctx.profile_reset()
    out = fn()
    prof = ctx.profile_read()
    ctx.profile(False)
    return out, prof


# odd widths run one voxel per thread: whole rows up to 341 columns, x tiles of
# 126 columns beyond (401: three full tiles and one of 23 columns)
SHAPES = [(9, 13, 32), (17, 20, 64), (6, 7, 128), (5, 11, 256), (3, 4, 8), (1, 12, 36),
          (20, 3, 16), (9, 13, 21), (6, 5, 1), (1, 7, 33), (4, 9, 401)]


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("data_term", ["l1", "l2"])
def test_fused_equals_two_sweeps_and_oracle(ctx, orc, shape, data_term, monkeypatch):
    from ambient_image_processor_amd import native
    _say("fused", shape, data_term)
    f = _f(shape, 3)
    p = _params(native, 23, data_term)
    ctx.set_override("fused", 0)
    ref = ctx.tgv2_deshade(f, p)
    ctx.set_override("fused", -1)
    out, prof = _profile_counts(ctx, lambda: ctx.tgv2_deshade(f, p))
    assert prof["fused_sweep"][1] == 22 and prof["dual_sweep"][1] == 1
    assert prof["primal_sweep"][1] == 1
    for a, b in zip(out, ref):
        np.testing.assert_array_equal(a, b)
    st = orc.tgv2(f, 23, data_term=data_term, dtype=np.float32, **PRM)
    np.testing.assert_array_equal(out[0], st.u)
    np.testing.assert_array_equal(out[3], st.v[2])


@pytest.mark.parametrize("vx,ty,zchunk,pf,shape", [
    ("4", "1", "1", "1", (7, 9, 32)), ("4", "3", "2", "0", (11, 10, 64)),
    ("4", "2", "5", "1", (13, 17, 40)), ("4", "6", "3", "1", (8, 5, 256)),
    ("4", "4", "7", "0", (1, 9, 64)), ("4", "4", "64", "1", (70, 13, 256)),
    ("2", "6", "4", "0", (9, 23, 256)),      # 128 x 8 = 1024 threads per block
    ("2", "1", "1", "0", (7, 9, 30)), ("2", "3", "5", "0", (6, 13, 66)),
    ("2", "2", "3", "0", (5, 4, 8)),
    ("1", "2", "3", "0", (7, 9, 31)), ("1", "5", "4", "0", (5, 6, 255))])
def test_fused_forced_plans(ctx, monkeypatch, vx, ty, zchunk, pf, shape):
    from ambient_image_processor_amd import native
    _say("fused plan vx", vx, "ty", ty, "zchunk", zchunk, "pf", pf, shape)
    f = _f(shape, 5)
    p = _params(native, 9)
    ctx.set_override("fused", 0)
    ref = ctx.tgv2_deshade(f, p)
    ctx.set_override("fused", -1)
    for k, v in (("VX", vx), ("TY", ty), ("ZCHUNK", zchunk), ("PF", pf)):
        ctx.set_override("fused_" + k.lower(), int(v))
    out, prof = _profile_counts(ctx, lambda: ctx.tgv2_deshade(f, p))
    assert prof["fused_sweep"][1] == 8
    for a, b in zip(out, ref):
        np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("vx", ["1", "2", "4"])
def test_fused_masked_and_bf16_per_width(ctx, monkeypatch, vx):
    from ambient_image_processor_amd import native
    shape = (9, 11, 47) if vx == "1" else (9, 11, 48)
    f = _f(shape, 12)
    m = (np.random.default_rng(13).random(shape) > 0.4).astype(np.float32)
    for storage in (native.STORAGE_F32, native.STORAGE_BF16):
        p = _params(native, 14, storage=storage)
        ctx.set_override("fused", 0)
        ref = ctx.tgv2_deshade_masked(f, p, m >= 1)
        ctx.set_override("fused", -1)
        ctx.set_override("fused_vx", int(vx))
        out, prof = _profile_counts(ctx, lambda: ctx.tgv2_deshade_masked(f, p, m >= 1))
        ctx.set_override("fused_vx", -1)
        assert prof["fused_sweep"][1] == 13
        for a, b in zip(out, ref):
            np.testing.assert_array_equal(a, b)


def test_fused_bf16_storage(ctx, monkeypatch):
    from ambient_image_processor_amd import native
    f = _f((12, 18, 64), 6)
    p = _params(native, 30, storage=native.STORAGE_BF16)
    ctx.set_override("fused", 0)
    ref = ctx.tgv2_deshade(f, p)
    ctx.set_override("fused", -1)
    out = ctx.tgv2_deshade(f, p)
    for a, b in zip(out, ref):
        np.testing.assert_array_equal(a, b)


def test_fused_masked(ctx, orc, monkeypatch):
    from ambient_image_processor_amd import native
    shape = (10, 14, 32)
    f = _f(shape, 7)
    m = (np.random.default_rng(8).random(shape) > 0.3).astype(np.float32)
    p = _params(native, 25)
    ctx.set_override("fused", 0)
    ref = ctx.tgv2_deshade_masked(f, p, m >= 1)
    ctx.set_override("fused", -1)
    out, prof = _profile_counts(ctx, lambda: ctx.tgv2_deshade_masked(f, p, m >= 1))
    assert prof["fused_sweep"][1] == 24
    for a, b in zip(out, ref):
        np.testing.assert_array_equal(a, b)
    st = orc.tgv2(f, 25, data_term="l1", dtype=np.float32, mask=m, **PRM)
    np.testing.assert_array_equal(out[0], st.u)


def test_fused_with_callback(ctx, monkeypatch):
    from ambient_image_processor_amd import native
    f = _f((8, 12, 32), 9)
    p = _params(native, 40, callback_interval=7)
    seen = {}

    def make(store):
        def cb(it, count, u, vx, vy, vz):
            store[it] = (u.copy(), vz.copy())
            return it >= 28
        return cb

    ctx.set_override("fused", 0)
    ref_seen = {}
    ref = ctx.tgv2_deshade(f, p, make(ref_seen))
    ctx.set_override("fused", -1)
    out = ctx.tgv2_deshade(f, p, make(seen))
    assert sorted(seen) == sorted(ref_seen) == [7, 14, 21, 28]
    for it in seen:
        np.testing.assert_array_equal(seen[it][0], ref_seen[it][0])
        np.testing.assert_array_equal(seen[it][1], ref_seen[it][1])
    for a, b in zip(out, ref):
        np.testing.assert_array_equal(a, b)


def test_fused_state_continues_and_energy(ctx, monkeypatch):
    """begin + iterate(7) + iterate(6) (the second call starts from the copy the
    first left the state in) equals 13 two-sweep iterations in every field,
    and the FP64 energy of the final state agrees."""
    from ambient_image_processor_amd import native
    f = _f((9, 10, 64), 10)
    p = _params(native, 0)
    ctx.set_override("fused", 0)
    ctx.begin(f, p)
    ctx.iterate(13)
    ref_state = ctx.get_state()
    ref_e = ctx.energy()
    ctx.set_override("fused", -1)
    ctx.begin(f, p)
    ctx.iterate(7)
    ctx.iterate(6)
    state = ctx.get_state()
    np.testing.assert_array_equal(state, ref_state)
    np.testing.assert_array_equal(ctx.energy(), ref_e)


# ---------------------------------------------------------------------------
# x tiles (rows too wide for whole-row bands; forced on small volumes)
# ---------------------------------------------------------------------------
TILE_CASES = [((9, 13, 64), "16", "2", "3"), ((9, 13, 64), "30", "5", "0"),
              ((5, 20, 130), "62", "0", "2"), ((7, 9, 34), "8", "1", "1"),
              ((6, 7, 40), "38", "3", "4"), ((7, 9, 35), "8", "1", "1"),
              ((6, 7, 41), "15", "3", "4"), ((5, 20, 131), "63", "0", "2")]


@pytest.mark.parametrize("shape,tile,band,zchunk", TILE_CASES)
@pytest.mark.parametrize("data_term", ["l1", "l2"])
def test_fused_x_tiles_against_oracle(ctx, orc, monkeypatch, shape, tile, band, zchunk,
                                      data_term):
    """Halo lanes on both sides of a tile, partial last tiles, band and chunk
    edges: u and v equal the FP32 oracle bit for bit."""
    from ambient_image_processor_amd import native
    ctx.set_override("fused_tx", int(tile))
    if band != "0":
        ctx.set_override("fused_ty", int(band))
    if zchunk != "0":
        ctx.set_override("fused_zchunk", int(zchunk))
    f = _f(shape, 50)
    dt = native.DATA_L1 if data_term == "l1" else native.DATA_L2
    p = native.default_params(lam=0.7, alpha0=2.0, alpha1=1.0, tau=TAU, sigma=TAU, theta=1.0,
                              iterations=14, data_term=dt)
    ctx.profile(True)
    ctx.profile_reset()
    u, vx, vy, vz = ctx.tgv2_deshade(f, p)
    prof = ctx.profile_read()
    ctx.profile(False)
    assert prof["fused_sweep"][1] == 13, prof
    st = orc.tgv2(f, 14, lam=0.7, alpha0=2.0, alpha1=1.0, tau=TAU, sigma=TAU,
                  data_term=data_term, dtype=np.float32)
    np.testing.assert_array_equal(u, st.u)
    np.testing.assert_array_equal(vx, st.v[0])
    np.testing.assert_array_equal(vy, st.v[1])
    np.testing.assert_array_equal(vz, st.v[2])


@pytest.mark.parametrize("tile", ["30", "62"])
def test_fused_x_tiles_bf16_masked_planar(ctx, orc, monkeypatch, tile):
    """BF16 fields, the masked loop and the 2-D loop in x tiles equal the
    two-sweep loop (and, for the FP32 forms, the oracle) bit for bit."""
    from ambient_image_processor_amd import native
    f = _f((6, 11, 96), 51)
    pb = native.default_params(lam=0.7, alpha0=2.0, alpha1=1.0, tau=TAU, sigma=TAU, theta=1.0,
                               iterations=12, data_term=native.DATA_L1,
                               storage=native.STORAGE_BF16)
    act = np.random.default_rng(52).random(f.shape) > 0.3
    pm = native.default_params(lam=0.7, alpha0=2.0, alpha1=1.0, tau=TAU, sigma=TAU, theta=1.0,
                               iterations=12, data_term=native.DATA_L1)
    f2 = _f((1, 30, 96), 53)
    p2 = native.default_params(lam=0.7, alpha0=2.0, alpha1=1.0, tau=TAU, sigma=TAU, theta=1.0,
                               iterations=12, data_term=native.DATA_L1, flags=native.FLAG_PLANAR)
    ctx.set_override("fused_tx", int(tile))
    tiled = [ctx.tgv2_deshade(f, pb), ctx.tgv2_deshade_masked(f, pm, act),
             ctx.tgv2_deshade(f2, p2)]
    ctx.set_override("fused", 0)
    two = [ctx.tgv2_deshade(f, pb), ctx.tgv2_deshade_masked(f, pm, act), ctx.tgv2_deshade(f2, p2)]
    for a, b in zip(tiled, two):
        for x, y in zip(a, b):
            np.testing.assert_array_equal(x, y)
    st = orc.tgv2(f2, 12, lam=0.7, alpha0=2.0, alpha1=1.0, tau=TAU, sigma=TAU, data_term="l1",
                  dtype=np.float32, planar=True)
    np.testing.assert_array_equal(tiled[2][0], st.u)


def test_fused_odd_tiles_bf16_and_masked(ctx, orc):
    """Odd rows wider than whole-row bands allow (263: three 126-column tiles,
    the last of 11 columns), one voxel per thread: BF16 fields and the masked
    loop equal the two-sweep loop bit for bit, and the masked FP32 run equals
    the oracle."""
    from ambient_image_processor_amd import native
    f = _f((5, 9, 263), 55)
    act = np.random.default_rng(56).random(f.shape) > 0.35
    kw = dict(lam=0.7, alpha0=2.0, alpha1=1.0, tau=TAU, sigma=TAU, theta=1.0, iterations=11,
              data_term=native.DATA_L1)
    pb = native.default_params(storage=native.STORAGE_BF16, **kw)
    pm = native.default_params(**kw)
    ctx.profile(True)
    ctx.profile_reset()
    fused = [ctx.tgv2_deshade(f, pb), ctx.tgv2_deshade_masked(f, pm, act)]
    prof = ctx.profile_read()
    ctx.profile(False)
    assert prof["fused_sweep"][1] == 20, prof
    ctx.set_override("fused", 0)
    two = [ctx.tgv2_deshade(f, pb), ctx.tgv2_deshade_masked(f, pm, act)]
    for a, b in zip(fused, two):
        for x, y in zip(a, b):
            np.testing.assert_array_equal(x, y)
    st = orc.tgv2(f, 11, lam=0.7, alpha0=2.0, alpha1=1.0, tau=TAU, sigma=TAU, data_term="l1",
                  dtype=np.float32, mask=act.astype(np.float32))
    np.testing.assert_array_equal(fused[1][0], st.u)


@pytest.mark.parametrize("shape", [(3, 6, 1400), (4, 19, 760)])
def test_fused_wide_rows_run_tiled(ctx, orc, shape):
    """Rows too wide for whole-row bands take x tiles of 124 columns and bands
    of up to 8 rows by themselves (760: six full tiles and one of 16 columns,
    bands of 8, 8 and 3 rows)."""
    from ambient_image_processor_amd import native
    f = _f(shape, 54)
    p = native.default_params(lam=0.7, alpha0=2.0, alpha1=1.0, tau=TAU, sigma=TAU, theta=1.0,
                              iterations=10, data_term=native.DATA_L1)
    ctx.profile(True)
    ctx.profile_reset()
    u, vx, vy, vz = ctx.tgv2_deshade(f, p)
    prof = ctx.profile_read()
    ctx.profile(False)
    assert prof["fused_sweep"][1] == 9, prof
    st = orc.tgv2(f, 10, lam=0.7, alpha0=2.0, alpha1=1.0, tau=TAU, sigma=TAU, data_term="l1",
                  dtype=np.float32)
    np.testing.assert_array_equal(u, st.u)
    for k, v in enumerate((vx, vy, vz)):
        np.testing.assert_array_equal(v, st.v[k])
