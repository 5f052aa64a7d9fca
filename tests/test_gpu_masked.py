"""GPU parity of the masked TGV2-L1 deshade (reference tgv_deshade_masked/):
the masked sweeps (amb_tgv2_deshade_masked), the masked pipeline
composition (AMB_FLAG_MASKED_ITERATION), the Python and C++ reference-signature
callers and the CLI's mask path, against the oracle's masked loop on the same
seeded inputs.

The masked loop performs, at every active voxel, the unmasked per-voxel
operations in the same order (tgv2_masked_common.cu:53-88, 170-222, 313-361),
so the FP32 GPU result is compared with the FP32 oracle bit for bit.  One
difference is not reproduced: the reference's masked p update takes |p| with
norm3df (tgv2_masked_common.cu:79) where the unmasked one sums squares; both
are within an ulp of |p|, so against the reference itself parity is at the
1e-4 tolerance, not bitwise."""
import os
import sys

import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu

TAU = float(np.float32(1 / np.sqrt(20)))
PRM = dict(lam=0.5, alpha0=2.0, alpha1=1.0, tau=TAU, sigma=TAU, theta=1.0)


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _params(native, iterations, **kw):
    return native.default_params(lam=PRM["lam"], alpha0=PRM["alpha0"], alpha1=PRM["alpha1"],
                                 tau=TAU, sigma=TAU, theta=1.0, iterations=iterations,
                                 data_term=native.DATA_L1, **kw)


def _mask(shape, seed, kind="blob"):
    """Float mask: a ball-like blob (the usual case: a foreground region), or
    scattered voxels, or fractional values (only values >= 1 are active)."""
    rng = np.random.default_rng(seed)
    d, h, w = shape
    if kind == "random":
        return (rng.random(shape) > 0.35).astype(np.float32)
    z, y, x = np.meshgrid(np.arange(d), np.arange(h), np.arange(w), indexing="ij")
    c = np.array(shape) / 2.0
    r2 = ((z - c[0]) / max(c[0], 1)) ** 2 + ((y - c[1]) / c[1]) ** 2 + ((x - c[2]) / c[2]) ** 2
    m = (r2 < 0.8).astype(np.float32)
    if kind == "fractional":
        m = np.where(m > 0, 1.0, rng.random(shape) * 0.999).astype(np.float32)
        m[::3, ::2, ::5] = 1.5
    return m


# odd widths: one voxel per thread; widths % 4 == 0: the vector sweeps
SHAPES = [(9, 13, 21), (8, 12, 32), (6, 10, 68), (5, 7, 132)]


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("kind", ["blob", "random", "fractional"])
def test_masked_iteration_matches_oracle(ctx, orc, shape, kind):
    from ambient_image_processor_amd import native, phantoms
    _say("masked iteration", shape, kind)
    f = phantoms.deshade_phantom(shape, seed=11)[0]
    m = _mask(shape, 3, kind)
    p = _params(native, 40)
    u, vx, vy, vz = ctx.tgv2_deshade_masked(f, p, m >= 1)
    st = orc.tgv2(f, 40, data_term="l1", dtype=np.float32, mask=m, **PRM)
    np.testing.assert_array_equal(u, st.u)
    for a, b in zip((vx, vy, vz), st.v):
        np.testing.assert_array_equal(a, b)
    out = m < 1
    np.testing.assert_array_equal(u[out], f[out])          # u = f outside the mask
    for a in (vx, vy, vz):
        assert np.all(a[out] == 0)                          # v = 0 outside the mask


def test_all_active_equals_unmasked(ctx):
    from ambient_image_processor_amd import native, phantoms
    f = phantoms.deshade_phantom((10, 12, 36), seed=2)[0]
    p = _params(native, 25)
    a = ctx.tgv2_deshade(f, p)
    b = ctx.tgv2_deshade_masked(f, p, np.ones(f.shape, np.uint8))
    for x, y in zip(a, b):
        np.testing.assert_array_equal(x, y)


def test_masked_bf16_storage_close_to_fp32(ctx, orc):
    from ambient_image_processor_amd import native, phantoms
    shape = (12, 16, 32)
    f = phantoms.deshade_phantom(shape, seed=5)[0]
    m = _mask(shape, 4)
    p = _params(native, 60, storage=native.STORAGE_BF16)
    u, vx, vy, vz = ctx.tgv2_deshade_masked(f, p, m >= 1)
    st = orc.tgv2(f, 60, data_term="l1", dtype=np.float32, mask=m, **PRM)
    e = rel_err(u, st.u)
    _say("bf16 masked u rel err", e)
    assert e < 2e-2
    out = m < 1
    assert np.all(vx[out] == 0) and np.all(vz[out] == 0)


@pytest.mark.parametrize("add_back", [False, True])
def test_masked_pipeline_matches_oracle(ctx, orc, add_back):
    from ambient_image_processor_amd import native, phantoms
    shape = (14, 18, 24)
    f = phantoms.deshade_phantom(shape, seed=6)[0]
    m = _mask(shape, 7, "fractional")
    flags = native.FLAG_MASKED_ITERATION | native.FLAG_SET_NEGATIVE_ZERO
    if add_back:
        flags |= native.FLAG_ADD_BACKGROUND_BACK
    p = _params(native, 50, flags=flags)
    u, l, r = ctx.deshade_pipeline(f, p, m)
    uo, lo, ro, _ = orc.deshade_masked(f, 50, mask=m, add_background_back=add_back, **PRM)
    np.testing.assert_array_equal(u, uo)
    np.testing.assert_array_equal(l, lo)
    np.testing.assert_array_equal(r, ro)
    if not add_back:
        assert np.all(r[m == 0] == 0)
    else:
        zero = m == 0
        np.testing.assert_array_equal(r[zero], f[zero])     # background = u = f there


def test_masked_processor_mirror_and_callback(ctx, orc):
    """TGVDeshadeMaskedProcessor.processTGV2L1GPUCuda: outputs, the callback's
    (u, l, r) at iteration 20 (21 iterations done), and div v."""
    from ambient_image_processor_amd import phantoms, processor
    shape = (10, 14, 20)
    f = phantoms.deshade_phantom(shape, seed=8)[0]
    m = _mask(shape, 9)
    seen = {}

    def cb(it, count, u, l, r):
        seen[it] = (u.copy(), l.copy(), r.copy())
        return False

    r, u, l, div_v = processor.TGVDeshadeMaskedProcessor.processTGV2L1GPUCuda(
        f, PRM["lam"], PRM["alpha0"], PRM["alpha1"], 30, -1, 10, cb, m,
        set_negative_values_to_zero=True, add_background_back=True, calculate_div_v=True)
    assert sorted(seen) == [10, 20]
    uo, lo, ro, st = orc.deshade_masked(f, 30, mask=m, add_background_back=True, **PRM)
    np.testing.assert_array_equal(u, uo)
    np.testing.assert_array_equal(l, lo)
    np.testing.assert_array_equal(r, ro)
    np.testing.assert_array_equal(div_v, orc.divergence(*st.v, dtype=np.float32))
    u21, l21, r21, _ = orc.deshade_masked(f, 21, mask=m, add_background_back=True, **PRM)
    np.testing.assert_array_equal(seen[20][0], u21)
    np.testing.assert_array_equal(seen[20][1], l21)
    np.testing.assert_array_equal(seen[20][2], r21)


def test_masked_launch_mirror_with_index_lists(ctx, orc):
    from ambient_image_processor_amd import native, phantoms, processor
    shape = (7, 9, 16)
    f = phantoms.deshade_phantom(shape, seed=10)[0]
    m = _mask(shape, 12, "random")
    lists = processor.TGVDeshadeMaskedProcessor.buildMaskIndices(m)
    d, h, w = shape
    u, vx, vy, vz = processor.tgv2_l1_deshade_masked_launch(
        f.ravel(), w, h, d, PRM["lam"], 30, 0, -1, None, PRM["alpha0"], PRM["alpha1"], *lists)
    st = orc.tgv2(f, 30, data_term="l1", dtype=np.float32, mask=m, **PRM)
    np.testing.assert_array_equal(u, st.u)
    np.testing.assert_array_equal(vy, st.v[1])


def test_masked_reference_launcher_caller(orc, tmp_path):
    """tests/compat/processor_demo.cpp with a mask: buildMaskIndices-style
    lists into tgv2_l1_deshade_masked_launch, then the masked composition."""
    import subprocess
    from ambient_image_processor_amd import _build, phantoms
    assert os.path.exists(_build.COMPAT_DEMO), "run __graft_entry__.build() first"
    shape = (9, 12, 20)
    f = phantoms.deshade_phantom(shape, seed=13)[0]
    m = _mask(shape, 14)
    d, h, w = shape
    f.tofile(tmp_path / "f.raw")
    m.tofile(tmp_path / "m.raw")
    out = [str(tmp_path / n) for n in ("u.raw", "l.raw", "r.raw")]
    _say("running", _build.COMPAT_DEMO, "with mask")
    r = subprocess.run([_build.COMPAT_DEMO, str(tmp_path / "f.raw"), str(w), str(h), str(d),
                        "0.5", "2.0", "1.0", "40"] + out + [str(tmp_path / "m.raw")],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    u, l, rr = (np.fromfile(p, dtype=np.float32).reshape(shape) for p in out)
    uo, lo, ro, _ = orc.deshade_masked(f, 40, mask=m, add_background_back=False, **PRM)
    np.testing.assert_array_equal(u, uo)
    np.testing.assert_array_equal(l, lo)
    np.testing.assert_array_equal(rr, ro)


def test_cli_with_mask(orc, tmp_path):
    """tgv2_deshade_application with a mask argument: the reference CLI's masked
    branch (negatives to zero, background added back)."""
    import subprocess
    from ambient_image_processor_amd import _build, metaimage, phantoms
    shape = (8, 14, 18)
    f = phantoms.deshade_phantom(shape, seed=15)[0]
    m = _mask(shape, 16)
    src, msk = str(tmp_path / "in.mha"), str(tmp_path / "mask.mha")
    metaimage.write(src, f)
    metaimage.write(msk, m)
    outs = [str(tmp_path / n) for n in ("u.mha", "l.mha", "r.mha")]
    _say("running", _build.CLI, "with mask")
    r = subprocess.run([_build.CLI, src, "0.5", "2", "1", "30", "-1", msk] + outs,
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    u, l, rr = (metaimage.read(p)[0] for p in outs)
    uo, lo, ro, _ = orc.deshade_masked(f, 30, mask=m, add_background_back=True, **PRM)
    np.testing.assert_array_equal(u, uo)
    np.testing.assert_array_equal(l, lo)
    np.testing.assert_array_equal(rr, ro)


def test_deshade_epilogue_entry(ctx, orc):
    from ambient_image_processor_amd import native
    rng = np.random.default_rng(3)
    shape = (5, 6, 7)
    u = rng.standard_normal(shape).astype(np.float32)
    l = rng.standard_normal(shape).astype(np.float32)
    m = (rng.random(shape) * 1.5).astype(np.float32)
    for flags, kw in [
        (native.FLAG_SET_NEGATIVE_ZERO, dict(mask_shading=False, set_negative_zero=True)),
        (native.FLAG_MASK_SHADING | native.FLAG_ADD_BACKGROUND_BACK,
         dict(mask_shading=True, set_negative_zero=False, add_background_back=True)),
        (native.FLAG_MASKED_ITERATION | native.FLAG_SET_NEGATIVE_ZERO,
         dict(mask_shading=True, set_negative_zero=True, mask_again=True)),
    ]:
        r, lo = ctx.deshade_epilogue(u, l, m, flags)
        ro, lo2 = orc.deshade_epilogue(u, l, m, **kw)
        np.testing.assert_array_equal(r, ro)
        np.testing.assert_array_equal(lo, lo2)


def test_masked_rejects_bad_arguments(ctx):
    from ambient_image_processor_amd import native
    f = np.zeros((4, 4, 4), np.float32)
    with pytest.raises(ValueError):
        ctx.tgv2_deshade_masked(f, _params(native, 2), np.ones((4, 4, 5), np.uint8))
    with pytest.raises(native.AmbError):
        native._check(native.lib().amb_tgv2_deshade_masked(
            ctx._h, native.MEM_HOST, native._ptr(f), 4, 4, 4,
            native.ctypes.byref(_params(native, 2)), None, None, None, None, None,
            native.CALLBACK(), None))


@pytest.mark.parametrize("zchunk,storage,cb_every", [(0, 0, 0), (4, 0, 7), (3, 1, 0), (5, 0, 0)])
def test_masked_fused_tiles_without_active_voxels_skip(ctx, orc, zchunk, storage, cb_every):
    """A small ball mask leaves whole bands and z chunks of the fused sweep
    inactive; from the second fused launch of a call those tiles skip (both
    copies of the state then agree outside the mask).  u and v equal the two-sweep
    masked loop bit for bit (and the FP32 oracle's masked loop for FP32 fields),
    also with callbacks, which end iterations with the primal sweep alone."""
    from ambient_image_processor_amd import native
    shape = (40, 36, 64)
    d, h, w = shape
    z, y, x = np.meshgrid(np.arange(d), np.arange(h), np.arange(w), indexing="ij")
    ball = ((z - 17) ** 2 + (y - 20) ** 2 + (x - 30) ** 2 <= 9 ** 2).astype(np.float32)
    f = (0.4 + 0.2 * np.random.default_rng(21).standard_normal(shape)).astype(np.float32)
    _say("tile skip zchunk", zchunk, "storage", storage, "callback every", cb_every)
    p = _params(native, 26, storage=storage, callback_interval=cb_every)
    seen = []

    def cb(it, count, u, vx, vy, vz):
        seen.append((it, u.copy()))
        return False

    if zchunk:
        ctx.set_override("fused_zchunk", zchunk)
    ctx.profile(True)
    ctx.profile_reset()
    out = ctx.tgv2_deshade_masked(f, p, ball >= 1, cb if cb_every else None)
    launches = ctx.profile_counts()["fused_sweep"]
    swept, total = ctx.profile_fused_tiles()
    ctx.profile(False)
    # the first fused launch of the call sweeps every tile, the others the active ones
    assert launches >= 2 and total % launches == 0
    per = total // launches
    assert (swept - per) % (launches - 1) == 0
    assert 0 < (swept - per) // (launches - 1) < per
    seen_fused = list(seen)
    seen.clear()
    ctx.clear_overrides()
    ctx.set_override("fused", 0)
    ref = ctx.tgv2_deshade_masked(f, p, ball >= 1, cb if cb_every else None)
    for a, b in zip(out, ref):
        np.testing.assert_array_equal(a, b)
    assert [s[0] for s in seen_fused] == [s[0] for s in seen]
    for (_, a), (_, b) in zip(seen_fused, seen):
        np.testing.assert_array_equal(a, b)
    if storage == 0:
        st = orc.tgv2(f, 26, data_term="l1", dtype=np.float32, mask=ball, **PRM)
        np.testing.assert_array_equal(out[0], st.u)
        np.testing.assert_array_equal(np.stack(out[1:]), st.v)


def test_unmasked_fused_sweep_counts_every_tile(ctx):
    """Without a mask every fused launch sweeps its whole grid."""
    from ambient_image_processor_amd import native
    f = (0.4 + 0.2 * np.random.default_rng(5).standard_normal((24, 20, 40))).astype(np.float32)
    ctx.profile(True)
    ctx.profile_reset()
    ctx.tgv2_deshade(f, _params(native, 9))
    launches = ctx.profile_counts()["fused_sweep"]
    swept, total = ctx.profile_fused_tiles()
    ctx.profile(False)
    assert launches >= 2 and total > 0 and total % launches == 0
    assert swept == total
