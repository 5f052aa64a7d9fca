"""GPU parity tests: the gfx950 HIP path (through the C ABI) against the CPU
oracle on the same seeded inputs.  Tolerances are norm-wise relative
(max |a - b| / max |b|); the FP32 path targets 1e-4 against the reference
algorithm, and its per-voxel op order is the oracle's FP32 mirror order, so
most comparisons with the FP32 oracle are much tighter."""
import os
import sys

import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu

PRM12 = dict(lam=1.0, alpha0=2.0, alpha1=1.0, tau=float(1 / np.sqrt(12)),
             sigma=float(1 / np.sqrt(12)), theta=1.0)
PRM20 = dict(lam=0.5, alpha0=2.0, alpha1=1.0, tau=float(1 / np.sqrt(20)),
             sigma=float(1 / np.sqrt(20)), theta=1.0)


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _params(native, prm, iterations, data_term, **kw):
    return native.default_params(lam=prm["lam"], alpha0=prm["alpha0"], alpha1=prm["alpha1"],
                                 tau=prm["tau"], sigma=prm["sigma"], theta=prm["theta"],
                                 iterations=iterations,
                                 data_term=native.DATA_L1 if data_term == "l1" else native.DATA_L2,
                                 **kw)


def _random_state(shape, seed):
    rng = np.random.default_rng(seed)
    st = rng.standard_normal((17,) + shape).astype(np.float32) * 0.3
    f = (0.5 + 0.2 * rng.standard_normal(shape)).astype(np.float32)
    return f, st


# (D, H, W): odd widths use the one-voxel-per-thread tiles, widths that are a
# multiple of 4 the vector tiles (partial tiles in x and y included).
SHAPES = [(11, 19, 37), (1, 20, 70), (9, 8, 64), (6, 5, 130), (7, 9, 132), (5, 13, 260),
          (3, 6, 8)]
# tile ids (tile_shape in csrc/amb_internal.hpp) x next-plane prefetch
VARIANTS = [("auto", "1"), ("auto", "0"), ("0", "1"), ("1", "1"), ("5", "1"), ("6", "0"),
            ("7", "1")]


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("data_term", ["l1", "l2"])
@pytest.mark.parametrize("zchunk", ["auto", "3"])
@pytest.mark.parametrize("variant", VARIANTS)
def test_one_iteration_from_random_state(ctx, orc, shape, data_term, zchunk, variant,
                                         monkeypatch):
    """Both sweeps on an arbitrary state (odd sizes, tile and z-chunk edges)."""
    from ambient_image_processor_amd import native
    _say("case", shape, data_term, zchunk, variant)
    if zchunk != "auto":
        ctx.set_override("zchunk", int(zchunk))
    else:
        ctx.set_override("zchunk", -1)
    if variant[0] != "auto":
        ctx.set_override("tile", int(variant[0]))
    else:
        ctx.set_override("tile", -1)
    ctx.set_override("sweep_pf", int(variant[1]))
    f, st = _random_state(shape, seed=sum(shape))
    p = _params(native, PRM12, 1, data_term)
    ctx.begin(f, p)
    ctx.set_state(st)
    ctx.iterate(2)
    got = ctx.get_state()
    ref = orc.TGVState(f, np.float32)
    ref.fields[:] = st
    ref.iterate(2, data_term=data_term, **PRM12)
    for k, name in enumerate(native.FIELD_NAMES):
        e = rel_err(got[k], ref.fields[k])
        assert e < 1e-6, (name, e)
    ndiff = int(np.sum(got != ref.fields))
    _say("fields differing bitwise from the FP32 oracle:", ndiff, "of", got.size)


def test_single_sweeps_order(ctx, orc):
    """Dual sweep changes only p, q; primal sweep only u, u_bar, v, v_bar."""
    from ambient_image_processor_amd import native
    f, st = _random_state((7, 9, 70), 3)
    ctx.begin(f, _params(native, PRM12, 1, "l2"))
    ctx.set_state(st)
    ctx.sweep(0)
    a = ctx.get_state()
    changed = [native.FIELD_NAMES[k] for k in range(17) if not np.array_equal(a[k], st[k])]
    assert set(changed) == {"px", "py", "pz", "qxx", "qyy", "qzz", "qxy", "qxz", "qyz"}
    ctx.sweep(1)
    b = ctx.get_state()
    changed = [native.FIELD_NAMES[k] for k in range(17) if not np.array_equal(b[k], a[k])]
    assert set(changed) == {"u", "ubar", "vx", "vy", "vz", "vbx", "vby", "vbz"}


def test_deshade_64_iteration_parity(ctx, orc):
    """64^3 L2 deshade workload, tau = sigma = 1/sqrt(12), 500 iterations."""
    from ambient_image_processor_amd import native, phantoms
    w = phantoms.workload("deshade_64")
    p = _params(native, PRM12, 500, "l2")
    u, vx, vy, vz = ctx.tgv2_deshade(w["f"], p)
    r32 = orc.tgv2(w["f"], 500, data_term="l2", dtype=np.float32, **PRM12)
    r64 = orc.tgv2(w["f"], 500, data_term="l2", dtype=np.float64, **PRM12)
    v = np.stack([vx, vy, vz])
    e32u, e32v = rel_err(u, r32.u), rel_err(v, r32.v)
    e64u, e64v = rel_err(u, r64.u), rel_err(v, r64.v)
    _say(f"64^3 L2 500 it: vs FP32 oracle u {e32u:.2e} v {e32v:.2e}; vs FP64 u {e64u:.2e} v {e64v:.2e}")
    assert e32u < 1e-6 and e32v < 1e-5
    assert e64u < 1e-4 and e64v < 1e-4


def test_l1_deshade_parity_32(ctx, orc):
    from ambient_image_processor_amd import native, phantoms
    f = phantoms.deshade_phantom((32, 32, 32), seed=2, kind="spheres", noise=0.03,
                                 bias_order=3, bias_sigma=0.05)[0]
    p = _params(native, PRM20, 200, "l1")
    u, vx, vy, vz = ctx.tgv2_deshade(f, p)
    r64 = orc.tgv2(f, 200, data_term="l1", dtype=np.float64, **PRM20)
    e = rel_err(u, r64.u), rel_err(np.stack([vx, vy, vz]), r64.v)
    _say("32^3 L1 200 it vs FP64:", e)
    assert e[0] < 1e-4 and e[1] < 1e-4


def test_callback_interval_and_stop(ctx, orc):
    """Callback at iteration_index > 0 with index % interval == 0; returning True
    after the k-th call equals running (index + 1) iterations."""
    from ambient_image_processor_amd import native, phantoms
    f = phantoms.deshade_phantom((12, 16, 20), seed=5)[0]
    seen = []

    def cb(it, count, u, vx, vy, vz):
        seen.append((it, count, float(u.sum())))
        return it >= 20

    p = _params(native, PRM12, 100, "l2", callback_interval=10)
    u, vx, vy, vz = ctx.tgv2_deshade(f, p, cb)
    assert [s[0] for s in seen] == [10, 20]
    assert all(s[1] == 100 for s in seen)
    ref = orc.tgv2(f, 21, data_term="l2", dtype=np.float32, **PRM12)
    assert rel_err(u, ref.u) < 1e-6
    assert rel_err(np.stack([vx, vy, vz]), ref.v) < 1e-5


def test_bf16_storage_runs_and_is_close(ctx, orc):
    from ambient_image_processor_amd import native, phantoms
    w = phantoms.workload("deshade_64")
    p = _params(native, PRM12, 200, "l2", storage=native.STORAGE_BF16)
    u, vx, vy, vz = ctx.tgv2_deshade(w["f"], p)
    r64 = orc.tgv2(w["f"], 200, data_term="l2", dtype=np.float64, **PRM12)
    eu = rel_err(u, r64.u)
    _say(f"BF16 storage 64^3 200 it: u rel err vs FP64 {eu:.2e}")
    assert eu < 2e-2


# ---------------------------------------------------------------------------
# DCT Poisson (A14-A18)
# ---------------------------------------------------------------------------
def test_divergence_bitwise(ctx, orc):
    rng = np.random.default_rng(1)
    v = rng.standard_normal((3, 9, 13, 70)).astype(np.float32)
    np.testing.assert_array_equal(ctx.divergence(*v), orc.divergence(*v, dtype=np.float32))


@pytest.mark.parametrize("shape", [(5, 6, 8), (16, 16, 16), (20, 24, 32), (7, 1, 9),
                                   (3, 70, 130), (65, 3, 2), (1, 1, 200), (17, 64, 65)])
def test_dct_reference_order(ctx, orc, shape):
    """The reference-numerics passes (tiled 64 lines x 64 outputs, 16-term chunks)
    equal the oracle's one-output-at-a-time FP32 sequence bit for bit, on tile
    and chunk edges."""
    import scipy.fft as sf
    from ambient_image_processor_amd import native
    rng = np.random.default_rng(7)
    x = rng.standard_normal(shape).astype(np.float32)
    X = ctx.cosine_transform(x, native.DCT_REFERENCE)
    Xo = orc.cosine_transform(x, np.float32)
    Y = ctx.inverse_cosine_transform(x, native.DCT_REFERENCE)
    Yo = orc.inverse_cosine_transform(x, np.float32)
    np.testing.assert_array_equal(X, Xo)
    np.testing.assert_array_equal(Y, Yo)
    assert rel_err(X, sf.dctn(x.astype(np.float64), type=2)) < 2e-6


def test_poisson_divide_bitwise_and_quirk(ctx, orc):
    rng = np.random.default_rng(2)
    x = rng.standard_normal((10, 12, 14)).astype(np.float32)
    np.testing.assert_array_equal(ctx.solve_poisson(x), orc.solve_poisson(x, dtype=np.float32))
    np.testing.assert_array_equal(ctx.solve_poisson(x, index_quirk=True),
                                  orc.solve_poisson(x, index_quirk=True, dtype=np.float32))


def test_poisson_quirk_above_2_24(ctx, orc):
    """At 2^24 + voxels the float coordinate recovery of the reference differs
    from the exact coordinates; the quirk flag reproduces the reference."""
    shape = (260, 256, 256)   # 17.0 M voxels
    x = np.random.default_rng(3).standard_normal(shape).astype(np.float32)
    a = ctx.solve_poisson(x, index_quirk=True)
    b = orc.solve_poisson(x, index_quirk=True, dtype=np.float32)
    np.testing.assert_array_equal(a, b)
    c = ctx.solve_poisson(x, index_quirk=False)
    nd = int(np.sum(a != c))
    _say("voxels whose quirk eigenvalue differs:", nd)
    assert nd > 0


def test_integrate_gradients_parity(ctx, orc):
    from ambient_image_processor_amd import native
    rng = np.random.default_rng(4)
    v = (0.01 * rng.standard_normal((3, 12, 16, 20))).astype(np.float32)
    got = ctx.integrate_gradients_poisson(*v, precision=native.DCT_REFERENCE)
    ref = orc.integrate_poisson(*v, dtype=np.float32)
    assert rel_err(got, ref) < 1e-5
    ref64 = orc.integrate_poisson(*v, dtype=np.float64)
    assert rel_err(got, ref64) < 1e-4


# ---------------------------------------------------------------------------
# Full deshade (A19) and energy (A21)
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("with_mask", [False, True])
def test_deshade_pipeline_parity(ctx, orc, with_mask):
    from ambient_image_processor_amd import native, phantoms
    f = phantoms.deshade_phantom((24, 28, 32), seed=9)[0]
    mask = None
    if with_mask:
        mask = (np.random.default_rng(1).random(f.shape) > 0.3).astype(np.float32)
    p = _params(native, PRM20, 120, "l1", flags=native.FLAG_SET_NEGATIVE_ZERO)
    u, l, r = ctx.deshade_pipeline(f, p, mask)
    st = orc.tgv2(f, 120, data_term="l1", dtype=np.float32, **PRM20)
    lo = orc.integrate_poisson(*st.v, dtype=np.float32)
    ro, lo2 = orc.deshade_epilogue(st.u, lo, mask, mask_shading=False, set_negative_zero=True)
    assert rel_err(u, st.u) < 1e-6
    assert rel_err(l, lo2) < 1e-4
    assert rel_err(r, ro) < 1e-4
    assert np.all(r >= 0)


def test_pipeline_flags_mask_shading_and_add_back(ctx, orc):
    from ambient_image_processor_amd import native, phantoms
    f = phantoms.deshade_phantom((16, 16, 16), seed=3)[0]
    mask = (np.random.default_rng(2).random(f.shape) > 0.5).astype(np.float32)
    flags = native.FLAG_MASK_SHADING | native.FLAG_ADD_BACKGROUND_BACK
    p = _params(native, PRM20, 50, "l1", flags=flags)
    u, l, r = ctx.deshade_pipeline(f, p, mask)
    st = orc.tgv2(f, 50, data_term="l1", dtype=np.float32, **PRM20)
    lo = orc.integrate_poisson(*st.v, dtype=np.float32)
    ro, lo2 = orc.deshade_epilogue(st.u, lo, mask, mask_shading=True, set_negative_zero=False,
                                   add_background_back=True)
    assert rel_err(l, lo2) < 1e-4 and rel_err(r, ro) < 1e-4
    assert np.all(l[mask == 0] == 0)


@pytest.mark.parametrize("data_term", ["l1", "l2"])
def test_energy_matches_fp64_oracle(ctx, orc, data_term):
    from ambient_image_processor_amd import native, phantoms
    f = phantoms.deshade_phantom((20, 24, 28), seed=4)[0]
    p = _params(native, PRM20, 30, data_term)
    ctx.begin(f, p)
    ctx.iterate(30)
    e_gpu = ctx.energy()
    st = ctx.get_state()
    ref = orc.TGVState(f, np.float64)
    ref.fields[:] = st
    e_ref = ref.energy(PRM20["lam"], PRM20["alpha0"], PRM20["alpha1"], data_term)
    for k in range(10):
        scale = max(abs(e_ref[k]), 1e-12 * abs(e_ref[0]))
        assert abs(e_gpu[k] - e_ref[k]) <= 1e-9 * scale + 1e-12, (k, e_gpu[k], e_ref[k])


def test_reference_named_launchers(orc):
    """The Python mirror of the reference's launcher signatures."""
    from ambient_image_processor_amd import processor, phantoms
    f = phantoms.deshade_phantom((8, 12, 16), seed=6)[0]
    d, h, w = f.shape
    u, vx, vy, vz = processor.tgv2_l1_deshade_launch(f.ravel(), w, h, d, 0.5, 40, 41, -1, None,
                                                     2.0, 1.0)
    ref = orc.tgv2(f, 40, lam=0.5, alpha0=2.0, alpha1=1.0, tau=float(np.float32(1 / np.sqrt(20))),
                   sigma=float(np.float32(1 / np.sqrt(20))), data_term="l1", dtype=np.float32)
    assert rel_err(u, ref.u) < 1e-6
    u2 = processor.tgv2_l2_launch(f, w, h, d, 1.0, 30, 0, None, 2.0, 1.0)
    ref2 = orc.tgv2(f, 30, lam=1.0, alpha0=2.0, alpha1=1.0, tau=float(np.float32(1 / np.sqrt(20))),
                    sigma=float(np.float32(1 / np.sqrt(20))), data_term="l2", dtype=np.float32)
    assert rel_err(u2, ref2.u) < 1e-6
    X = processor.cosine_transform_kernel_launch(f, w, h, d)
    assert rel_err(X, orc.cosine_transform(f, np.float32)) < 1e-6


# ---------------------------------------------------------------------------
# MFMA DCT passes (F32 and BF16x3 split precision)
# ---------------------------------------------------------------------------
MFMA_SHAPES = [(16, 16, 16), (5, 6, 8), (20, 24, 32), (7, 1, 9), (33, 17, 70), (64, 64, 64),
               (40, 130, 96), (300, 7, 12), (3, 520, 24), (9, 10, 600), (2, 131, 129)]


@pytest.mark.parametrize("shape", MFMA_SHAPES)
@pytest.mark.parametrize("precision,tol", [(1, 3e-6), (2, 4e-5)])
def test_dct_mfma_against_fp64(ctx, orc, shape, precision, tol):
    rng = np.random.default_rng(11)
    x = rng.standard_normal(shape).astype(np.float32)
    X = ctx.cosine_transform(x, precision)
    Y = ctx.inverse_cosine_transform(x, precision)
    X64 = orc.cosine_transform(x.astype(np.float64))
    Y64 = orc.inverse_cosine_transform(x.astype(np.float64))
    ex, ey = rel_err(X, X64), rel_err(Y, Y64)
    _say(f"MFMA DCT precision {precision} {shape}: fwd {ex:.2e} inv {ey:.2e}")
    assert ex < tol and ey < tol


@pytest.mark.parametrize("precision,tol", [(1, 1e-5), (2, 1e-4)])
def test_poisson_solve_mfma_known_answer(ctx, orc, precision, tol):
    """RHS = Neumann Laplacian of 64 cosine modes: the solve returns the field
    (fused eigenvalue divide in the forward z pass)."""
    from ambient_image_processor_amd import phantoms
    rhs, phi = phantoms.poisson_field((48, 80, 96), seed=3, modes=64, kmax=15)
    # integrate_gradients with v = (D+ phi) reproduces rhs via the B-divergence
    v = [orc.forward_difference(phi, a).astype(np.float32) for a in range(3)]
    l = ctx.integrate_gradients_poisson(*v, precision=precision)
    e = rel_err(l, phi - phi.mean())
    _say(f"MFMA Poisson precision {precision}: rel err vs closed form {e:.2e}")
    assert e < tol


def test_mfma_extent_limit_is_an_error(ctx):
    from ambient_image_processor_amd import native
    x = np.zeros((2, 2, 4100), dtype=np.float32)
    with pytest.raises(native.AmbError):
        ctx.cosine_transform(x, native.DCT_BF16X3)


def test_reference_launcher_caller_links_unchanged(ctx, orc, tmp_path):
    """tests/compat/processor_demo.cpp composes the reference's processor from the
    launcher templates of include/ambient_amd_reference_compat.hpp; its outputs
    equal the oracle's and the device-resident pipeline's."""
    import subprocess
    from ambient_image_processor_amd import _build, native, phantoms
    assert os.path.exists(_build.COMPAT_DEMO), "run __graft_entry__.build() first"
    f = phantoms.deshade_phantom((12, 20, 24), seed=8)[0]
    d, h, w = f.shape
    f.tofile(tmp_path / "f.raw")
    out = [str(tmp_path / n) for n in ("u.raw", "l.raw", "r.raw")]
    _say("running", _build.COMPAT_DEMO)
    r = subprocess.run([_build.COMPAT_DEMO, str(tmp_path / "f.raw"), str(w), str(h), str(d),
                        "0.5", "2.0", "1.0", "60"] + out, capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    u, l, rr = (np.fromfile(p, dtype=np.float32).reshape(f.shape) for p in out)
    tau = float(np.float32(1 / np.sqrt(20)))
    st = orc.tgv2(f, 60, lam=0.5, alpha0=2.0, alpha1=1.0, tau=tau, sigma=tau,
                  data_term="l1", dtype=np.float32)
    lo = orc.integrate_poisson(*st.v, dtype=np.float32)
    ro, _ = orc.deshade_epilogue(st.u, lo, None, False, True, False)
    np.testing.assert_array_equal(u, st.u)
    np.testing.assert_array_equal(l, lo)
    np.testing.assert_array_equal(rr, ro)
    p = native.default_params(lam=0.5, iterations=60)
    u2, l2, r2 = ctx.deshade_pipeline(f, p)
    np.testing.assert_array_equal(u2, u)
    np.testing.assert_array_equal(l2, l)
    np.testing.assert_array_equal(r2, rr)


def test_cli_deshade_metaimage(orc, tmp_path):
    """tgv2_deshade_application on a MetaImage input: the output trio equals the
    oracle's TGV2-L1 deshade (tau = sigma = 1/sqrt(20)) bit for bit."""
    import subprocess
    from ambient_image_processor_amd import _build, metaimage, phantoms
    f = phantoms.deshade_phantom((10, 16, 20), seed=4)[0]
    src = str(tmp_path / "in.mha")
    metaimage.write(src, f, spacing=[0.9, 0.9, 2.0])
    outs = [str(tmp_path / n) for n in ("u.mha", "l.mhd", "r.mha")]
    _say("running", _build.CLI)
    r = subprocess.run([_build.CLI, src, "0.5", "2", "1", "40", "-1"] + outs,
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    (u, mu), (l, _), (rr, _) = (metaimage.read(p) for p in outs)
    assert mu["spacing"] == [0.9, 0.9, 2.0]
    tau = float(np.float32(1 / np.sqrt(20)))
    st = orc.tgv2(f, 40, lam=0.5, alpha0=2.0, alpha1=1.0, tau=tau, sigma=tau,
                  data_term="l1", dtype=np.float32)
    lo = orc.integrate_poisson(*st.v, dtype=np.float32)
    ro, _ = orc.deshade_epilogue(st.u, lo, None, False, True, False)
    np.testing.assert_array_equal(u, st.u)
    np.testing.assert_array_equal(l, lo)
    np.testing.assert_array_equal(rr, ro)


def test_dct_extent_beyond_table_budget_is_rejected(ctx):
    """An extent whose cosine tables would exceed the stated budget (16384 per
    axis) returns AMB_ERR_NOT_SUPPORTED instead of allocating them; the context
    stays usable."""
    from ambient_image_processor_amd import native
    a = np.zeros((1, 1, 16385), np.float32)
    with pytest.raises(native.AmbError) as ei:
        ctx.cosine_transform(a)
    assert ei.value.status == 4 and "16384" in str(ei.value)
    b = np.random.default_rng(2).standard_normal((3, 4, 5)).astype(np.float32)
    from oracle import oracle
    np.testing.assert_array_equal(ctx.cosine_transform(b), oracle.cosine_transform(b, dtype=np.float32))
