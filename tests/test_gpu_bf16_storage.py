"""BF16 field storage on the GPU against the oracle's BF16 mode, bit for bit.

The oracle's BF16 mode (FP32 arithmetic, every field store rounded to bfloat16,
nearest even, at the stores of the reference's update kernels:
tgv_cuda/tgv2_common.cu:124-130 p, :262-275 q, :338-347 v and v_bar;
tgv_cuda/tgv2_l2.cu:44-46 u and u_bar) is pinned in tests/test_oracle_bf16.py.
The GPU rounds with gfx950's packed conversion (two values per instruction) in
the vector stores and in the fused sweep's in-register u_bar / v_bar, and with
the scalar conversion in the one-voxel tiles; every launch form is run here.

Workload: the mixed-precision TGV2-L2 deshade configuration's generator (384^3
spheres, bias rescaled to max |b| = 0.2, noise 0.02, lambda 1, alpha0 2, alpha1
1, tau = sigma = 1/sqrt(20)) at sizes the oracle runs in seconds, and a crop of
the 384^3 volume itself; at 384^3, 800 iterations, checks that do not depend on
the size (dual feasibility up to the bfloat16 rounding, the FP64 gap falling).
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TAU = float(1 / np.sqrt(20))
PRM = dict(lam=1.0, alpha0=2.0, alpha1=1.0, tau=TAU, sigma=TAU, theta=1.0)

# launch forms: (name, plan overrides)
FORMS = [
    ("fused_vx2", {}),                                   # default plan, packed rounding
    ("fused_vx4", {"fused_vx": 4}),                      # four-voxel fused lanes
    ("two_sweeps_vec", {"fused": 0}),                    # vector tiles (W % 4 == 0)
    ("two_sweeps_scalar", {"fused": 0, "tile": 0}),      # one voxel per thread
]


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _params(native, iterations, **kw):
    return native.default_params(lam=PRM["lam"], alpha0=PRM["alpha0"], alpha1=PRM["alpha1"],
                                 tau=PRM["tau"], sigma=PRM["sigma"], theta=PRM["theta"],
                                 iterations=iterations, data_term=native.DATA_L2,
                                 storage=native.STORAGE_BF16, **kw)


def _c4_generator(shape):
    from ambient_image_processor_amd import phantoms
    return phantoms.deshade_phantom(shape, seed=4, kind="spheres", noise=0.02, bias_order=3,
                                    bias_sigma=0.05, bias_max=0.2)[0]


@pytest.mark.parametrize("form", FORMS, ids=[f[0] for f in FORMS])
def test_c4_generator_bitwise(ctx, orc, form):
    """64 x 48 x 64 (W % 4 == 0), 150 iterations: all 17 fields equal the oracle's
    BF16 mode bit for bit."""
    from ambient_image_processor_amd import native
    name, ov = form
    ctx.clear_overrides()
    for k, v in ov.items():
        ctx.set_override(k, v)
    shape = (64, 48, 64)
    f = _c4_generator(shape)
    _say("case", name, shape, "150 iterations")
    ctx.begin(f, _params(native, 0))
    ctx.iterate(150)
    got = ctx.get_state()
    ref = orc.tgv2(f, 150, data_term="l2", dtype=np.float32, bf16=True, **PRM)
    for k, fname in enumerate(native.FIELD_NAMES):
        np.testing.assert_array_equal(got[k], ref.fields[k], err_msg=f"{name} {fname}")


def test_c4_generator_odd_width_and_callback(ctx, orc):
    """Odd width (one-voxel tiles, scalar rounding) and a callback every 40
    iterations: u and v seen by the callback and returned equal the oracle."""
    from ambient_image_processor_amd import native
    ctx.clear_overrides()
    f = _c4_generator((20, 33, 45))
    seen = {}

    def cb(it, count, u, vx, vy, vz):
        seen[it] = (u.copy(), vz.copy())
        return False

    _say("case odd width 20x33x45, 120 iterations, callback 40")
    u, vx, vy, vz = ctx.tgv2_deshade(f, _params(native, 120, callback_interval=40), cb)
    # the callback of iteration index k runs after that iteration's updates,
    # i.e. after k + 1 iterations (reference tgv_cuda/tgv2_common.cu:374-408)
    ref = orc.TGVState(f, np.float32, bf16=True)
    assert sorted(seen) == [40, 80]
    done = 0
    for it in (40, 80):
        ref.iterate(it + 1 - done, data_term="l2", **PRM)
        done = it + 1
        np.testing.assert_array_equal(seen[it][0], ref.u)
        np.testing.assert_array_equal(seen[it][1], ref.v[2])
    ref.iterate(120 - done, data_term="l2", **PRM)
    np.testing.assert_array_equal(u, ref.u)
    np.testing.assert_array_equal(np.stack([vx, vy, vz]), ref.v)


@pytest.fixture(scope="module")
def w384():
    from ambient_image_processor_amd import phantoms
    _say("generating the 384^3 workload")
    return phantoms.workload("deshade_384")


def test_c4_crop_pipeline_bitwise(ctx, orc, w384):
    """A 64^3 crop of the 384^3 volume through the whole deshade (BF16 fields,
    reference-numerics DCT): u, shading and deshaded equal the oracle's BF16
    loop followed by its FP32 Poisson solve and epilogue, bit for bit."""
    from ambient_image_processor_amd import native
    f = np.ascontiguousarray(w384["f"][160:224, 160:224, 160:224])
    _say("case 384^3 crop 64^3, 200 iterations, pipeline")
    u, l, r = ctx.deshade_pipeline(f, _params(native, 200, flags=native.FLAG_SET_NEGATIVE_ZERO))
    st = orc.tgv2(f, 200, data_term="l2", dtype=np.float32, bf16=True, **PRM)
    lo = orc.integrate_poisson(*st.v, dtype=np.float32)
    ro, _ = orc.deshade_epilogue(st.u, lo, None, False, True, False)
    np.testing.assert_array_equal(u, st.u)
    np.testing.assert_array_equal(l, lo)
    np.testing.assert_array_equal(r, ro)


def test_c4_full_size_feasible_and_gap_falls(ctx, w384):
    """384^3, 800 iterations in BF16 storage: every p and q lies in its ball up
    to the bfloat16 rounding of the stored components (relative 2^-8), and the
    FP64 primal-dual gap of the stored state falls."""
    from ambient_image_processor_amd import native
    f = w384["f"]
    ctx.begin(f, _params(native, 0))
    gaps = []
    done = 0
    for n in (50, 200, 800):
        ctx.iterate(n - done)
        done = n
        e = ctx.energy()
        _say(f"384^3 BF16 after {n} iterations: E {e[0]:.6e} gap {e[4]:.6e}")
        assert np.all(np.isfinite(e))
        gaps.append(e[4])
    assert gaps[2] < gaps[1] < gaps[0]
    st = ctx.get_state()
    bound = 1.0 + 2.0 ** -7
    pn = np.sqrt(st[8].astype(np.float64) ** 2 + st[9].astype(np.float64) ** 2 +
                 st[10].astype(np.float64) ** 2)
    assert pn.max() <= PRM["alpha1"] * bound
    del pn
    q = st[11:17].astype(np.float64)
    qn = np.sqrt(q[0] ** 2 + q[1] ** 2 + q[2] ** 2 + 2 * (q[3] ** 2 + q[4] ** 2 + q[5] ** 2))
    assert qn.max() <= PRM["alpha0"] * bound
