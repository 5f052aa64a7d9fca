"""GPU tests at the benchmark configurations' full sizes, through checks whose
meaning does not depend on the size: the first iterations against the FP32
oracle bit for bit (TGV2, TGV-k, and the one-voxel fused sweep at odd sizes
next to the benchmark's 256^3), schedule and masking equivalences bit for
bit, the dual feasibility and the decrease of the FP64 primal-dual gap over the
iterations,
the Poisson solve against its closed form, and the adjoint against central
differences.  (Thousand-iteration parity with the oracle at 256^3 would take
the CPU minutes; the same loop is checked over many iterations at smaller
sizes in the other test modules.)"""
import sys

import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def w256():
    from ambient_image_processor_amd import phantoms
    _say("generating the 256^3 workload")
    return phantoms.workload("deshade_256")


def _p256(native, w, iterations, **kw):
    return native.default_params(lam=w["lam"], alpha0=w["alpha0"], alpha1=w["alpha1"],
                                 tau=w["tau"], sigma=w["tau"], theta=1.0, iterations=iterations,
                                 data_term=native.DATA_L1, **kw)


def test_256_first_iterations_bitwise(ctx, orc, w256):
    """TGV2-L1 at 256^3: three iterations (fused sweep) equal the FP32 oracle."""
    from ambient_image_processor_amd import native
    f = w256["f"]
    u, vx, vy, vz = ctx.tgv2_deshade(f, _p256(native, w256, 3))
    st = orc.tgv2(f, 3, lam=w256["lam"], alpha0=w256["alpha0"], alpha1=w256["alpha1"],
                  tau=w256["tau"], sigma=w256["tau"], data_term="l1", dtype=np.float32)
    np.testing.assert_array_equal(u, st.u)
    np.testing.assert_array_equal(vz, st.v[2])


def test_256_schedules_and_all_ones_mask_agree(ctx, monkeypatch, w256):
    """12 iterations at 256^3: fused == two sweeps == masked loop with every
    voxel active, bit for bit."""
    from ambient_image_processor_amd import native
    f = w256["f"]
    p = _p256(native, w256, 12)
    fused = ctx.tgv2_deshade(f, p)
    masked = ctx.tgv2_deshade_masked(f, p, np.ones(f.shape, np.uint8))
    ctx.set_override("fused", 0)
    two = ctx.tgv2_deshade(f, p)
    for a, b, c in zip(fused, two, masked):
        np.testing.assert_array_equal(a, b)
        np.testing.assert_array_equal(a, c)


def test_256_gap_decreases_and_duals_stay_feasible(ctx, w256):
    from ambient_image_processor_amd import native
    f = w256["f"]
    ctx.begin(f, _p256(native, w256, 0))
    gaps = []
    for n in (25, 75, 150):
        ctx.iterate(n)
        e = ctx.energy()
        _say("gap", e[4], "brackets", e[5:8])
        assert e[6] >= -1e-6 * abs(e[0]) and e[7] >= -1e-6 * abs(e[0])
        gaps.append(e[4])
    assert gaps[2] < gaps[1] < gaps[0]
    st = ctx.get_state()
    pn = np.sqrt(st[8].astype(np.float64) ** 2 + st[9] ** 2 + st[10] ** 2)
    q = st[11:17].astype(np.float64)
    qn = np.sqrt(q[0] ** 2 + q[1] ** 2 + q[2] ** 2 + 2 * (q[3] ** 2 + q[4] ** 2 + q[5] ** 2))
    assert pn.max() <= w256["alpha1"] * (1 + 1e-6)
    assert qn.max() <= w256["alpha0"] * (1 + 1e-6)


def test_tgvk3_256_fused_bitwise(ctx, orc, monkeypatch, w256):
    """The TGV-k (k = 3) configuration at 256^3: three iterations (dual, two
    fused sweeps, primal) equal the FP32 oracle bit for bit, and 12 iterations
    of the fused loop equal the two-launch loop's."""
    from ambient_image_processor_amd import native
    f = w256["f"]
    alpha = (1.0, 4.0, 7.0)
    tau = float(np.float32(1 / np.sqrt(20)))

    def params(it):
        return native.default_params(lam=0.5, tau=tau, sigma=tau, theta=1.0, iterations=it,
                                     data_term=native.DATA_L1)

    u, vx, vy, vz = ctx.tgvk_deshade(f, params(3), alpha)
    _say("FP32 oracle, 3 TGV-k iterations at 256^3")
    st = orc.tgvk(f, 3, alpha=alpha, lam=0.5, tau=tau, sigma=tau, theta=1.0, dtype=np.float32)
    np.testing.assert_array_equal(u, st.u)
    np.testing.assert_array_equal(vx, st.v[0])
    np.testing.assert_array_equal(vz, st.v[2])
    fused = ctx.tgvk_deshade(f, params(12), alpha)
    ctx.set_override("tgvk_fused", 0)
    two = ctx.tgvk_deshade(f, params(12), alpha)
    for a, b in zip(fused, two):
        np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("precision,tol", [(1, 1e-5), (2, 1e-4)])
def test_poisson_512x512x256_known_answer(ctx, precision, tol):
    """The standalone Poisson configuration at full size: DCT, eigenvalue divide
    and inverse DCT of the Neumann Laplacian of 64 cosine modes return the
    field (MFMA FP32 and BF16x3 passes)."""
    from ambient_image_processor_amd import phantoms
    _say("generating the 512x512x256 Poisson field")
    w = phantoms.workload("poisson_512")
    rhs, phi = w["rhs"], w["phi"]
    X = ctx.cosine_transform(rhs, precision)
    sol = ctx.inverse_cosine_transform(ctx.solve_poisson(X), precision)
    e = rel_err(sol, phi - phi.mean())
    _say(f"Poisson 512x512x256 precision {precision}: rel err {e:.2e}")
    assert e < tol


def test_384_bf16_storage_schedules_agree(ctx, monkeypatch):
    """The BF16-storage configuration at 384^3 (TGV2-L2): fused == two sweeps,
    bit for bit, over 6 iterations."""
    from ambient_image_processor_amd import native, phantoms
    w = phantoms.workload("deshade_384")
    p = native.default_params(lam=w["lam"], alpha0=w["alpha0"], alpha1=w["alpha1"],
                              tau=w["tau"], sigma=w["tau"], theta=1.0, iterations=6,
                              data_term=native.DATA_L2, storage=native.STORAGE_BF16)
    fused = ctx.tgv2_deshade(w["f"], p)
    ctx.set_override("fused", 0)
    two = ctx.tgv2_deshade(w["f"], p)
    for a, b in zip(fused, two):
        np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("n", [255, 257])
def test_odd_width_fused_sweep_full_size(ctx, orc, n):
    """A sphere phantom at 255^3 (one voxel per thread, whole rows) and 257^3
    (126-column x tiles): three FP32 iterations equal the FP32 oracle, and over
    6 iterations the fused sweep equals the two-sweep loop bit for bit, FP32
    and BF16 fields."""
    from ambient_image_processor_amd import native, phantoms
    f = phantoms.deshade_phantom((n, n, n), seed=2)[0]
    tau = float(np.float32(1 / np.sqrt(20)))
    p3 = native.default_params(lam=0.5, alpha0=2.0, alpha1=1.0, tau=tau, sigma=tau, theta=1.0,
                               iterations=3, data_term=native.DATA_L1)
    u, vx, vy, vz = ctx.tgv2_deshade(f, p3)
    st = orc.tgv2(f, 3, lam=0.5, alpha0=2.0, alpha1=1.0, tau=tau, sigma=tau, data_term="l1",
                  dtype=np.float32)
    np.testing.assert_array_equal(u, st.u)
    np.testing.assert_array_equal(vx, st.v[0])
    for storage in (native.STORAGE_F32, native.STORAGE_BF16):
        _say("odd width", n, "storage", storage)
        p = native.default_params(lam=0.5, alpha0=2.0, alpha1=1.0, tau=tau, sigma=tau, theta=1.0,
                                  iterations=6, data_term=native.DATA_L1, storage=storage)
        ctx.clear_overrides()
        ctx.profile(True)
        ctx.profile_reset()
        fused = ctx.tgv2_deshade(f, p)
        prof = ctx.profile_read()
        ctx.profile(False)
        assert prof["fused_sweep"][1] == 5, prof
        ctx.set_override("fused", 0)
        two = ctx.tgv2_deshade(f, p)
        ctx.clear_overrides()
        for a, b in zip(fused, two):
            np.testing.assert_array_equal(a, b)


def test_adjoint_128_against_central_differences(ctx, orc):
    """The differentiable configuration at full size (128^3, 200 unrolled
    TGV2-L2 iterations, checkpoints every 20, MSE to the clean phantom): the
    adjoint's directional derivative along a seeded unit direction agrees with
    central differences (eps = 1e-3) of the FP64 oracle's loss to 1e-3.  (Along
    a random unit direction the loss moves by ~1e-10, below the FP32 forward's
    rounding noise, so the differences use the FP64 forward, as at 32^3.)"""
    from ambient_image_processor_amd import native, phantoms
    w = phantoms.workload("denoise_128x4")
    f, clean = w["f"][0], w["clean"][0]
    n = f.size
    prm = dict(lam=w["lam"], alpha0=w["alpha0"], alpha1=w["alpha1"], tau=w["tau"],
               sigma=w["tau"])
    p = native.default_params(theta=1.0, iterations=w["iterations"], data_term=native.DATA_L2,
                              **prm)

    def loss(ff):
        st = orc.tgv2(ff, w["iterations"], data_term="l2", dtype=np.float64, **prm)
        return float(np.mean((st.u - clean) ** 2)), st.u

    f64 = f.astype(np.float64)
    _say("FP64 forward at 128^3")
    L0, u64 = loss(f64)
    u, gf, gp = ctx.tgv2_denoise_adjoint(f, p, ((2.0 / n) * (u64 - clean)).astype(np.float32),
                                         w["checkpoint_every"])
    assert rel_err(u, u64) < 1e-5
    d = np.random.default_rng(3).standard_normal(f.shape)
    d /= np.linalg.norm(d)
    eps = 1e-3
    fd = (loss(f64 + eps * d)[0] - loss(f64 - eps * d)[0]) / (2 * eps)
    gd = float(np.sum(gf.astype(np.float64) * d))
    _say(f"128^3 directional derivative: adjoint {gd:.6e} fd(1e-3) {fd:.6e}")
    assert abs(gd - fd) <= 1e-3 * abs(fd)
