"""GPU tests of the hand-written adjoint of the unrolled TGV2-L2 denoise
(amb_tgv2_denoise_adjoint) against the FP64 CPU adjoint (oracle/adjoint.py)
and against central finite differences (eps = 1e-3) of the FP64 oracle."""
import sys

import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu

TAU = float(np.float32(1 / np.sqrt(20)))


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _vol(shape, seed):
    from ambient_image_processor_amd import phantoms
    rng = phantoms.SplitMix64(seed)
    clean = phantoms._boxes(rng, shape, 6, max(shape[2] - 6, 1), (3, max(4, shape[2] // 2)),
                            (0.1, 1.0), 0.0)
    f = clean + rng.normal(int(np.prod(shape)), 0.05).reshape(shape)
    return f.astype(np.float32), clean.astype(np.float32)


def _params(native, iters, alpha0, alpha1, lam):
    return native.default_params(lam=lam, alpha0=alpha0, alpha1=alpha1, tau=TAU, sigma=TAU,
                                 theta=1.0, iterations=iters, data_term=native.DATA_L2)


@pytest.mark.parametrize("shape", [(10, 14, 18), (9, 11, 19)])   # two / one voxel per thread
@pytest.mark.parametrize("alphas", [(2.0, 1.0), (0.02, 0.01)])
def test_adjoint_matches_fp64_cpu_adjoint(ctx, alphas, shape):
    from ambient_image_processor_amd import native
    from oracle import adjoint as adj
    f, clean = _vol(shape, 5)
    a0, a1 = alphas
    p = _params(native, 40, a0, a1, 1.0)
    n = f.size
    u_ref = adj.forward(f, 40, a0, a1, 1.0, TAU, TAU)["u"]
    g = (2.0 / n) * (u_ref - clean)
    u, gf, gp = ctx.tgv2_denoise_adjoint(f, p, g.astype(np.float32), checkpoint_every=7)
    _, ref = adj.adjoint(f.astype(np.float64), 40, a0, a1, 1.0, TAU, TAU, 1.0, g)
    _say("dL/df rel", rel_err(gf, ref["f"]), "params", gp,
         [ref["alpha0"], ref["alpha1"], ref["lam"]])
    assert rel_err(u, u_ref) < 1e-5
    assert rel_err(gf, ref["f"]) < 2e-3
    for k, name in enumerate(("alpha0", "alpha1", "lam")):
        assert abs(gp[k] - ref[name]) <= 2e-3 * max(abs(ref[name]), 1e-9 + 1e-3 * np.abs(ref["f"]).sum()), name


@pytest.mark.parametrize("k", [1, 20])
def test_adjoint_of_zero_iterations_is_the_identity(ctx, k):
    """No iterations: u = f, dL/df = dL/du, no parameter gradient."""
    from ambient_image_processor_amd import native
    f, _ = _vol((6, 8, 10), 4)
    g = np.random.default_rng(5).standard_normal(f.shape).astype(np.float32)
    u, gf, gp = ctx.tgv2_denoise_adjoint(f, _params(native, 0, 2.0, 1.0, 1.0), g, checkpoint_every=k)
    np.testing.assert_array_equal(u, f)
    np.testing.assert_array_equal(gf, g)
    np.testing.assert_array_equal(gp, np.zeros(3))


@pytest.mark.parametrize("shape", [(8, 12, 20), (7, 9, 21)])   # two / one voxel per thread
def test_adjoint_independent_of_checkpoint_interval(ctx, shape):
    from ambient_image_processor_amd import native
    f, clean = _vol(shape, 6)
    p = _params(native, 23, 0.05, 0.02, 1.0)
    g = np.random.default_rng(1).standard_normal(f.shape).astype(np.float32)
    outs = [ctx.tgv2_denoise_adjoint(f, p, g, checkpoint_every=k) for k in (1, 5, 23, 40)]
    for u, gf, gp in outs[1:]:
        np.testing.assert_array_equal(u, outs[0][0])
        np.testing.assert_array_equal(gf, outs[0][1])
        np.testing.assert_array_equal(gp, outs[0][2])


@pytest.mark.parametrize("shape,form", [
    ((8, 12, 20), {}), ((8, 12, 20), {"fused_pf": 0}), ((7, 10, 40), {"fused_tx": 12}),
    ((7, 10, 40), {"fused_tx": 12, "fused_pf": 0}), ((6, 9, 34), {"fused_ty": 3}),
    ((7, 10, 41), {}), ((5, 6, 301), {}), ((6, 7, 45), {"fused_tx": 13})])
def test_adjoint_recompute_forms_agree(ctx, shape, form):
    """The reverse pass recomputes each segment with the fused sweep out of
    place (u_bar, v_bar stored too); with the fused sweep off it copies the
    state and runs the two sweeps in place.  Both give the same states, so u,
    dL/df and the scalar gradients agree bit for bit.  Odd widths take the
    one-voxel form: whole rows (41), 126-column tiles (301), forced tiles (45)."""
    from ambient_image_processor_amd import native
    f, clean = _vol(shape, 8)
    p = _params(native, 17, 0.05, 0.02, 1.0)
    g = np.random.default_rng(2).standard_normal(f.shape).astype(np.float32)
    _say("recompute forms", shape, form)
    for k, v in form.items():
        ctx.set_override(k, v)
    a = ctx.tgv2_denoise_adjoint(f, p, g, checkpoint_every=6)
    ctx.clear_overrides()
    ctx.set_override("fused", 0)
    b = ctx.tgv2_denoise_adjoint(f, p, g, checkpoint_every=6)
    for x, y in zip(a, b):
        np.testing.assert_array_equal(x, y)


@pytest.mark.parametrize("shape", [(8, 12, 20), (6, 9, 130), (5, 7, 600)])
def test_adjoint_two_voxels_per_thread_equals_one(ctx, shape):
    """On even widths the reverse phases run two voxels per thread (paired
    loads); the plan override adjoint_vx = 1 forces one.  Each voxel's
    arithmetic is the same, so u and dL/df agree bit for bit; the scalar
    gradients are FP64 sums over differently sized blocks and agree to FP64
    rounding.  Widths: one row piece, a row piece with idle lanes, several
    row pieces per row."""
    from ambient_image_processor_amd import native
    f, clean = _vol(shape, 9)
    p = _params(native, 13, 0.05, 0.02, 1.0)
    g = np.random.default_rng(3).standard_normal(f.shape).astype(np.float32)
    _say("adjoint voxels per thread", shape)
    a = ctx.tgv2_denoise_adjoint(f, p, g, checkpoint_every=5)
    ctx.set_override("adjoint_vx", 1)
    b = ctx.tgv2_denoise_adjoint(f, p, g, checkpoint_every=5)
    np.testing.assert_array_equal(a[0], b[0])
    np.testing.assert_array_equal(a[1], b[1])
    np.testing.assert_allclose(a[2], b[2], rtol=1e-12, atol=1e-15)


@pytest.mark.parametrize("alphas", [(2.0, 1.0), (0.02, 0.01)])
def test_adjoint_against_central_differences(ctx, orc, alphas):
    """32^3 volume, 200 iterations, checkpoints every 20, MSE loss to the clean
    phantom; central differences of the FP64 oracle forward.

    * Directional derivative along a seeded unit direction: FD at eps = 1e-3
      agrees to 1e-3.
    * Scalars: the loss is piecewise smooth (projection kinks).  Once a p or q
      has been projected it sits on the ball |.| = alpha and later iterates
      hover there, so FD steps of 1e-6..1e-3 cross kinks; FD converges to the
      adjoint only as eps -> 1e-8 (FP64).  Checked: converged FD (eps = 1e-8)
      to 1e-3 in both cases; FD at eps = 1e-3 to 1e-2 (the kink-crossing
      spread) with alpha0 = 2, alpha1 = 1, where eps is 1e-3 of the scalar.
      With alpha0 = 0.02 and alpha1 = 0.01 a step of 1e-3 changes the scalar by
      5-10 % and crosses many kinks (FD(1e-3) of dL/dalpha0 is then off by
      ~75 %), so only the converged FD is compared there.
      With alpha0 = 2 the q projection never activates and dL/dalpha0 is 0;
      in the second case many q's hover at the ball, where the FP32 forward's
      active set can differ from the FP64 one, so dL/dalpha0 is held to 2e-2."""
    from ambient_image_processor_amd import native
    f, clean = _vol((32, 32, 32), 7)
    a0, a1 = alphas
    lam, iters = 1.0, 200
    n = f.size
    f64 = f.astype(np.float64)

    def loss(ff, a0_, a1_, lam_):
        st = orc.tgv2(ff, iters, lam=lam_, alpha0=a0_, alpha1=a1_, tau=TAU, sigma=TAU,
                      data_term="l2", dtype=np.float64)
        return float(np.mean((st.u - clean) ** 2)), st.u

    L0, u64 = loss(f64, a0, a1, lam)
    p = _params(native, iters, a0, a1, lam)
    u, gf, gp = ctx.tgv2_denoise_adjoint(f, p, ((2.0 / n) * (u64 - clean)).astype(np.float32), 20)

    def fd_scalar(k, eps):
        plus, minus = [a0, a1, lam], [a0, a1, lam]
        plus[k] += eps
        minus[k] -= eps
        return (loss(f64, *plus)[0] - loss(f64, *minus)[0]) / (2 * eps)

    d = np.random.default_rng(3).standard_normal(f.shape)
    d /= np.linalg.norm(d)
    eps = 1e-3
    fdf = (loss(f64 + eps * d, a0, a1, lam)[0] - loss(f64 - eps * d, a0, a1, lam)[0]) / (2 * eps)
    gd = float(np.sum(gf.astype(np.float64) * d))
    _say(f"alphas {alphas}: directional adjoint {gd:.6e} fd(1e-3) {fdf:.6e}")
    assert abs(gd - fdf) <= 1e-3 * abs(fdf)
    scale = max(abs(x) for x in gp)
    for k, name in enumerate(("alpha0", "alpha1", "lambda")):
        fd3, fd8 = fd_scalar(k, 1e-3), fd_scalar(k, 1e-8)
        _say(f"  d/d{name}: adjoint {gp[k]:.6e} fd(1e-3) {fd3:.6e} fd(1e-8) {fd8:.6e}")
        tol = 2e-2 if (k == 0 and a0 < 1) else 1e-3
        assert abs(gp[k] - fd8) <= tol * abs(fd8) + 1e-6 * scale, name
        if a0 >= 1:
            assert abs(gp[k] - fd3) <= 1e-2 * abs(fd3) + 1e-6 * scale, name
    if a0 < 1:
        assert abs(gp[0]) > 1e-6 * scale      # the q projection is active


def test_adjoint_rejects_l1_and_bf16(ctx):
    from ambient_image_processor_amd import native
    f = np.zeros((4, 4, 4), np.float32)
    p = _params(native, 3, 2.0, 1.0, 1.0)
    p.data_term = native.DATA_L1
    with pytest.raises(native.AmbError):
        ctx.tgv2_denoise_adjoint(f, p, f)
    p = _params(native, 3, 2.0, 1.0, 1.0)
    p.storage = native.STORAGE_BF16
    with pytest.raises(native.AmbError):
        ctx.tgv2_denoise_adjoint(f, p, f)
