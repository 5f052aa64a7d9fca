"""CPU tests that pin the oracle's TGV-k (k >= 3) deshade loop
(oracle/tgv_oracle.cpp, reference tgvk_deshade/tgvk_l1_deshade.cu) before it
checks the GPU path: the level operators against dense matrices, the loop
against an independent vectorised restatement for k = 3, 4, 5 (odd and even
levels, the last-level rule, q~ = r_{i-1}), invariants, projections, and
the FP32 mirror against FP64."""
import numpy as np
import pytest

from conftest import rel_err

TAU = 1 / np.sqrt(20)
KPRM = dict(lam=0.7, tau=TAU, sigma=TAU, theta=1.0)
AXES = [(2, None), (1, None), (0, None), (2, 1), (2, 0), (1, 0)]   # numpy axes per component


def _dplus(a, ax):
    out = np.zeros_like(a)
    sl, sl2 = [slice(None)] * a.ndim, [slice(None)] * a.ndim
    sl[ax], sl2[ax] = slice(0, -1), slice(1, None)
    out[tuple(sl)] = a[tuple(sl2)] - a[tuple(sl)]
    return out


def _bminus(a, ax):
    out = -a.copy()
    sl, sl2 = [slice(None)] * a.ndim, [slice(None)] * a.ndim
    sl[ax], sl2[ax] = slice(1, None), slice(0, -1)
    out[tuple(sl)] += a[tuple(sl2)]
    return out


def _phantom(shape, seed):
    rng = np.random.default_rng(seed)
    z, y, x = np.indices(shape)
    return (0.4 + 0.1 * np.sin(0.5 * x) * np.cos(0.3 * y + 0.2 * z)
            + 0.05 * rng.standard_normal(shape))


# ---------------------------------------------------------------------------
# level operators G (halve) and H against dense difference matrices
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("forward", [False, True])
@pytest.mark.parametrize("halve", [False, True])
def test_level_operators_match_dense(orc, forward, halve):
    shape = (3, 4, 5)
    n = int(np.prod(shape))
    mats = []
    for axis in range(3):   # oracle axis 0 = x
        M = np.zeros((n, n))
        for j in range(n):
            e = np.zeros(n)
            e[j] = 1.0
            op = orc.forward_difference if forward else orc.backward_difference
            M[:, j] = op(e.reshape(shape), axis).ravel()
        mats.append(M)
    comp_axes = [(0,), (1,), (2,), (0, 1), (0, 2), (1, 2)]   # x, y, z, xy, xz, yz
    t = np.random.default_rng(3).standard_normal((6,) + shape)
    got = orc.tensor_op(t, forward, halve)
    for c, axes in enumerate(comp_axes):
        ref = sum(mats[a] @ t[c].ravel() for a in axes)
        if halve and len(axes) == 2:
            ref = ref * 0.5
        np.testing.assert_allclose(got[c].ravel(), ref, rtol=0, atol=1e-15)


# ---------------------------------------------------------------------------
# the loop against an independent restatement
# ---------------------------------------------------------------------------
def _numpy_tgvk(f, iters, lam, alpha, tau, sigma, theta, mask=None):
    """tgvk_l1_deshade.cu:152-351 written with numpy slicing: the TGV2 part
    (p with alpha[0], u L1 prox, q with alpha[1] shifted by w_bar_0, v), then
    per level i: r_i from G(w_bar_i) (B for even i, D+ for odd i) minus
    w_bar_{i+1} (not on the last level), projection with alpha[i+2], w_i from
    H(r_i) (D+ for even i, B for odd i) minus q~ (q, or r_{i-1}).  With a mask
    (tgvk_l1_deshade_masked.cu) every update keeps its old value where the
    mask is below 1."""
    act = np.ones(f.shape, bool) if mask is None else np.asarray(mask) >= 1

    def upd(new, old):
        return np.where(act, new, old)

    X, Y, Z = 2, 1, 0
    L = len(alpha) - 2
    u = f.copy(); ub = f.copy()
    v = np.zeros((3,) + f.shape); vb = v.copy(); p = v.copy()
    q = np.zeros((6,) + f.shape)
    w = np.zeros((L, 6) + f.shape); wb = w.copy(); r = w.copy()

    def wnorm(t):
        return np.sqrt(t[0] ** 2 + t[1] ** 2 + t[2] ** 2 + 2 * (t[3] ** 2 + t[4] ** 2 + t[5] ** 2))

    def G(t, fwd):
        d = _dplus if fwd else _bminus
        out = [d(t[c], a) if b is None else 0.5 * (d(t[c], a) + d(t[c], b))
               for c, (a, b) in enumerate(AXES)]
        return np.stack(out)

    def H(t, fwd):
        d = _dplus if fwd else _bminus
        out = [d(t[c], a) if b is None else d(t[c], a) + d(t[c], b)
               for c, (a, b) in enumerate(AXES)]
        return np.stack(out)

    for _ in range(iters):
        u_old, v_old = u.copy(), v.copy()
        pn = p + sigma * (np.stack([_dplus(ub, X), _dplus(ub, Y), _dplus(ub, Z)]) - vb)
        p = upd(pn / np.maximum(1, np.sqrt((pn ** 2).sum(0)) / alpha[0]), p)
        un = u - tau * (_bminus(p[0], X) + _bminus(p[1], Y) + _bminus(p[2], Z))
        d = un - f
        un = np.where(d > tau * lam, un - tau * lam, np.where(d < -tau * lam, un + tau * lam, f))
        u = upd(un, u)
        ub = upd(u + theta * (u - u_old), ub)
        e = np.stack([_bminus(vb[0], X), _bminus(vb[1], Y), _bminus(vb[2], Z),
                      0.5 * (_bminus(vb[1], X) + _bminus(vb[0], Y)),
                      0.5 * (_bminus(vb[2], X) + _bminus(vb[0], Z)),
                      0.5 * (_bminus(vb[2], Y) + _bminus(vb[1], Z))])
        qn = q + sigma * (e - wb[0])
        q = upd(qn / np.maximum(1, wnorm(qn) / alpha[1]), q)
        div = np.stack([_dplus(q[0], X) + _dplus(q[3], Y) + _dplus(q[4], Z),
                        _dplus(q[1], Y) + _dplus(q[3], X) + _dplus(q[5], Z),
                        _dplus(q[2], Z) + _dplus(q[4], X) + _dplus(q[5], Y)])
        v = upd(v - tau * (div - p), v)
        vb = upd(v + theta * (v - v_old), vb)
        wb_old = wb.copy()
        for i in range(L):
            odd = i % 2 == 1
            g = G(wb_old[i], fwd=odd)
            if i < L - 1:
                g = g - wb_old[i + 1]
            rn = r[i] + sigma * g
            r[i] = upd(rn / np.maximum(1, wnorm(rn) / alpha[i + 2]), r[i])
            qt = q if i == 0 else r[i - 1]
            w_old = w[i].copy()
            w[i] = upd(w[i] - tau * (H(r[i], fwd=not odd) - qt), w[i])
            wb[i] = upd(w[i] + theta * (w[i] - w_old), wb[i])
    return u, v, p, q, w, wb, r


@pytest.mark.parametrize("alpha", [(1.0, 2.0, 0.6), (1.0, 2.0, 0.8, 0.5),
                                   (0.8, 1.5, 0.7, 0.4, 0.3)])
def test_tgvk_matches_vectorised_restatement(orc, alpha):
    f = _phantom((5, 6, 7), 1)
    st = orc.tgvk(f, 12, alpha=alpha, **KPRM)
    u, v, p, q, w, wb, r = _numpy_tgvk(f, 12, alpha=alpha, **KPRM)
    assert rel_err(st.u, u) < 1e-13
    assert rel_err(st.v, v) < 1e-12
    assert rel_err(st.p, p) < 1e-12
    assert rel_err(st.q, q) < 1e-12
    for i in range(len(alpha) - 2):
        sw, swb, sr = st.level(i)
        assert rel_err(sw, w[i]) < 1e-12, i
        assert rel_err(swb, wb[i]) < 1e-12, i
        assert rel_err(sr, r[i]) < 1e-12, i
    # the levels move: the check is not vacuous
    assert np.abs(st.level(len(alpha) - 3)[2]).max() > 1e-3


@pytest.mark.parametrize("alpha", [(1.0, 2.0, 0.6), (0.8, 1.5, 0.7, 0.4, 0.3)])
def test_masked_tgvk_matches_vectorised_restatement(orc, alpha):
    """The masked loop: updates only where mask >= 1 (fractional values below 1
    count as outside), stencils read the unmasked neighbours' initial values."""
    f = _phantom((6, 7, 9), 8)
    rng = np.random.default_rng(9)
    m = np.where(rng.random(f.shape) > 0.35, 1.0, rng.random(f.shape) * 0.999)
    st = orc.tgvk(f, 12, alpha=alpha, mask=m, **KPRM)
    u, v, p, q, w, wb, r = _numpy_tgvk(f, 12, alpha=alpha, mask=m, **KPRM)
    assert rel_err(st.u, u) < 1e-13
    assert rel_err(st.v, v) < 1e-12
    assert rel_err(st.q, q) < 1e-12
    for i in range(len(alpha) - 2):
        assert rel_err(st.level(i)[0], w[i]) < 1e-12, i
        assert rel_err(st.level(i)[2], r[i]) < 1e-12, i
    out = m < 1
    np.testing.assert_array_equal(st.u[out], f[out])
    assert not np.any(st.v[:, out]) and not np.any(st.level(0)[2][:, out])
    unmasked = orc.tgvk(f, 12, alpha=alpha, **KPRM)
    assert rel_err(st.u, unmasked.u) > 1e-4


def test_tgvk_planar_matches_3d_restatement_without_z(orc):
    """The 2-D TGV-k loop (tgvk_l1_deshade_2d.cu) at depth 1: the 3-D loop with
    the z backward differences of eps_B(v_bar) removed; every z component
    stays 0."""
    f = _phantom((1, 11, 13), 2)
    alpha = (1.0, 2.0, 0.7, 0.5)
    st = orc.tgvk(f, 15, alpha=alpha, planar=True, **KPRM)
    st3 = orc.tgvk(f, 15, alpha=alpha, **KPRM)
    assert not np.any(st.v[2]) and not np.any(st.q[2]) and not np.any(st.q[4:])
    for i in range(2):
        for t in st.level(i):
            assert not np.any(t[2]) and not np.any(t[4:])
    assert np.abs(st3.v[2]).max() > 1e-4
    with pytest.raises(ValueError):
        orc.tgvk(np.zeros((2, 3, 4)), 1, alpha=alpha, planar=True, **KPRM)


# ---------------------------------------------------------------------------
# invariants
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_tgvk_constant_image_is_fixed_point(orc, dtype):
    f = np.full((4, 5, 6), 0.37)
    st = orc.tgvk(f, 20, alpha=(1.0, 2.0, 0.5, 0.5), dtype=dtype, **KPRM)
    np.testing.assert_array_equal(st.u, np.asarray(f, dtype))
    assert not np.any(st.fields[2:])


def test_tgvk_zero_iterations_is_the_initial_state(orc):
    f = _phantom((3, 4, 5), 3)
    st = orc.tgvk(f, 0, alpha=(1.0, 2.0, 0.5), **KPRM)
    np.testing.assert_array_equal(st.u, f)
    np.testing.assert_array_equal(st.fields[1], f)
    assert not np.any(st.fields[2:])


def test_tgvk_projections_hold_and_bind(orc):
    """|p| <= alpha[0], |q|_w <= alpha[1], |r_i|_w <= alpha[i+2] after every
    iteration, and each bound is active somewhere."""
    f = _phantom((8, 9, 10), 4) * 3
    alpha = (0.05, 0.03, 0.02, 0.01)
    st = orc.TGVKState(f, 4)
    wn = lambda t: np.sqrt(t[0] ** 2 + t[1] ** 2 + t[2] ** 2            # noqa: E731
                           + 2 * (t[3] ** 2 + t[4] ** 2 + t[5] ** 2))
    bound = np.zeros(4, bool)
    for _ in range(4):
        st.iterate(10, alpha=alpha, **KPRM)
        norms = [np.sqrt((st.p ** 2).sum(0)), wn(st.q), wn(st.level(0)[2]), wn(st.level(1)[2])]
        for k, (a, nrm) in enumerate(zip(alpha, norms)):
            assert nrm.max() <= a * (1 + 1e-12)
            bound[k] |= np.mean(nrm > 0.999 * a) > 0.001
    assert bound.all(), bound


def test_tgvk_fp32_mirror_tracks_fp64(orc):
    f = _phantom((10, 12, 14), 5)
    alpha = (1.0, 2.0, 0.6, 0.4)
    a = orc.tgvk(f, 100, alpha=alpha, dtype=np.float64, **KPRM)
    b = orc.tgvk(f, 100, alpha=alpha, dtype=np.float32, **KPRM)
    assert rel_err(b.u, a.u) < 1e-5
    assert rel_err(b.v, a.v) < 1e-4


def test_deshade_k_order_2_is_tgv2_with_swapped_alphas(orc):
    f = _phantom((4, 6, 8), 6).astype(np.float32)
    u, l, r, _ = orc.deshade_k(f, 20, alpha=(1.0, 2.0), **KPRM)
    st = orc.tgv2(f, 20, alpha0=2.0, alpha1=1.0, data_term="l1", dtype=np.float32, **KPRM)
    np.testing.assert_array_equal(u, st.u)
    with pytest.raises(ValueError):
        orc.TGVKState(f, 2)
