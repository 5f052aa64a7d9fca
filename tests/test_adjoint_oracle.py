"""CPU tests pinning the FP64 adjoint (oracle/adjoint.py): forward map against
the C++ oracle, operator transposes against dense matrices, and gradients
against central finite differences."""
import numpy as np
import pytest

from oracle import adjoint as adj

PRM = dict(alpha0=0.005, alpha1=0.02, lam=1.0, tau=1 / np.sqrt(20), sigma=1 / np.sqrt(20),
           theta=1.0)


def _f(shape=(5, 6, 7), seed=3):
    rng = np.random.default_rng(seed)
    base = np.zeros(shape)
    base[:, : shape[1] // 2, :] = 0.6
    return base + 0.05 * rng.standard_normal(shape)


def test_forward_matches_cpp_oracle(orc):
    f = _f()
    s = adj.forward(f, 12, **PRM)
    ref = orc.tgv2(f, 12, lam=PRM["lam"], alpha0=PRM["alpha0"], alpha1=PRM["alpha1"],
                   tau=PRM["tau"], sigma=PRM["sigma"], theta=1.0, data_term="l2")
    assert np.max(np.abs(s["u"] - ref.u)) < 1e-13
    assert np.max(np.abs(s["q"] - ref.q)) < 1e-12


@pytest.mark.parametrize("pair", [("dplus", "dplus_t"), ("bm", "bm_t")])
def test_scalar_transposes(pair):
    shape = (3, 4, 5)
    rng = np.random.default_rng(0)
    fwd, tr = getattr(adj, pair[0]), getattr(adj, pair[1])
    for ax in range(3):
        a, b = rng.standard_normal(shape), rng.standard_normal(shape)
        assert abs(np.sum(fwd(a, ax) * b) - np.sum(a * tr(b, ax))) < 1e-12


def test_tensor_transposes():
    shape = (3, 4, 5)
    rng = np.random.default_rng(1)
    v, Q = rng.standard_normal((3,) + shape), rng.standard_normal((6,) + shape)
    assert abs(np.sum(adj.eps(v) * Q) - np.sum(v * adj.eps_t(Q))) < 1e-12
    W = rng.standard_normal((3,) + shape)
    assert abs(np.sum(adj.div2(Q) * W) - np.sum(Q * adj.div2_t(W))) < 1e-12


def _loss(f, it, prm, target):
    u = adj.forward(f, it, **prm)["u"]
    return 0.5 * np.sum((u - target) ** 2), u


@pytest.mark.parametrize("iters", [1, 3, 25])
def test_gradients_against_central_differences(iters):
    f = _f()
    target = np.zeros_like(f)
    target[:, : f.shape[1] // 2, :] = 0.6
    L0, u = _loss(f, iters, PRM, target)
    _, g = adj.adjoint(f, iters, dL_du=u - target, **PRM)
    rng = np.random.default_rng(5)
    d = rng.standard_normal(f.shape)
    d /= np.linalg.norm(d)
    h = 1e-5
    fd = (_loss(f + h * d, iters, PRM, target)[0] - _loss(f - h * d, iters, PRM, target)[0]) / (2 * h)
    assert abs(fd - np.sum(g["f"] * d)) <= 1e-6 * max(1.0, abs(fd))
    # scalar steps small enough not to move any voxel across a projection kink
    hs = 1e-7
    for name in ("alpha0", "alpha1", "lam"):
        prm_p = dict(PRM); prm_p[name] += hs
        prm_m = dict(PRM); prm_m[name] -= hs
        fd = (_loss(f, iters, prm_p, target)[0] - _loss(f, iters, prm_m, target)[0]) / (2 * hs)
        assert abs(fd - g[name]) <= 1e-6 * max(1.0, abs(fd)), (name, fd, g[name])
    # the projections were active (otherwise alpha gradients are trivially 0)
    if iters == 25:
        assert abs(g["alpha0"]) > 0 and abs(g["alpha1"]) > 0
