"""Pins the oracle's BF16 field-storage mode (FP32 arithmetic, every field store
rounded to bfloat16, nearest even), the checker of the GPU's BF16 storage.

* the rounding rule against torch's float32 -> bfloat16 cast (round to nearest
  even) on ties, denormals, overflow, signed zeros and random values;
* the loop against an independent vectorised restatement in float32 with the
  reference's fmaf placement (tgv_cuda/tgv2_common.cu:96-131, 222-276, 324-354;
  tgv_cuda/tgv2_l1.cu:27-60, tgv2_l2.cu:27-56) and a bfloat16 round (torch) at
  each field store - bit for bit;
* invariants: every stored field is a bfloat16 value; the first dual step
  (which reads only the initial state) is the FP32 step rounded.
"""
import numpy as np
import pytest
import torch

from ambient_image_processor_amd import phantoms

PRM20 = dict(lam=1.0, alpha0=2.0, alpha1=1.0, tau=float(1 / np.sqrt(20)),
             sigma=float(1 / np.sqrt(20)), theta=1.0)


def _torch_bf16(a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).to(
        torch.bfloat16).to(torch.float32).numpy()


def test_rounding_rule_matches_torch(orc):
    rng = np.random.default_rng(0)
    bits = [rng.integers(0, 2 ** 32, 200000, dtype=np.uint64).astype(np.uint32)]
    # exact ties with an even and an odd bfloat16 mantissa, and their neighbours
    hi = rng.integers(0, 2 ** 16, 4000, dtype=np.uint64).astype(np.uint32) << 16
    for low in (0x8000, 0x7fff, 0x8001, 0x0000, 0xffff):
        bits.append(hi | np.uint32(low))
    # denormals, the largest finite values (round up to infinity), zeros, infinities
    bits.append(np.arange(0, 0x00020000, 7, dtype=np.uint32))
    bits.append(np.array([0x7f7fffff, 0x7f7f8000, 0x7f7f7fff, 0xff7fffff, 0x80000000, 0,
                          0x7f800000, 0xff800000, 0x00008000, 0x80018000], dtype=np.uint32))
    b = np.concatenate(bits)
    a = b.view(np.float32)
    a = a[~np.isnan(a)]
    got = orc.bf16_rne(a)
    ref = _torch_bf16(a)
    np.testing.assert_array_equal(got.view(np.uint32), ref.view(np.uint32))


def _fma(a, b, c):
    """float32 fmaf: the float64 product of two float32 values is exact, so one
    rounding of the float64 sum to float32 (double rounding is possible only at
    a float32 tie after a float64 rounding, which these inputs do not reach)."""
    return (a.astype(np.float64) * b + c).astype(np.float32)


def _restated_bf16_loop(f, iters, lam, alpha0, alpha1, tau, sigma, theta, data_term, rnd):
    """The reference loop (p, u, q, v order of tgv_deshade/tgv2_l1_deshade.cu:169-264)
    in float32 numpy slices, ``rnd`` applied where each kernel stores a field."""
    f32 = np.float32
    X, Y, Z = 2, 1, 0

    def dplus(a, ax):
        out = np.zeros_like(a)
        s0 = [slice(None)] * 3
        s1 = [slice(None)] * 3
        s0[ax] = slice(0, -1)
        s1[ax] = slice(1, None)
        out[tuple(s0)] = a[tuple(s1)] - a[tuple(s0)]
        return out

    def bminus(a, ax):
        out = -a
        s0 = [slice(None)] * 3
        s1 = [slice(None)] * 3
        s0[ax] = slice(1, None)
        s1[ax] = slice(0, -1)
        out[tuple(s0)] = out[tuple(s0)] + a[tuple(s1)]
        return out

    tau, sigma, theta = f32(tau), f32(sigma), f32(theta)
    lam, a0, a1 = f32(lam), f32(alpha0), f32(alpha1)
    tl = f32(tau * lam)
    f = f.astype(f32)
    u = rnd(f)
    ub = u.copy()
    z = np.zeros_like(f)
    v = [z.copy() for _ in range(3)]
    vb = [z.copy() for _ in range(3)]
    p = [z.copy() for _ in range(3)]
    q = [z.copy() for _ in range(6)]
    for _ in range(iters):
        # p (tgv_kernel_part22)
        g = [dplus(ub, X), dplus(ub, Y), dplus(ub, Z)]
        a = [_fma(sigma, g[c] - vb[c], p[c]) for c in range(3)]
        n2 = a[0] * a[0]
        n2 = _fma(a[1], a[1], n2)
        n2 = _fma(a[2], a[2], n2)
        nrm = np.maximum(f32(1), np.sqrt(n2) / a1)
        p = [rnd(a[c] / nrm) for c in range(3)]
        u_old = u
        # u (tgv_kernel_part4_tgv2_l1 / _l2)
        div = (bminus(p[0], X) + bminus(p[1], Y)) + bminus(p[2], Z)
        un = _fma(-tau, div, u)
        if data_term == "l1":
            d = un - f
            un = np.where(d > tl, un - tl, np.where(d < -tl, un + tl, f))
        else:
            un = _fma(tl, f, un) / (f32(1) + tl)
        ub = rnd(_fma(theta, un - u_old, un))
        u = rnd(un)
        # q (gradient2 + tgv_kernel_part5)
        e = [bminus(vb[0], X), bminus(vb[1], Y), bminus(vb[2], Z),
             (bminus(vb[1], X) + bminus(vb[0], Y)) * f32(0.5),
             (bminus(vb[2], X) + bminus(vb[0], Z)) * f32(0.5),
             (bminus(vb[2], Y) + bminus(vb[1], Z)) * f32(0.5)]
        b = [_fma(sigma, e[c], q[c]) for c in range(6)]
        m2 = b[0] * b[0]
        m2 = _fma(b[1], b[1], m2)
        m2 = _fma(f32(2) * b[3], b[3], m2)
        m2 = _fma(b[2], b[2], m2)
        m2 = _fma(f32(2) * b[4], b[4], m2)
        m2 = _fma(f32(2) * b[5], b[5], m2)
        nrm = np.maximum(f32(1), np.sqrt(m2) / a0)
        q = [rnd(b[c] / nrm) for c in range(6)]
        v_old = v
        # v (divergence2 + tgv_kernel_part6)
        w = [(dplus(q[0], X) + dplus(q[3], Y)) + dplus(q[4], Z),
             (dplus(q[1], Y) + dplus(q[3], X)) + dplus(q[5], Z),
             (dplus(q[2], Z) + dplus(q[4], X)) + dplus(q[5], Y)]
        vn = [_fma(-tau, w[c] - p[c], v_old[c]) for c in range(3)]
        vb = [rnd(_fma(theta, vn[c] - v_old[c], vn[c])) for c in range(3)]
        v = [rnd(vn[c]) for c in range(3)]
    return np.stack([u, ub] + v + vb + p + q)


@pytest.mark.parametrize("data_term", ["l1", "l2"])
@pytest.mark.parametrize("bf16", [False, True])
def test_bf16_loop_matches_restatement(orc, data_term, bf16):
    f = phantoms.deshade_phantom((7, 9, 12), seed=4, kind="spheres", noise=0.02,
                                 bias_order=3, bias_sigma=0.05, bias_max=0.2)[0]
    st = orc.tgv2(f, 25, data_term=data_term, dtype=np.float32, bf16=bf16, **PRM20)
    ref = _restated_bf16_loop(f, 25, data_term=data_term,
                              rnd=_torch_bf16 if bf16 else (lambda a: a), **PRM20)
    for k, name in enumerate(orc.FIELD_NAMES):
        np.testing.assert_array_equal(st.fields[k], ref[k], err_msg=name)


def test_bf16_loop_matches_restatement_c4_crop(orc):
    """Config 4's generator (spheres, bias max 0.2, noise 0.02; L2, lambda 1) on a
    16 x 20 x 24 volume, 40 iterations, with projections active."""
    f = phantoms.deshade_phantom((16, 20, 24), seed=4, kind="spheres", noise=0.02,
                                 bias_order=3, bias_sigma=0.05, bias_max=0.2)[0]
    prm = dict(PRM20, alpha0=0.05, alpha1=0.02)
    st = orc.tgv2(f, 40, data_term="l2", dtype=np.float32, bf16=True, **prm)
    ref = _restated_bf16_loop(f, 40, data_term="l2", rnd=_torch_bf16, **prm)
    np.testing.assert_array_equal(st.fields, ref)
    pn = np.sqrt((st.p.astype(np.float64) ** 2).sum(0))
    assert np.mean(pn > 0.99 * prm["alpha1"]) > 0.05     # the p projection is active


def test_bf16_fields_are_bfloat16_values(orc):
    f = phantoms.deshade_phantom((12, 10, 16), seed=4, kind="spheres", noise=0.02,
                                 bias_order=3, bias_sigma=0.05, bias_max=0.2)[0]
    st = orc.tgv2(f, 30, data_term="l2", dtype=np.float32, bf16=True, **PRM20)
    for k, name in enumerate(orc.FIELD_NAMES):
        np.testing.assert_array_equal(orc.bf16_rne(st.fields[k]), st.fields[k], err_msg=name)
    # f stays FP32: the prox reads the unrounded input
    assert np.any(orc.bf16_rne(st.f) != st.f)


def test_bf16_first_dual_step_is_rounded_fp32(orc):
    """The first dual step reads u_bar = f and zero v_bar, p, q only; for an f of
    bfloat16 values both modes start from the same state, so p and q after one
    iteration are the FP32 mode's rounded once."""
    f = phantoms.deshade_phantom((9, 8, 12), seed=4, kind="spheres", noise=0.02,
                                 bias_order=3, bias_sigma=0.05, bias_max=0.2)[0]
    f = orc.bf16_rne(f)
    a = orc.tgv2(f, 1, data_term="l2", dtype=np.float32, bf16=True, **PRM20)
    b = orc.tgv2(f, 1, data_term="l2", dtype=np.float32, **PRM20)
    np.testing.assert_array_equal(a.p, orc.bf16_rne(b.p))
    np.testing.assert_array_equal(a.q, orc.bf16_rne(b.q))
    assert np.any(a.p != b.p)
