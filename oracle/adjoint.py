"""FP64 CPU adjoint of the unrolled TGV2-L2 iteration (test infrastructure only).

The forward map is the reference loop (tgv_deshade/tgv2_l1_deshade.cu:169-264 with the
L2 u-update of tgv_cuda/tgv2_l2.cu:27-56), written with numpy slicing in FP64; the
reverse map differentiates it exactly as implemented: B is the reference's
"backward difference" (not the transpose of D+), so the reverse sweeps use the
true transposes of D+, B, eps_B and div2.  Gradients of L(u_N) with respect to
f, alpha0, alpha1 and lambda, given dL/du_N.  Non-smooth points (|p| = alpha1,
|q|_w = alpha0) take the branch the forward map takes (inactive at equality).
Pinned by central finite differences (tests/test_adjoint_oracle.py).
"""
from __future__ import annotations

import numpy as np

X, Y, Z = 2, 1, 0   # numpy axes of x, y, z for (D, H, W) volumes


def _sl(ax, s):
    t = [slice(None)] * 3
    t[ax] = s
    return tuple(t)


def dplus(a, ax):
    out = np.zeros_like(a)
    out[_sl(ax, slice(0, -1))] = a[_sl(ax, slice(1, None))] - a[_sl(ax, slice(0, -1))]
    return out


def dplus_t(b, ax):
    """(D+)^T b [j] = b[j-1] [j >= 1] - b[j] [j < n-1]"""
    out = np.zeros_like(b)
    out[_sl(ax, slice(1, None))] += b[_sl(ax, slice(0, -1))]
    out[_sl(ax, slice(0, -1))] -= b[_sl(ax, slice(0, -1))]
    return out


def bm(a, ax):
    """reference backward difference: -a[i] + a[i-1], -a[0] at 0"""
    out = -a.copy()
    out[_sl(ax, slice(1, None))] += a[_sl(ax, slice(0, -1))]
    return out


def bm_t(b, ax):
    """B^T b [j] = b[j+1] [j < n-1] - b[j]"""
    out = -b.copy()
    out[_sl(ax, slice(0, -1))] += b[_sl(ax, slice(1, None))]
    return out


def eps(vb):
    return np.stack([bm(vb[0], X), bm(vb[1], Y), bm(vb[2], Z),
                     0.5 * (bm(vb[1], X) + bm(vb[0], Y)),
                     0.5 * (bm(vb[2], X) + bm(vb[0], Z)),
                     0.5 * (bm(vb[2], Y) + bm(vb[1], Z))])


def eps_t(Q):
    return np.stack([bm_t(Q[0], X) + 0.5 * bm_t(Q[3], Y) + 0.5 * bm_t(Q[4], Z),
                     bm_t(Q[1], Y) + 0.5 * bm_t(Q[3], X) + 0.5 * bm_t(Q[5], Z),
                     bm_t(Q[2], Z) + 0.5 * bm_t(Q[4], X) + 0.5 * bm_t(Q[5], Y)])


def div2(q):
    return np.stack([dplus(q[0], X) + dplus(q[3], Y) + dplus(q[4], Z),
                     dplus(q[1], Y) + dplus(q[3], X) + dplus(q[5], Z),
                     dplus(q[2], Z) + dplus(q[4], X) + dplus(q[5], Y)])


def div2_t(W):
    return np.stack([dplus_t(W[0], X), dplus_t(W[1], Y), dplus_t(W[2], Z),
                     dplus_t(W[0], Y) + dplus_t(W[1], X),
                     dplus_t(W[0], Z) + dplus_t(W[2], X),
                     dplus_t(W[1], Z) + dplus_t(W[2], Y)])


_QW = np.array([1.0, 1.0, 1.0, 2.0, 2.0, 2.0])[:, None, None, None]


def init_state(f):
    f = np.asarray(f, dtype=np.float64)
    z3 = np.zeros((3,) + f.shape)
    return dict(u=f.copy(), ub=f.copy(), v=z3.copy(), vb=z3.copy(), p=z3.copy(),
                q=np.zeros((6,) + f.shape))


def step(s, f, alpha0, alpha1, lam, tau, sigma, theta):
    """One L2 iteration; returns the new state."""
    pt = s["p"] + sigma * (np.stack([dplus(s["ub"], X), dplus(s["ub"], Y), dplus(s["ub"], Z)])
                           - s["vb"])
    npn = np.sqrt((pt ** 2).sum(0))
    p = pt / np.maximum(1.0, npn / alpha1)
    qt = s["q"] + sigma * eps(s["vb"])
    nq = np.sqrt((_QW * qt ** 2).sum(0))
    q = qt / np.maximum(1.0, nq / alpha0)
    a = s["u"] - tau * (bm(p[0], X) + bm(p[1], Y) + bm(p[2], Z))
    u = (a + tau * lam * f) / (1.0 + tau * lam)
    ub = u + theta * (u - s["u"])
    v = s["v"] - tau * (div2(q) - p)
    vb = v + theta * (v - s["v"])
    return dict(u=u, ub=ub, v=v, vb=vb, p=p, q=q)


def forward(f, iterations, alpha0, alpha1, lam, tau, sigma, theta=1.0):
    s = init_state(f)
    for _ in range(iterations):
        s = step(s, np.asarray(f, np.float64), alpha0, alpha1, lam, tau, sigma, theta)
    return s


def reverse_step(s, A, f, alpha0, alpha1, lam, tau, sigma, theta, grads):
    """Given state s = x_k and adjoints A of x_{k+1} (keys u, ub, v, vb, p, q),
    return the adjoints of x_k; accumulate into grads (f, alpha0, alpha1, lam)."""
    # recompute the forward step's intermediates
    pt = s["p"] + sigma * (np.stack([dplus(s["ub"], X), dplus(s["ub"], Y), dplus(s["ub"], Z)])
                           - s["vb"])
    npn = np.sqrt((pt ** 2).sum(0))
    act_p = npn / alpha1 > 1.0
    p = np.where(act_p, pt / np.where(act_p, npn / alpha1, 1.0), pt)
    qt = s["q"] + sigma * eps(s["vb"])
    nq = np.sqrt((_QW * qt ** 2).sum(0))
    act_q = nq / alpha0 > 1.0
    q = np.where(act_q, qt / np.where(act_q, nq / alpha0, 1.0), qt)
    a = s["u"] - tau * (bm(p[0], X) + bm(p[1], Y) + bm(p[2], Z))
    tl = tau * lam
    un = (a + tl * f) / (1.0 + tl)

    Un, UBn, Vn, VBn = A["u"].copy(), A["ub"], A["v"].copy(), A["vb"]
    Pn, Qn = A["p"].copy(), A["q"].copy()
    # v_bar' = (1 + theta) v' - theta v ;  v' = v - tau (div2 q' - p')
    Vn += (1.0 + theta) * VBn
    V_in = -theta * VBn + Vn
    Wst = -tau * Vn
    Pn += tau * Vn
    Qn += div2_t(Wst)
    # u_bar' = (1 + theta) u' - theta u ;  u' = (a + tl f) / (1 + tl) ; a = u - tau B p'
    Un += (1.0 + theta) * UBn
    U_in = -theta * UBn
    Ast = Un / (1.0 + tl)
    grads["f"] += tl / (1.0 + tl) * Un
    grads["lam"] += float(np.sum(Un * tau * (f - un) / (1.0 + tl)))
    U_in += Ast
    Pn += -tau * np.stack([bm_t(Ast, X), bm_t(Ast, Y), bm_t(Ast, Z)])
    # q' = q~ / max(1, |q~|_w / alpha0)
    dot_q = (Qn * qt).sum(0)
    nq_safe = np.where(act_q, nq, 1.0)
    Qt = np.where(act_q, alpha0 / nq_safe * (Qn - dot_q * _QW * qt / nq_safe ** 2), Qn)
    grads["alpha0"] += float(np.sum(np.where(act_q, dot_q / nq_safe, 0.0)))
    # p' = p~ / max(1, |p~| / alpha1)
    dot_p = (Pn * pt).sum(0)
    np_safe = np.where(act_p, npn, 1.0)
    Pt = np.where(act_p, alpha1 / np_safe * (Pn - dot_p * pt / np_safe ** 2), Pn)
    grads["alpha1"] += float(np.sum(np.where(act_p, dot_p / np_safe, 0.0)))
    # q~ = q + sigma eps_B(v_bar) ; p~ = p + sigma (D+ u_bar - v_bar)
    Q_in = Qt
    VB_in = sigma * eps_t(Qt) - sigma * Pt
    P_in = Pt
    UB_in = sigma * (dplus_t(Pt[0], X) + dplus_t(Pt[1], Y) + dplus_t(Pt[2], Z))
    return dict(u=U_in, ub=UB_in, v=V_in, vb=VB_in, p=P_in, q=Q_in)


def adjoint(f, iterations, alpha0, alpha1, lam, tau, sigma, theta, dL_du, checkpoint_every=0):
    """Returns (u_N, grads) with grads = dict(f, alpha0, alpha1, lam).  All states
    are kept (checkpoint_every is accepted for interface parity and ignored)."""
    f = np.asarray(f, dtype=np.float64)
    states = [init_state(f)]
    for _ in range(iterations):
        states.append(step(states[-1], f, alpha0, alpha1, lam, tau, sigma, theta))
    z3 = np.zeros((3,) + f.shape)
    A = dict(u=np.asarray(dL_du, np.float64).copy(), ub=np.zeros_like(f), v=z3.copy(),
             vb=z3.copy(), p=z3.copy(), q=np.zeros((6,) + f.shape))
    grads = dict(f=np.zeros_like(f), alpha0=0.0, alpha1=0.0, lam=0.0)
    for k in range(iterations - 1, -1, -1):
        A = reverse_step(states[k], A, f, alpha0, alpha1, lam, tau, sigma, theta, grads)
    # x_0: u = u_bar = f
    grads["f"] += A["u"] + A["ub"]
    return states[-1]["u"], grads
