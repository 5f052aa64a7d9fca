"""torch.autograd.Function for the unrolled TGV2-L2 denoise, differentiated by
the hand-written adjoint of the C ABI (amb_tgv2_denoise_adjoint).

    u = tgv2_l2_denoise(f, alpha0, alpha1, lam, iterations=200, checkpoint_every=20)

f is a float32 tensor of shape (D, H, W) or (B, D, H, W) on the GPU (device
pointers go straight to the library) or on the CPU (host buffers).  alpha0,
alpha1 and lam may be Python floats or 0-d tensors; gradients flow to f and to
any of them that require grad.  tau = sigma = 1/sqrt(20) and theta = 1 by
default, as the reference hard-codes (tgv_deshade/tgv2_l1_deshade.cu:148-150).
The backward pass re-runs the forward iteration with a checkpoint every
`checkpoint_every` iterations and recomputes each segment in reverse;
checkpoint_every = 1 keeps every state instead (the fastest backward, with
iterations + 3 copies of the 17-field state in device memory).
"""
from __future__ import annotations

import contextlib
import math
import threading

import numpy as np
import torch

from . import native

# One library context per device, shared by every thread; each holds a lock
# around the calls of one forward or backward (a context keeps the iteration
# state, scratch and adjoint workspace of the call in progress, and autograd
# may run a backward on its own thread while another thread runs a forward).
_contexts = {}
_contexts_lock = threading.Lock()


@contextlib.contextmanager
def _context(device_index: int):
    with _contexts_lock:
        entry = _contexts.get(device_index)
        if entry is None:
            entry = (native.Context(device_index), threading.Lock())
            _contexts[device_index] = entry
    c, lock = entry
    with lock:
        yield c


def _scalar(x) -> float:
    return float(x.detach().cpu().item()) if torch.is_tensor(x) else float(x)


def _params(alpha0, alpha1, lam, iterations, tau, sigma, theta):
    return native.default_params(lam=lam, alpha0=alpha0, alpha1=alpha1, tau=tau, sigma=sigma,
                                 theta=theta, iterations=iterations, data_term=native.DATA_L2)


class TGV2L2DenoiseFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, f, alpha0, alpha1, lam, iterations, tau, sigma, theta, checkpoint_every):
        if f.dtype != torch.float32:
            raise TypeError("f must be float32")
        f = f.contiguous()
        vols = f if f.dim() == 4 else f.unsqueeze(0)
        if vols.dim() != 4:
            raise ValueError("f must have shape (D, H, W) or (B, D, H, W)")
        p = _params(_scalar(alpha0), _scalar(alpha1), _scalar(lam), int(iterations), tau, sigma,
                    theta)
        shape = tuple(vols.shape[1:])
        u = torch.empty_like(vols)
        if vols.is_cuda:
            torch.cuda.current_stream(vols.device).synchronize()
            with _context(vols.device.index or 0) as c:
                for b in range(vols.shape[0]):
                    native._check(native.lib().amb_tgv2_denoise(
                        c._h, native.MEM_DEVICE, native.c_void_p(vols[b].data_ptr()),
                        *native._dims(shape), native.ctypes.byref(p),
                        native.c_void_p(u[b].data_ptr()), native.CALLBACK(), None))
        else:
            with _context(0) as c:
                for b in range(vols.shape[0]):
                    u[b] = torch.from_numpy(c.tgv2_denoise(vols[b].numpy(), p))
        ctx.save_for_backward(vols)
        ctx.meta = (p, shape, int(checkpoint_every), f.dim(),
                    [torch.is_tensor(a) for a in (alpha0, alpha1, lam)],
                    [a.device if torch.is_tensor(a) else None for a in (alpha0, alpha1, lam)])
        return u if f.dim() == 4 else u[0]

    @staticmethod
    def backward(ctx, grad_u):
        (vols,) = ctx.saved_tensors
        p, shape, k, fdim, is_t, devs = ctx.meta
        g = grad_u.contiguous().to(torch.float32)
        g = g if g.dim() == 4 else g.unsqueeze(0)
        grad_f = torch.empty_like(vols)
        gp = np.zeros(3)
        if vols.is_cuda:
            torch.cuda.current_stream(vols.device).synchronize()
            with _context(vols.device.index or 0) as c:
                for b in range(vols.shape[0]):
                    gp += c.tgv2_denoise_adjoint_device(vols[b].data_ptr(), shape, p,
                                                        g[b].data_ptr(), None,
                                                        grad_f[b].data_ptr(), k)
        else:
            with _context(0) as c:
                for b in range(vols.shape[0]):
                    _, gf, gpb = c.tgv2_denoise_adjoint(vols[b].numpy(), p, g[b].numpy(), k)
                    grad_f[b] = torch.from_numpy(gf)
                    gp += gpb
        outs = []
        for j in range(3):
            outs.append(torch.tensor(gp[j], dtype=torch.float32, device=devs[j]) if is_t[j] else None)
        gf_out = grad_f if fdim == 4 else grad_f[0]
        return (gf_out, outs[0], outs[1], outs[2], None, None, None, None, None)


def tgv2_l2_denoise(f, alpha0=2.0, alpha1=1.0, lam=1.0, iterations=200,
                    tau=1.0 / math.sqrt(20.0), sigma=1.0 / math.sqrt(20.0), theta=1.0,
                    checkpoint_every=20):
    """Differentiable TGV2-L2 denoise (see the module docstring)."""
    return TGV2L2DenoiseFunction.apply(f, alpha0, alpha1, lam, iterations, tau, sigma, theta,
                                       checkpoint_every)
