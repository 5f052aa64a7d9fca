"""amb_tgv2_energy_noalloc (the FP64 energy / gap with no device allocation:
partials in the library's fixed per-device scratch array) against
amb_tgv2_energy on the same state.  Both run the same kernels over the same
grid, so the ten values agree bit for bit.  Two calls in a row, the second with
a smaller volume (the scratch array keeps the first call's rows); the first
volume is large enough for the grid cap of 4096 blocks."""
import ctypes

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

TAU = float(1 / np.sqrt(20))


def _params(native, data_term, storage):
    return native.default_params(lam=0.75, alpha0=0.3125, alpha1=0.15625, tau=TAU, sigma=TAU,
                                 theta=1.0, iterations=0,
                                 data_term=native.DATA_L1 if data_term == "l1" else native.DATA_L2,
                                 storage=storage)


def _state(ctx, native, shape, seed, data_term, storage, iters):
    """The context's state after `iters` iterations, its amb_tgv2_energy, and
    device copies of f and the 17 fields in the storage type."""
    from ambient_image_processor_amd import phantoms
    f = phantoms.deshade_phantom(shape, seed=seed, kind="spheres", noise=0.03, bias_order=3,
                                 bias_sigma=0.05)[0]
    p = _params(native, data_term, storage)
    ctx.begin(f, p)
    ctx.iterate(iters)
    ref = ctx.energy()
    st = ctx.get_state()
    if storage == native.STORAGE_BF16:
        bits = (st.view(np.uint32) >> 16).astype(np.uint16)
        assert np.array_equal((bits.astype(np.uint32) << 16).view(np.float32), st)
        fields = torch.from_numpy(bits.view(np.int16)).to("cuda:0")
    else:
        fields = torch.from_numpy(st).to("cuda:0")
    return p, ref, torch.from_numpy(f).to("cuda:0"), fields


def _noalloc(native, p, f_dev, fields, shape):
    out = torch.full((10,), float("nan"), dtype=torch.float64, device="cuda:0")
    ptrs = (ctypes.c_void_p * 17)(*[fields[k].data_ptr() for k in range(17)])
    d, h, w = shape
    torch.cuda.synchronize()
    native._check(native.lib().amb_tgv2_energy_noalloc(
        ctypes.c_void_p(f_dev.data_ptr()), ptrs, w, h, d, ctypes.byref(p), None,
        ctypes.c_void_p(out.data_ptr())))
    torch.cuda.synchronize()
    return out.cpu().numpy()


@pytest.mark.parametrize("first,second", [
    (((96, 128, 128), "l1", 0), ((9, 10, 12), "l2", 1)),
    (((40, 48, 64), "l2", 1), ((7, 5, 6), "l1", 0)),
])
def test_noalloc_energy_equals_context_energy(ctx, first, second):
    from ambient_image_processor_amd import native
    for k, (shape, data_term, storage) in enumerate((first, second)):
        print("call", k, shape, data_term, "bf16" if storage else "f32", flush=True)
        p, ref, f_dev, fields = _state(ctx, native, shape, 11 + k, data_term, storage, 25)
        got = _noalloc(native, p, f_dev, fields, shape)
        np.testing.assert_array_equal(got, ref)


def test_noalloc_rejects_bad_arguments(ctx):
    from ambient_image_processor_amd import native
    p = _params(native, "l1", 0)
    ptrs = (ctypes.c_void_p * 17)()
    rc = native.lib().amb_tgv2_energy_noalloc(None, ptrs, 4, 4, 4, ctypes.byref(p), None, None)
    assert rc == 1
