"""CPU tests of the host side: synthetic inputs, and that the C-ABI library loads
and exports every function include/ambient_amd.h declares (no GPU calls)."""
import ctypes
import hashlib
import os
import re

import numpy as np
import pytest

from conftest import ROOT, rel_err
from ambient_image_processor_amd import phantoms


def test_splitmix64_published_vectors():
    r = phantoms.SplitMix64(0)
    assert [int(x) for x in r.next_u64(3)] == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4,
                                              0x06C45D188009454F]
    r = phantoms.SplitMix64(1234567)
    assert [int(x) for x in r.next_u64(3)] == [6457827717110365317, 3203168211198807973,
                                              9817491932198370423]


def test_stream_chunking_is_consistent():
    a = phantoms.SplitMix64(9).next_u64(10)
    r = phantoms.SplitMix64(9)
    b = np.concatenate([r.next_u64(3), r.next_u64(7)])
    np.testing.assert_array_equal(a, b)


def test_normal_moments():
    z = phantoms.SplitMix64(5).normal(200000)
    assert abs(z.mean()) < 0.01 and abs(z.std() - 1) < 0.01


def test_phantom_deterministic_and_shaped():
    f1, c1, b1 = phantoms.deshade_phantom((16, 20, 24), seed=1)
    f2, _, _ = phantoms.deshade_phantom((16, 20, 24), seed=1)
    assert f1.shape == (16, 20, 24) and f1.dtype == np.float32
    assert hashlib.sha256(f1.tobytes()).hexdigest() == hashlib.sha256(f2.tobytes()).hexdigest()
    assert abs(b1.mean()) < 0.05          # no DC term in the bias
    vals = np.unique(c1)
    assert 0.1 in vals.astype(np.float64).round(6)


def test_poisson_field_rhs_is_neumann_laplacian():
    rhs, phi = phantoms.poisson_field((8, 10, 12), seed=3, modes=10, kmax=4)
    lap = phantoms.neumann_laplacian(phi)
    assert rel_err(rhs, lap) < 1e-6        # rhs is stored float32
    assert abs(phi.mean()) < 1e-12


def _declared_functions():
    hdr = open(os.path.join(ROOT, "include", "ambient_amd.h")).read()
    return sorted(set(re.findall(r"AMB_API\s+[\w\s\*]+?\b(amb_\w+)\s*\(", hdr)))


def test_header_and_binding_agree():
    from ambient_image_processor_amd import native
    assert _declared_functions() == sorted(native.EXPORTS)


def test_library_exports_every_declared_symbol():
    from ambient_image_processor_amd import native
    if not os.path.exists(native.LIB_PATH):
        pytest.fail("libambient_amd.so not built (run __graft_entry__.build())")
    lib = ctypes.CDLL(native.LIB_PATH)
    for name in _declared_functions():
        assert hasattr(lib, name), name


def test_params_default_matches_reference_constants():
    from ambient_image_processor_amd import native
    p = native.default_params()
    assert p.tau == pytest.approx(1 / np.sqrt(20), rel=1e-7)
    assert p.sigma == p.tau and p.theta == 1.0
    assert (p.lambda_, p.alpha0, p.alpha1) == (1.0, 2.0, 1.0)
    assert p.data_term == native.DATA_L1 and p.flags == native.FLAG_SET_NEGATIVE_ZERO


def test_context_without_gpu_raises():
    """No silent fallback: on a machine without a GPU the context cannot be made."""
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    from ambient_image_processor_amd import native
    with pytest.raises(native.AmbError):
        native.Context(0)


def test_build_mask_indices_matches_brute_force():
    """TGVDeshadeMaskedProcessor.buildMaskIndices mirror (reference
    TGVDeshadeMaskedProcessor.cpp:337-385) against a per-voxel loop."""
    from ambient_image_processor_amd import processor
    rng = np.random.default_rng(0)
    d, h, w = 4, 5, 6
    m = (rng.random((d, h, w)) * 1.6).astype(np.float32)
    got = processor.TGVDeshadeMaskedProcessor.buildMaskIndices(m)
    want = [[] for _ in range(13)]
    for z in range(d):
        for y in range(h):
            for x in range(w):
                if m[z, y, x] < 1:
                    continue
                i = (z * h + y) * w + x
                want[0].append(i)
                for k, edge in enumerate((x == 0, x == w - 1, y == 0, y == h - 1, z == 0,
                                          z == d - 1)):
                    want[1 + 2 * k + (0 if edge else 1)].append(i)
    assert len(got) == 13
    for g, ww in zip(got, want):
        assert g.dtype == np.uint32
        assert g.tolist() == ww


def test_masked_launch_rejects_out_of_range_index():
    from ambient_image_processor_amd import processor
    f = np.zeros(2 * 3 * 4, np.float32)
    with pytest.raises(IndexError):
        processor.tgv2_l1_deshade_masked_launch(f, 4, 3, 2, 1.0, 1, 0, -1, None, 2.0, 1.0, [24])
