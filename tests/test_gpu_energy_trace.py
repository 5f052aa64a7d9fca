"""The FP64 energy / gap reduced inside the fused sweep (amb_tgv2_set_energy_interval,
amb_tgv2_energy_trace) against the standalone FP64 pass (amb_tgv2_energy) and the
FP64 oracle's energy of the same state.

Each trace row is the state after iteration c of the call.  The fused sweep's
energy form computes the standalone pass's per-voxel terms from the values the
sweep holds, so the two differ only in the order of the FP64 sums (1e-9
relative, scaled by the energy).  The iteration itself must be unchanged bit for
bit with the reduction on.  (The reference computes no energy; the closest
quantity is the host NCC/TV of l, tgv_deshade/TGVDeshadeProcessor.cpp:509-517.)
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TAU = float(1 / np.sqrt(20))
# weights exact in float32, so the oracle's FP64 energy uses the library's values
LAM, A0, A1 = 0.75, 0.3125, 0.15625


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _params(native, iterations, data_term, storage=0, **kw):
    return native.default_params(lam=LAM, alpha0=A0, alpha1=A1, tau=TAU, sigma=TAU,
                                 theta=1.0, iterations=iterations,
                                 data_term=native.DATA_L1 if data_term == "l1" else native.DATA_L2,
                                 storage=storage, **kw)


def _volume(shape, seed):
    from ambient_image_processor_amd import phantoms
    return phantoms.deshade_phantom(shape, seed=seed, kind="spheres", noise=0.03, bias_order=3,
                                    bias_sigma=0.05)[0]


def _close(a, b, scale):
    """|a - b| <= 1e-9 * scale for each of the ten quantities."""
    return np.all(np.abs(a - b) <= 1e-9 * scale)


# (D, H, W): whole-row bands, x tiles (rows wider than 510), odd tile and band
# edges; odd widths run the one-voxel energy form (whole rows up to 170, then
# 126-column tiles)
SHAPES = [(20, 24, 32), (9, 13, 130), (6, 11, 600), (7, 10, 33), (5, 9, 263)]


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("data_term", ["l1", "l2"])
@pytest.mark.parametrize("storage", [0, 1])
def test_trace_matches_standalone_pass(ctx, orc, shape, data_term, storage):
    from ambient_image_processor_amd import native
    _say("case", shape, data_term, "bf16" if storage else "f32")
    f = _volume(shape, seed=sum(shape))
    iters, every = 37, 6
    ctx.set_energy_interval(0)
    plain = ctx.tgv2_deshade(f, _params(native, iters, data_term, storage))
    ctx.set_energy_interval(every)
    ctx.profile(True)
    ctx.profile_reset()
    traced = ctx.tgv2_deshade(f, _params(native, iters, data_term, storage))
    prof = ctx.profile_read()
    ctx.profile(False)
    trace = ctx.energy_trace()
    ctx.set_energy_interval(0)
    for a, b in zip(plain, traced):          # the iteration is unchanged
        np.testing.assert_array_equal(a, b)
    counts = [int(c) for c in trace[:, 0]]
    assert counts == [6, 12, 18, 24, 30, 36, 37]
    # the fused sweep ran, with its energy form for the rows
    assert prof["fused_sweep_energy"][1] == len(counts) - 1, prof
    # the standalone pass on the same states
    ctx.begin(f, _params(native, 0, data_term, storage))
    done = 0
    for row in trace:
        c = int(row[0])
        ctx.iterate(c - done)
        done = c
        ref = ctx.energy()
        scale = max(abs(ref[0]), 1e-30)
        assert _close(row[1:], ref, scale), (c, row[1:], ref)
        if c == counts[-1]:
            st = orc.TGVState(f, np.float32)
            st.fields[:] = ctx.get_state()
            e64 = st.energy(lam=LAM, alpha0=A0, alpha1=A1, data_term=data_term)
            assert _close(row[1:], e64, scale), (row[1:], e64)
    _say("final E", trace[-1, 1], "gap", trace[-1, 5])


def test_trace_masked_and_callback(ctx):
    """Masked iteration (energy form of the masked sweep) with a callback every
    10 iterations that stops at 20: rows at 5, 10, 15, 20; the callback
    iterations run the primal sweep alone and take the standalone pass."""
    from ambient_image_processor_amd import native
    f = _volume((12, 16, 48), seed=5)
    rng = np.random.default_rng(5)
    active = (rng.random(f.shape) < 0.6).astype(np.uint8)
    calls = []

    def cb(it, count, u, vx, vy, vz):
        calls.append(it)
        return it >= 20

    ctx.set_energy_interval(5)
    p = _params(native, 60, "l1", callback_interval=10)
    ctx.tgv2_deshade_masked(f, p, active, cb)
    trace = ctx.energy_trace()
    ctx.set_energy_interval(0)
    assert calls == [10, 20]
    # the callback of iteration index 20 ran after 21 iterations and stopped the loop
    assert [int(c) for c in trace[:, 0]] == [5, 10, 15, 20, 21]
    # reference: the masked loop run for c iterations, then the standalone pass
    ref_rows = []
    for c in (5, 10, 15, 20, 21):
        p2 = _params(native, c, "l1")
        ctx.tgv2_deshade_masked(f, p2, active)
        ref_rows.append(ctx.energy())
    for row, ref in zip(trace, ref_rows):
        assert _close(row[1:], ref, abs(ref[0])), (row, ref)


def test_trace_planar_uses_standalone_pass(ctx):
    """The 2-D loop (depth 1, planar) has no energy form: every row comes from
    the standalone pass, equal to amb_tgv2_energy of the same state."""
    from ambient_image_processor_amd import native
    f = _volume((1, 40, 64), seed=9)
    p = _params(native, 12, "l1", flags=native.FLAG_PLANAR)
    ctx.set_energy_interval(4)
    ctx.tgv2_deshade(f, p)
    trace = ctx.energy_trace()
    ctx.set_energy_interval(0)
    assert [int(c) for c in trace[:, 0]] == [4, 8, 12]
    ctx.tgv2_deshade(f, p)
    ref = ctx.energy()
    assert _close(trace[-1, 1:], ref, abs(ref[0]))


def test_trace_256_headline_configuration(ctx):
    """The 256^3 TGV2-L1 deshade configuration, 100 iterations with the energy
    every 25: the in-sweep rows equal the standalone pass, and the gap falls."""
    from ambient_image_processor_amd import native, phantoms
    _say("generating the 256^3 workload")
    w = phantoms.workload("deshade_256")
    f = w["f"]

    def params(it):
        return native.default_params(lam=w["lam"], alpha0=w["alpha0"], alpha1=w["alpha1"],
                                     tau=w["tau"], sigma=w["tau"], theta=1.0, iterations=it,
                                     data_term=native.DATA_L1)

    ctx.set_energy_interval(25)
    ctx.tgv2_deshade(f, params(100))
    trace = ctx.energy_trace()
    ctx.set_energy_interval(0)
    assert [int(c) for c in trace[:, 0]] == [25, 50, 75, 100]
    gaps = trace[:, 5]
    _say("256^3 gaps", gaps)
    assert np.all(np.diff(gaps) < 0)
    ctx.begin(f, params(0))
    ctx.iterate(75)
    ref = ctx.energy()
    assert _close(trace[2, 1:], ref, abs(ref[0])), (trace[2], ref)
