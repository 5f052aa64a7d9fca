"""BF16 field storage rounds every stored value to nearest, ties to even.

A state of chosen FP32 values goes in through amb_tgv2_set_state of a context
with BF16 storage (the device converts FP32 -> BF16) and comes back through
amb_tgv2_get_state (BF16 -> FP32, exact).  The result must equal

* hand-worked cases: ties at 1 (to even, both directions), just above a tie,
  the largest finite value overflowing to infinity, denormal ties and the
  smallest FP32 denormal rounding to a signed zero;
* an integer round-to-nearest-even of the FP32 bits, (u + 0x7fff + bit 16 of
  u) >> 16, on every sign/exponent class with random mantissas and on the
  mantissa patterns around each tie.

NaN inputs are not part of the contract (a NaN stays a NaN)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SHAPE = (2, 4, 64)


def _bits(a):
    return np.asarray(a, dtype=np.float32).view(np.uint32)


def _rne(x):
    u = _bits(x).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) & 0xFFFF
    return (r.astype(np.uint32) << 16).view(np.float32)


def _roundtrip(ctx, values):
    """Store `values` (any length) through the 17 BF16 fields and read them back."""
    from ambient_image_processor_amd import native
    n_per = int(np.prod(SHAPE))
    total = native.N_FIELDS * n_per
    vals = np.asarray(values, dtype=np.float32).ravel()
    assert vals.size <= total
    buf = np.zeros(total, dtype=np.float32)
    buf[:vals.size] = vals
    f = np.full(SHAPE, 0.5, dtype=np.float32)
    p = native.default_params(iterations=1, storage=native.STORAGE_BF16)
    ctx.begin(f, p)
    ctx.set_state(buf.reshape((native.N_FIELDS,) + SHAPE))
    out = ctx.get_state().ravel()
    return out[:vals.size]


def test_bf16_rounding_hand_cases(ctx):
    tiny = np.float32(2.0 ** -133)            # smallest BF16 denormal
    cases = [
        (1 + 2.0 ** -8, 1.0),                   # tie, 1 has an even mantissa: down
        (1 + 3 * 2.0 ** -8, 1 + 2.0 ** -6),     # tie between odd 1+2^-7 and even 1+2^-6: up
        (1 + 2.0 ** -8 + 2.0 ** -20, 1 + 2.0 ** -7),   # just above the tie: up
        (-(1 + 2.0 ** -8), -1.0),
        (np.finfo(np.float32).max, np.inf),     # beyond the largest BF16 + half an ulp
        (-np.finfo(np.float32).max, -np.inf),
        (np.float32(2.0 ** -149), 0.0),         # smallest FP32 denormal
        (np.float32(-2.0 ** -149), -0.0),
        (tiny, tiny),
        (np.float32(3 * 2.0 ** -134), np.float32(2.0 ** -132)),   # denormal tie to even
        (np.float32(2.0 ** -134), 0.0),         # denormal tie with zero: to even (zero)
        (-0.0, -0.0),
        (np.inf, np.inf),
        (-np.inf, -np.inf),
    ]
    got = _roundtrip(ctx, [c[0] for c in cases])
    want = np.array([c[1] for c in cases], dtype=np.float32)
    np.testing.assert_array_equal(_bits(got), _bits(want))


def test_bf16_rounding_matches_integer_form(ctx):
    rng = np.random.default_rng(7)
    from ambient_image_processor_amd import native
    total = native.N_FIELDS * int(np.prod(SHAPE))
    # every exponent (finite and infinite), both signs, random mantissas
    expo = rng.integers(0, 256, total // 2, dtype=np.uint32)
    mant = rng.integers(0, 1 << 23, total // 2, dtype=np.uint32)
    sign = rng.integers(0, 2, total // 2, dtype=np.uint32)
    mant[expo == 255] = 0                      # infinities only, no NaN
    a = (sign << 31) | (expo << 23) | mant
    # the low 16 bits at and around the tie (0x7fff, 0x8000, 0x8001) with both
    # parities of bit 16, over random high bits
    hi = rng.integers(0, 1 << 16, total - a.size, dtype=np.uint32)
    hi[(hi & 0x7F80) == 0x7F80] &= 0xFF7F      # keep the exponent finite
    lo = np.array([0x7FFF, 0x8000, 0x8001, 0x0000, 0xFFFF], dtype=np.uint32)
    b = (hi << 16) | lo[np.arange(hi.size) % lo.size]
    x = np.concatenate([a, b]).view(np.float32)
    assert not np.isnan(x).any()
    got = _roundtrip(ctx, x)
    np.testing.assert_array_equal(_bits(got), _bits(_rne(x)))
