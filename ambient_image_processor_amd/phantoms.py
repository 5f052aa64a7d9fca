"""Deterministic synthetic volumes for the TGV-DCT deshade workloads.

The reference ships no data (its ``data/`` folder is git-ignored), so every
workload here is synthetic and seeded.  Random numbers come from SplitMix64
(Steele, Lea & Flood 2014) with uniform doubles built from the top 53 bits and
normal deviates from the Box-Muller transform in FP64, so the same seed gives
the same bytes on every host.  Volumes are float32 arrays of shape
``(depth, height, width)`` in the reference's x-fastest layout
(``idx = z*W*H + y*W + x``, reference source/modules/itk/ITKImage.cpp:426-429).

Workloads (sizes and seeds):
  * ``deshade_64``      64^3, 8 boxes on 0.1 background + cosine bias (order <= 2)
                        + N(0, 0.02^2) noise                              (seed 1)
  * ``deshade_256``     256^3, 32 spheres + DCT bias (order <= 3, N(0, 0.05^2))
                        + N(0, 0.03^2) noise                              (seed 2)
  * ``poisson_512``     512x512x256 smooth field of 64 cosine modes; the
                        right-hand side is its Neumann Laplacian          (seed 3)
  * ``deshade_384``     384^3 spheres, bias rescaled to max |b| = 0.2,
                        noise 0.02                                         (seed 4)
  * ``denoise_128x4``   4 x 128^3 piecewise-constant volumes (12 boxes) +
                        N(0, 0.05^2)                                       (seeds 5..8)
"""
from __future__ import annotations

import numpy as np

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


class SplitMix64:
    """SplitMix64 stream with vectorised draws (wrapping uint64 arithmetic)."""

    def __init__(self, seed: int):
        self.state = np.uint64(seed & 0xFFFFFFFFFFFFFFFF)

    def next_u64(self, n: int) -> np.ndarray:
        with np.errstate(over="ignore"):
            k = np.arange(1, n + 1, dtype=np.uint64)
            z = self.state + k * _GOLDEN
            self.state = self.state + np.uint64(n) * _GOLDEN
            z = (z ^ (z >> np.uint64(30))) * _M1
            z = (z ^ (z >> np.uint64(27))) * _M2
            return z ^ (z >> np.uint64(31))

    def uniform(self, n: int, lo: float = 0.0, hi: float = 1.0) -> np.ndarray:
        u = (self.next_u64(n) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
        return lo + (hi - lo) * u

    def integers(self, n: int, lo: int, hi_inclusive: int) -> np.ndarray:
        u = self.uniform(n)
        return (lo + np.floor(u * (hi_inclusive - lo + 1))).astype(np.int64)

    def normal(self, n: int, sigma: float = 1.0) -> np.ndarray:
        m = (n + 1) // 2
        u1 = self.uniform(m)
        u2 = self.uniform(m)
        r = np.sqrt(-2.0 * np.log1p(-u1))
        z = np.empty(2 * m, dtype=np.float64)
        z[0::2] = r * np.cos(2.0 * np.pi * u2)
        z[1::2] = r * np.sin(2.0 * np.pi * u2)
        return sigma * z[:n]


def cosine_basis(n: int, k: int) -> np.ndarray:
    """c_k(x) = cos(pi k (x + 1/2) / n), the DCT-II basis vector of order k."""
    x = np.arange(n, dtype=np.float64)
    return np.cos(np.pi * k * (x + 0.5) / n)


def separable_modes(shape, coeffs: dict) -> np.ndarray:
    """sum_{(i,j,k)} a_ijk c_i(x) c_j(y) c_k(z) as float64 array (D, H, W),
    evaluated as one (D*H, M) x (M, W) matrix product over the M modes."""
    d, h, w = shape
    if not coeffs:
        return np.zeros((d, h, w), dtype=np.float64)
    keys = list(coeffs.keys())
    a = np.array([coeffs[k] for k in keys], dtype=np.float64)
    cx = np.stack([cosine_basis(w, i) for (i, _, _) in keys], axis=1)        # (W, M)
    cy = np.stack([cosine_basis(h, j) for (_, j, _) in keys], axis=1)        # (H, M)
    cz = np.stack([cosine_basis(d, k) for (_, _, k) in keys], axis=1)        # (D, M)
    t = (cz[:, None, :] * cy[None, :, :] * a[None, None, :]).reshape(d * h, len(keys))
    return (t @ cx.T).reshape(d, h, w)


def _boxes(rng: SplitMix64, shape, count, lo_max, size_range, inten, background):
    d, h, w = shape
    vol = np.full(shape, background, dtype=np.float64)
    for _ in range(count):
        lo = rng.integers(3, 0, lo_max)
        sz = rng.integers(3, size_range[0], size_range[1])
        val = rng.uniform(1, inten[0], inten[1])[0]
        x0, y0, z0 = lo
        x1, y1, z1 = np.minimum(lo + sz, [w, h, d])
        vol[z0:z1, y0:y1, x0:x1] = val
    return vol


def _spheres(rng: SplitMix64, shape, count, r_range, inten):
    d, h, w = shape
    vol = np.zeros(shape, dtype=np.float64)
    for _ in range(count):
        c = rng.uniform(3, 0.0, 1.0) * np.array([w, h, d], dtype=np.float64)
        r = rng.uniform(1, r_range[0], r_range[1])[0]
        val = rng.uniform(1, inten[0], inten[1])[0]
        lo = np.maximum(np.floor(c - r).astype(int), 0)
        hi = np.minimum(np.ceil(c + r).astype(int) + 1, [w, h, d])
        zz, yy, xx = np.meshgrid(np.arange(lo[2], hi[2]), np.arange(lo[1], hi[1]),
                                 np.arange(lo[0], hi[0]), indexing="ij")
        inside = (xx - c[0]) ** 2 + (yy - c[1]) ** 2 + (zz - c[2]) ** 2 <= r * r
        sub = vol[lo[2]:hi[2], lo[1]:hi[1], lo[0]:hi[0]]
        sub[inside] = val
    return vol


def _dct_bias(rng: SplitMix64, shape, order, sigma=None, amp=None):
    coeffs = {}
    for k in range(order + 1):
        for j in range(order + 1):
            for i in range(order + 1):
                if i == j == k == 0:
                    continue
                if sigma is not None:
                    coeffs[(i, j, k)] = rng.normal(1, sigma)[0]
                else:
                    coeffs[(i, j, k)] = rng.uniform(1, -amp, amp)[0]
    return separable_modes(shape, coeffs)


def deshade_phantom(shape=(64, 64, 64), seed=1, kind="boxes", noise=0.02,
                    bias_order=2, bias_amp=0.1, bias_sigma=None, bias_max=None,
                    n_objects=None):
    """Return (f, clean, bias) float32 volumes of ``shape`` = (D, H, W).

    kind="boxes": 8 axis-aligned boxes, corners U{0..W-9}, sizes U{4..24},
    intensities U[0.2, 1] on a 0.1 background, later boxes overwrite earlier.
    kind="spheres": 32 spheres, centre U[0, n)^3, radius U[n/32, 3n/16),
    intensities U[0.1, 1] on background 0.
    Bias: cosine modes up to ``bias_order`` per axis without the DC term, with
    amplitudes U[-bias_amp, bias_amp] (or N(0, bias_sigma^2)); optionally
    rescaled to max |b| = bias_max.  f = clean + bias + N(0, noise^2).
    """
    rng = SplitMix64(seed)
    d, h, w = shape
    if kind == "boxes":
        clean = _boxes(rng, shape, n_objects or 8, max(w - 9, 0), (4, 24), (0.2, 1.0), 0.1)
    elif kind == "spheres":
        n = max(shape)
        clean = _spheres(rng, shape, n_objects or 32, (n / 32.0, 3.0 * n / 16.0), (0.1, 1.0))
    else:
        raise ValueError(kind)
    bias = _dct_bias(rng, shape, bias_order, sigma=bias_sigma, amp=bias_amp)
    if bias_max is not None:
        bias *= bias_max / np.max(np.abs(bias))
    noise_v = rng.normal(d * h * w, noise).reshape(shape)
    f = clean + bias + noise_v
    return f.astype(np.float32), clean.astype(np.float32), bias.astype(np.float32)


def neumann_laplacian(phi: np.ndarray) -> np.ndarray:
    """(-Delta_N) phi with the reference's Neumann semantics: sum over axes of
    (D+)^T D+ phi, i.e. 2*phi - neighbours with mirrored missing neighbours.
    Same operator as reference conjugate_gradient/image_vector_operations_kernels.cu:181-238
    (with the opposite sign convention); the DCT diagonalises it."""
    phi = np.asarray(phi, dtype=np.float64)
    out = np.zeros_like(phi)
    for ax in range(3):
        b = np.moveaxis(np.zeros_like(phi), ax, 0)
        p = np.moveaxis(phi, ax, 0)
        b[:-1] = p[1:] - p[:-1]          # D+ phi, 0 at the last index
        t = -b.copy()                    # (D+)^T b [j] = b[j-1] [j >= 1] - b[j]
        t[1:] += b[:-1]
        out += np.moveaxis(t, 0, ax)     # (-Delta_N) = sum over axes of (D+)^T D+
    return out


def poisson_field(shape=(256, 512, 512), seed=3, modes=64, kmax=15):
    """Return (rhs, phi) with phi = sum of ``modes`` cosine modes a ~ N(0,1),
    k per axis U{0..kmax}, (0,0,0) excluded, and rhs = (-Delta_N) phi computed
    in closed form from the mode eigenvalues 6 - 2cos(pi kx/W) - 2cos(pi ky/H)
    - 2cos(pi kz/D).  phi is zero-mean, so the DCT Poisson solve returns phi."""
    rng = SplitMix64(seed)
    d, h, w = shape
    coeffs = {}
    eig = {}
    drawn = 0
    while drawn < modes:
        k = rng.integers(3, 0, kmax)
        if k[0] == 0 and k[1] == 0 and k[2] == 0:
            continue
        a = rng.normal(1, 1.0)[0]
        key = (int(k[0]), int(k[1]), int(k[2]))
        coeffs[key] = coeffs.get(key, 0.0) + a
        eig[key] = (6.0 - 2.0 * np.cos(np.pi * key[0] / w) - 2.0 * np.cos(np.pi * key[1] / h)
                    - 2.0 * np.cos(np.pi * key[2] / d))
        drawn += 1
    phi = separable_modes(shape, coeffs)
    rhs = separable_modes(shape, {k: a * eig[k] for k, a in coeffs.items()})
    return rhs.astype(np.float32), phi


def workload(name: str):
    """Named workloads; returns a dict with the input volume(s) and parameters."""
    if name == "deshade_64":
        f, clean, bias = deshade_phantom((64, 64, 64), seed=1, kind="boxes", noise=0.02,
                                         bias_order=2, bias_amp=0.1)
        return dict(f=f, clean=clean, bias=bias, data_term="l2", alpha0=2.0, alpha1=1.0,
                    lam=1.0, tau=1.0 / np.sqrt(12.0), sigma=1.0 / np.sqrt(12.0), theta=1.0,
                    iterations=500)
    if name == "deshade_256":
        f, clean, bias = deshade_phantom((256, 256, 256), seed=2, kind="spheres", noise=0.03,
                                         bias_order=3, bias_sigma=0.05)
        return dict(f=f, clean=clean, bias=bias, data_term="l1", alpha0=2.0, alpha1=1.0,
                    lam=0.5, tau=1.0 / np.sqrt(20.0), sigma=1.0 / np.sqrt(20.0), theta=1.0,
                    iterations=1000)
    if name == "deshade_384":
        f, clean, bias = deshade_phantom((384, 384, 384), seed=4, kind="spheres", noise=0.02,
                                         bias_order=3, bias_sigma=0.05, bias_max=0.2)
        return dict(f=f, clean=clean, bias=bias, data_term="l2", alpha0=2.0, alpha1=1.0,
                    lam=1.0, tau=1.0 / np.sqrt(20.0), sigma=1.0 / np.sqrt(20.0), theta=1.0,
                    iterations=800)
    if name == "poisson_512":
        rhs, phi = poisson_field((256, 512, 512), seed=3)
        return dict(rhs=rhs, phi=phi)
    if name == "denoise_128x4":
        vols, cleans = [], []
        for s in range(5, 9):
            rng = SplitMix64(s)
            clean = _boxes(rng, (128, 128, 128), 12, 119, (8, 48), (0.1, 1.0), 0.0)
            f = clean + rng.normal(128 ** 3, 0.05).reshape(clean.shape)
            vols.append(f.astype(np.float32))
            cleans.append(clean.astype(np.float32))
        return dict(f=np.stack(vols), clean=np.stack(cleans), data_term="l2", alpha0=2.0,
                    alpha1=1.0, lam=1.0, tau=1.0 / np.sqrt(20.0), sigma=1.0 / np.sqrt(20.0),
                    theta=1.0, iterations=200, checkpoint_every=20)
    raise KeyError(name)
