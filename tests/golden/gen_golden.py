"""Writes tests/golden/golden.npz: known-answer vectors for the oracle.

  dct_*   : scipy.fft.dctn / idctn (type 2, norm=None) of seeded FP64 volumes.
            The reference's separable DCT (cosine_transform.cu:142-223 with the
            *8 of the z pass) equals dctn(type=2, norm=None); its inverse
            (inverse_cosine_transform.cu:79-164 with /(W*H*D*8)) equals idctn.
  lap_*   : a zero-mean field of DCT-II modes and its Neumann Laplacian,
            computed by explicit neighbour sums (the operator of
            conjugate_gradient/image_vector_operations_kernels.cu:181-238,
            sign flipped) - the Poisson solve must return the field.
  sten_*  : forward/backward differences of an integer-valued 4x3x2 volume
            from dense matrices built entry by entry from the reference's
            definitions (tgv_cuda/tgv_common.cu:66-197).
Run:  python tests/golden/gen_golden.py
"""
import os

import numpy as np
import scipy.fft as sf

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden.npz")


def lap_neighbour_sums(phi):
    d, h, w = phi.shape
    out = np.zeros_like(phi)
    for z in range(d):
        for y in range(h):
            for x in range(w):
                s = 0.0
                for (dz, dy, dx) in ((1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)):
                    zz, yy, xx = z + dz, y + dy, x + dx
                    if 0 <= zz < d and 0 <= yy < h and 0 <= xx < w:
                        s += phi[z, y, x] - phi[zz, yy, xx]
                out[z, y, x] = s
    return out


def dense_diff(n, kind):
    M = np.zeros((n, n))
    for i in range(n):
        if kind == "fwd":
            if i < n - 1:
                M[i, i + 1] = 1.0
                M[i, i] = -1.0
        else:
            M[i, i] = -1.0
            if i >= 1:
                M[i, i - 1] = 1.0
    return M


def apply_axis(M, a, axis):
    return np.moveaxis(np.tensordot(M, np.moveaxis(a, axis, 0), axes=(1, 0)), 0, axis)


def main():
    rng = np.random.default_rng(20160101)
    g = {}
    for name, shape in (("a", (5, 6, 8)), ("b", (16, 16, 16)), ("c", (20, 24, 32)), ("d", (7, 1, 9))):
        x = rng.standard_normal(shape)
        g[f"dct_{name}_in"] = x
        g[f"dct_{name}_out"] = sf.dctn(x, type=2, norm=None)
        g[f"idct_{name}_out"] = sf.idctn(x, type=2, norm=None)
    # Poisson known answer on 12x10x14 (D, H, W)
    d, h, w = 12, 10, 14
    phi = np.zeros((d, h, w))
    for _ in range(6):
        k = rng.integers(0, 5, size=3)
        if not k.any():
            k[0] = 1
        a = rng.standard_normal()
        cz = np.cos(np.pi * k[0] * (np.arange(d) + 0.5) / d)
        cy = np.cos(np.pi * k[1] * (np.arange(h) + 0.5) / h)
        cx = np.cos(np.pi * k[2] * (np.arange(w) + 0.5) / w)
        phi += a * cz[:, None, None] * cy[None, :, None] * cx[None, None, :]
    g["lap_phi"] = phi
    g["lap_rhs"] = lap_neighbour_sums(phi)
    # stencils on an integer volume (D=2, H=3, W=4)
    a = rng.integers(-9, 10, size=(2, 3, 4)).astype(np.float64)
    g["sten_in"] = a
    for axis_name, axis in (("x", 2), ("y", 1), ("z", 0)):
        n = a.shape[axis]
        g[f"sten_fwd_{axis_name}"] = apply_axis(dense_diff(n, "fwd"), a, axis)
        g[f"sten_bwd_{axis_name}"] = apply_axis(dense_diff(n, "bwd"), a, axis)
    np.savez_compressed(OUT, **g)
    print("wrote", OUT, sorted(g))


if __name__ == "__main__":
    main()
