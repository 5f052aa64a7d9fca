"""ctypes binding of the CPU oracle (oracle/libtgv_oracle.so).

TEST INFRASTRUCTURE ONLY: used by tests/, by __graft_entry__.smoke() as the
checker, and by bench.py's cpu_baseline leg.  The product package never
imports this module.

Every function takes and returns numpy arrays shaped (D, H, W) (x fastest).
``dtype`` selects the mode: float32 = the reference's FP32 operation order
("mirror"), float64 = FP64 throughout.  See tgv_oracle.cpp for the reference
file:line each routine follows.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libtgv_oracle.so")
_lib = None

N_FIELDS = 17   # u, ubar, v[3], vbar[3], p[3], q[6]  (q: xx, yy, zz, xy, xz, yz)
FIELD_NAMES = ["u", "ubar", "vx", "vy", "vz", "vbx", "vby", "vbz", "px", "py", "pz",
               "qxx", "qyy", "qzz", "qxy", "qxz", "qyz"]


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (g++, OpenMP)."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-C", _HERE, "-B" if force else "-s"], check=True,
                       stdout=subprocess.DEVNULL)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = ctypes.CDLL(_LIB_PATH)
        _lib.orc_cospi.restype = ctypes.c_double
        _lib.orc_cospi.argtypes = [ctypes.c_double]
        _lib.orc_num_threads.restype = ctypes.c_int
    return _lib


def _suf(dtype):
    dt = np.dtype(dtype)
    if dt == np.float32:
        return "f32"
    if dt == np.float64:
        return "f64"
    raise TypeError(dtype)


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _c(a, dtype):
    return np.ascontiguousarray(a, dtype=dtype)


def _dims(a):
    d, h, w = a.shape
    return ctypes.c_longlong(w), ctypes.c_longlong(h), ctypes.c_longlong(d)


def set_num_threads(n: int):
    lib().orc_set_num_threads(int(n))


def num_threads() -> int:
    return lib().orc_num_threads()


def cospi(t: float) -> float:
    return lib().orc_cospi(float(t))


def bf16_rne(a):
    """The oracle's bfloat16 store rounding (nearest even) of float32 values,
    returned as float32."""
    a = np.ascontiguousarray(a, dtype=np.float32)
    out = np.empty_like(a)
    lib().orc_bf16_rne(_p(a), _p(out), ctypes.c_longlong(a.size))
    return out


def forward_difference(a, axis, dtype=np.float64):
    a = _c(a, dtype)
    out = np.empty_like(a)
    getattr(lib(), "orc_forward_difference_" + _suf(dtype))(_p(a), _p(out), *_dims(a), int(axis))
    return out


def backward_difference(a, axis, dtype=np.float64):
    a = _c(a, dtype)
    out = np.empty_like(a)
    getattr(lib(), "orc_backward_difference_" + _suf(dtype))(_p(a), _p(out), *_dims(a), int(axis))
    return out


class TGVState:
    """The 17 iteration fields as one contiguous (17, D, H, W) array."""

    def __init__(self, f, dtype=np.float64, bf16=False):
        """bf16: BF16 field storage - FP32 arithmetic, every field store rounded
        to bfloat16 (nearest even) where the reference's kernels store it."""
        self.dtype = np.dtype(dtype)
        if bf16 and self.dtype != np.float32:
            raise TypeError("BF16 storage runs FP32 arithmetic (dtype float32)")
        self.bf16 = bool(bf16)
        self.f = _c(f, dtype)
        self.fields = np.zeros((N_FIELDS,) + self.f.shape, dtype=dtype)
        self._ptrs()
        suf = "bf16" if self.bf16 else _suf(dtype)
        getattr(lib(), "orc_tgv2_init_" + suf)(self._pp, _p(self.f), *_dims(self.f))

    def _ptrs(self):
        n = self.fields[0].size * self.dtype.itemsize
        base = self.fields.ctypes.data
        self._pp = (ctypes.c_void_p * N_FIELDS)(*[base + i * n for i in range(N_FIELDS)])

    def __getitem__(self, name):
        return self.fields[FIELD_NAMES.index(name)]

    @property
    def u(self):
        return self.fields[0]

    @property
    def v(self):
        return self.fields[2:5]

    @property
    def p(self):
        return self.fields[8:11]

    @property
    def q(self):
        return self.fields[11:17]

    def copy(self):
        c = TGVState.__new__(TGVState)
        c.dtype = self.dtype
        c.bf16 = self.bf16
        c.f = self.f.copy()
        c.fields = self.fields.copy()
        c._ptrs()
        return c

    def _planar(self, planar):
        if planar and self.f.shape[0] != 1:
            raise ValueError("the planar (2-D) loop needs depth 1")
        return int(bool(planar))

    def iterate(self, iterations, lam, alpha0, alpha1, tau, sigma, theta=1.0, data_term="l2",
                mask=None, planar=False):
        """mask (optional): updates only where mask >= 1 (the reference's masked deshade).
        planar: the reference's 2-D loop at depth 1 (tgv_deshade/tgv2_l1_deshade_2d.cu)."""
        pl = self._planar(planar)
        prm = (ctypes.c_double * 6)(lam, alpha0, alpha1, tau, sigma, theta)
        dt = 0 if data_term == "l1" else 1
        act = None
        if mask is not None:
            act = np.ascontiguousarray(np.asarray(mask) >= 1, dtype=np.uint8)
        suf = "bf16" if self.bf16 else _suf(self.dtype)
        getattr(lib(), "orc_tgv2_iterate_" + suf)(
            self._pp, _p(self.f), *_dims(self.f), prm, dt, int(iterations),
            _p(act) if act is not None else None, pl)
        return self

    def energy(self, lam, alpha0, alpha1, data_term="l2", planar=False):
        """FP64 energy/gap: [E, E_data, E_first, E_second, gap, b1, b2, b3,
        infeasibility, v_residual]."""
        prm = (ctypes.c_double * 6)(lam, alpha0, alpha1, 0.0, 0.0, 0.0)
        out = (ctypes.c_double * 10)()
        getattr(lib(), "orc_tgv2_energy_" + _suf(self.dtype))(
            _p(self.f), self._pp, *_dims(self.f), prm, 0 if data_term == "l1" else 1,
            self._planar(planar), out)
        return np.array(out[:])


def tgv2(f, iterations, lam, alpha0, alpha1, tau, sigma, theta=1.0, data_term="l2",
         dtype=np.float64, mask=None, planar=False, bf16=False):
    """Run the reference loop from the reference initialisation; returns the state.
    With a mask: the masked loop (updates only where mask >= 1).  planar: the
    2-D loop at depth 1.  bf16: BF16 field storage (dtype float32)."""
    st = TGVState(f, dtype, bf16=bf16)
    st.iterate(iterations, lam, alpha0, alpha1, tau, sigma, theta, data_term, mask, planar)
    return st


class TGVKState:
    """TGV-k (k >= 3) iteration state: the 17 TGV2 fields, then per level
    i < k - 2 the tensors w_i, w_bar_i, r_i (components x, y, z, xy, xz, yz),
    as one contiguous (17 + 18 (k-2), D, H, W) array."""

    def __init__(self, f, order, dtype=np.float64):
        if order < 3:
            raise ValueError("TGV-k needs order >= 3")
        self.order = int(order)
        self.dtype = np.dtype(dtype)
        self.f = _c(f, dtype)
        nf = N_FIELDS + 18 * (self.order - 2)
        self.fields = np.zeros((nf,) + self.f.shape, dtype=dtype)
        n = self.fields[0].size * self.dtype.itemsize
        base = self.fields.ctypes.data
        self._pp = (ctypes.c_void_p * nf)(*[base + i * n for i in range(nf)])
        getattr(lib(), "orc_tgv2_init_" + _suf(dtype))(self._pp, _p(self.f), *_dims(self.f))

    @property
    def u(self):
        return self.fields[0]

    @property
    def v(self):
        return self.fields[2:5]

    @property
    def p(self):
        return self.fields[8:11]

    @property
    def q(self):
        return self.fields[11:17]

    def level(self, i):
        """(w, w_bar, r) of level i, each (6, D, H, W)."""
        b = N_FIELDS + 18 * i
        return self.fields[b:b + 6], self.fields[b + 6:b + 12], self.fields[b + 12:b + 18]

    def iterate(self, iterations, lam, alpha, tau, sigma, theta=1.0, planar=False, mask=None):
        """alpha: order weights; alpha[0] bounds p, alpha[1] q, alpha[i + 2] r_i.
        mask (optional): updates only where mask >= 1 (the masked TGV-k loop)."""
        alpha = np.ascontiguousarray(alpha, dtype=np.float64)
        if alpha.size != self.order:
            raise ValueError("alpha needs one weight per order")
        if planar and self.f.shape[0] != 1:
            raise ValueError("the planar (2-D) loop needs depth 1")
        prm = (ctypes.c_double * 4)(lam, tau, sigma, theta)
        act = None
        if mask is not None:
            act = np.ascontiguousarray(np.asarray(mask) >= 1, dtype=np.uint8)
        rc = getattr(lib(), "orc_tgvk_iterate_" + _suf(self.dtype))(
            self._pp, self.order, _p(self.f), *_dims(self.f), prm, _p(alpha), int(iterations),
            int(bool(planar)), _p(act) if act is not None else None)
        assert rc == iterations
        return self


def tensor_op(t, forward, halve):
    """The TGV-k level operators on a (6, D, H, W) symmetric tensor (x, y, z, xy,
    xz, yz): halve=True is G (w_bar -> r2), halve=False is H (r -> r2); forward
    selects D+ instead of the reference's backward difference."""
    t = np.ascontiguousarray(t, dtype=np.float64)
    out = np.empty_like(t)
    pin = (ctypes.c_void_p * 6)(*[t[c].ctypes.data for c in range(6)])
    pout = (ctypes.c_void_p * 6)(*[out[c].ctypes.data for c in range(6)])
    lib().orc_tensor_op_f64(pin, pout, *_dims(t[0]), int(forward), int(halve))
    return out


def tgvk(f, iterations, lam, alpha, tau, sigma, theta=1.0, dtype=np.float64, planar=False,
         mask=None):
    """TGV-k (k = len(alpha) >= 3) deshade loop from the reference initialisation;
    with a mask the masked loop (updates only where mask >= 1)."""
    st = TGVKState(f, len(alpha), dtype)
    st.iterate(iterations, lam, alpha, tau, sigma, theta, planar, mask)
    return st


def deshade_k(f, iterations, lam, alpha, tau, sigma, theta=1.0, mask=None,
              set_negative_zero=True, add_background_back=False, dtype=np.float32):
    """TGVKDeshadeProcessor::processTGVKL1Cuda (TGVKDeshadeProcessor.cpp:76-174):
    order 2 runs the TGV2 loop with alpha0 = alpha[1], alpha1 = alpha[0]; order
    >= 3 the TGV-k loop; depth 1 the 2-D loops.  Then the Poisson integration
    of v, l and r times the mask, clamp, optional add-back of u*(1 - mask).
    Returns (u, l, r, state)."""
    planar = np.shape(f)[0] == 1
    if len(alpha) == 2:
        st = tgv2(f, iterations, lam, alpha[1], alpha[0], tau, sigma, theta, "l1", dtype,
                  planar=planar)
    else:
        st = tgvk(f, iterations, lam, alpha, tau, sigma, theta, dtype, planar=planar)
    l = integrate_poisson(st.v[0], st.v[1], st.v[2], False, dtype)
    r, l2 = deshade_epilogue(st.u, l, mask, True, set_negative_zero, add_background_back)
    return st.u.copy(), l2, r, st


def deshade_k_masked(f, iterations, lam, alpha, tau, sigma, mask=None, theta=1.0,
                     set_negative_zero=True, add_background_back=False, dtype=np.float32):
    """TGVKDeshadeMaskedProcessor::processTGVKL1Cuda (TGVKDeshadeMaskedProcessor.cpp:76-226):
    order 2 is the masked TGV2 composition (deshade_masked with alpha0 = alpha[1],
    alpha1 = alpha[0]); order >= 3 the masked TGV-k loop on the unmasked f (2-D at
    depth 1), the Poisson integration, l and r times the mask, clamp, then the
    add-back of u*(1 - mask) or r times the mask again.  No mask: all ones.
    Returns (u, l, r, state)."""
    f = np.asarray(f)
    m = np.ones(f.shape, np.float32) if mask is None else np.asarray(mask, np.float32)
    if len(alpha) == 2:
        return deshade_masked(f, iterations, lam, alpha[1], alpha[0], tau, sigma, m, theta,
                              set_negative_zero, add_background_back, dtype=dtype)
    st = tgvk(f, iterations, lam, alpha, tau, sigma, theta, dtype, planar=f.shape[0] == 1,
              mask=m)
    l = integrate_poisson(st.v[0], st.v[1], st.v[2], False, dtype)
    r, l2 = deshade_epilogue(st.u, l, m, True, set_negative_zero, add_background_back,
                             mask_again=True)
    return st.u.copy(), l2, r, st


def spline3_coefficients(a):
    """Cubic B-spline coefficients (FP64) of a float32 (D, H, W) volume, the
    itk::BSplineDecompositionImageFilter restatement."""
    a = _c(a, np.float32)
    c = np.empty(a.shape, np.float64)
    lib().orc_spline3_coefficients(_p(a), _p(c), *_dims(a))
    return c


def resize_spline3(a, out_shape):
    """ResizeProcessor::process(image, width, height, depth, BSpline3): cubic
    B-spline resampling of a (D, H, W) volume to out_shape (D', H', W')."""
    a = _c(a, np.float32)
    d, h, w = (int(v) for v in out_shape)
    out = np.empty((d, h, w), np.float32)
    lib().orc_resize_spline3(_p(a), *_dims(a), _p(out), ctypes.c_longlong(w),
                             ctypes.c_longlong(h), ctypes.c_longlong(d))
    return out


def resize_extents(shape, factor):
    """ResizeProcessor::process(image, size_factor, ...) target extents
    (resize/ResizeProcessor.cpp:41-52): max(1, ceil(float(extent) * factor))."""
    f32 = np.float32(factor)
    return tuple(max(1, int(np.ceil(np.float32(n) * f32))) for n in shape)


def deshade_k_downsampled(f, factor, iterations, lam, alpha, tau, sigma, mask=None, theta=1.0,
                          set_negative_zero=True, add_background_back=False, masked=True):
    """TGVKDeshadeDownsampledProcessor::processTGVKL1Cuda
    (tgvk_deshade_downsampled/TGVKDeshadeDownsampledProcessor.cpp:29-129; masked=False:
    TGVKDeshadeDownsampledNotMaskedProcessor.cpp:29-120): resize f (and the mask) by
    the factor with cubic B-splines, deshade there (the masked TGV-k processor, or the
    unmasked one), resize u and l back, r = f - l with the mask, add-back and clamp.
    Returns (u, l, r)."""
    f = _c(f, np.float32)
    m = None if mask is None else _c(mask, np.float32)
    down = (lambda a: a) if factor == 1 else \
        (lambda a: resize_spline3(a, resize_extents(a.shape, factor)))
    up = (lambda a: a) if factor == 1 else (lambda a: resize_spline3(a, f.shape))
    if masked:
        background = None
        if add_background_back and m is not None:
            background = f * (np.float32(1) - m)   # invert.cu: 1 - mask, multiply.cu
        fm = f if m is None else f * m
        fd = down(fm)
        md = None if m is None else down(m)
        ud, ld, _, _ = deshade_k_masked(fd, iterations, lam, alpha, tau, sigma, md, theta,
                                        set_negative_zero, add_background_back)
        u, l = up(ud), up(ld)
        r = fm - l
        if m is not None:
            l = l * m
            r = r * m
        if background is not None:
            r = r + background
    else:
        fd = down(f)
        md = None if m is None else down(m)
        ud, ld, _, _ = deshade_k(fd, iterations, lam, alpha, tau, sigma, theta, md,
                                 set_negative_zero, add_background_back)
        u, l = up(ud), up(ld)
        r = f - l
        if m is not None:
            r = r * m
        if add_background_back and m is not None:
            r = r + f * (np.float32(1) - m)
    if set_negative_zero:
        r = np.where(r < 0, np.float32(0), r)
    return u, l, r.astype(np.float32)


def divergence(vx, vy, vz, dtype=np.float64):
    vx, vy, vz = (_c(a, dtype) for a in (vx, vy, vz))
    out = np.empty_like(vx)
    getattr(lib(), "orc_divergence_" + _suf(dtype))(_p(vx), _p(vy), _p(vz), _p(out), *_dims(vx))
    return out


def cosine_transform(a, dtype=np.float64):
    a = _c(a, dtype)
    out = np.empty_like(a)
    getattr(lib(), "orc_cosine_transform_" + _suf(dtype))(_p(a), _p(out), *_dims(a))
    return out


def inverse_cosine_transform(a, dtype=np.float64):
    a = _c(a, dtype)
    out = np.empty_like(a)
    getattr(lib(), "orc_inverse_cosine_transform_" + _suf(dtype))(_p(a), _p(out), *_dims(a))
    return out


def dct_pass(a, axis, inverse=False, last=False, dtype=np.float64):
    a = _c(a, dtype)
    out = np.empty_like(a)
    getattr(lib(), "orc_dct_pass_" + _suf(dtype))(_p(a), _p(out), *_dims(a), int(axis),
                                                  int(inverse), int(last))
    return out


def solve_poisson(a, index_quirk=False, dtype=np.float64):
    a = _c(a, dtype)
    out = np.empty_like(a)
    getattr(lib(), "orc_solve_poisson_" + _suf(dtype))(_p(a), _p(out), *_dims(a), int(index_quirk))
    return out


def integrate_poisson(vx, vy, vz, index_quirk=False, dtype=np.float64):
    vx, vy, vz = (_c(a, dtype) for a in (vx, vy, vz))
    out = np.empty_like(vx)
    getattr(lib(), "orc_integrate_poisson_" + _suf(dtype))(_p(vx), _p(vy), _p(vz), _p(out),
                                                          *_dims(vx), int(index_quirk))
    return out


def deshade_epilogue(u, l, mask=None, mask_shading=False, set_negative_zero=True,
                     add_background_back=False, mask_again=False):
    """Returns (r, l_out) in FP32 (reference epilogue op order)."""
    u = _c(u, np.float32)
    l = _c(l, np.float32).copy()
    r = np.empty_like(u)
    m = _c(mask, np.float32) if mask is not None else None
    lib().orc_deshade_epilogue_f32(_p(u), _p(l), _p(m) if m is not None else None, _p(r),
                                   ctypes.c_longlong(u.size), int(mask_shading),
                                   int(set_negative_zero), int(add_background_back),
                                   int(mask_again))
    return r, l


def deshade(f, iterations, lam, alpha0, alpha1, tau, sigma, theta=1.0, data_term="l1",
            mask=None, set_negative_zero=True, index_quirk=False, dtype=np.float32,
            planar=False):
    """Full reference pipeline (CLI overload, TGVDeshadeProcessor.cpp:122-166):
    TGV2 iteration -> Poisson integration of v -> r = u - l (mask, clamp).
    planar: the 2-D loop at depth 1.  Returns (u, l, r, state)."""
    st = tgv2(f, iterations, lam, alpha0, alpha1, tau, sigma, theta, data_term, dtype,
              planar=planar)
    l = integrate_poisson(st.v[0], st.v[1], st.v[2], index_quirk, dtype)
    r, l2 = deshade_epilogue(st.u, l, mask, False, set_negative_zero, False)
    return st.u.copy(), l2, r, st


def deshade_masked(f, iterations, lam, alpha0, alpha1, tau, sigma, mask, theta=1.0,
                   set_negative_zero=True, add_background_back=False, index_quirk=False,
                   dtype=np.float32):
    """Masked deshade composition (TGVDeshadeMaskedProcessor.cpp:76-210): the
    masked TGV2-L1 loop on the unmasked f (updates only where mask >= 1), the
    Poisson integration of v, then l and r times the mask, clamp, and either
    the add-back of u*(1 - mask) or r times the mask once more.  A depth-1
    image takes the 2-D loop (TGVDeshadeMaskedProcessor.cpp:93-98).
    Returns (u, l, r, state)."""
    st = tgv2(f, iterations, lam, alpha0, alpha1, tau, sigma, theta, "l1", dtype, mask=mask,
              planar=np.shape(f)[0] == 1)
    l = integrate_poisson(st.v[0], st.v[1], st.v[2], index_quirk, dtype)
    r, l2 = deshade_epilogue(st.u, l, mask, True, set_negative_zero, add_background_back,
                             mask_again=True)
    return st.u.copy(), l2, r, st
