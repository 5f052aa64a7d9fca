"""ctypes binding of libambient_amd.so (the C ABI in include/ambient_amd.h).

The product path has no CPU fallback: if the HIP library is not built, or no
GPU is visible, every entry point raises.  Volumes are numpy float32 arrays of
shape (D, H, W) (x fastest); device buffers are raw pointers (ints) obtained
from Context.alloc().
"""
from __future__ import annotations

import ctypes
import os
from ctypes import POINTER, c_double, c_float, c_int, c_size_t, c_uint32, c_uint64, c_void_p

import numpy as np

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libambient_amd.so")

MEM_HOST, MEM_DEVICE = 0, 1
DATA_L1, DATA_L2 = 0, 1
STORAGE_F32, STORAGE_BF16 = 0, 1
DCT_REFERENCE, DCT_F32, DCT_BF16X3 = 0, 1, 2
FLAG_SET_NEGATIVE_ZERO = 0x1
FLAG_MASK_SHADING = 0x2
FLAG_ADD_BACKGROUND_BACK = 0x4
FLAG_POISSON_INDEX_QUIRK = 0x8
FLAG_MASKED_ITERATION = 0x10
FLAG_PLANAR = 0x20
N_FIELDS = 17
FIELD_NAMES = ["u", "ubar", "vx", "vy", "vz", "vbx", "vby", "vbz", "px", "py", "pz",
               "qxx", "qyy", "qzz", "qxy", "qxz", "qyz"]
PROFILE_CLASSES = ["dual_sweep", "primal_sweep", "dct", "poisson", "divergence", "epilogue",
                   "init", "energy", "adjoint", "fused_sweep", "tgvk_dual", "tgvk_primal",
                   "resize", "tgvk_fused", "fused_sweep_energy"]

# Every function include/ambient_amd.h declares (checked by the CPU test suite).
EXPORTS = [
    "amb_context_create", "amb_context_destroy", "amb_last_error", "amb_version",
    "amb_tgv_params_default", "amb_device_alloc", "amb_device_free", "amb_memcpy",
    "amb_synchronize", "amb_context_stream", "amb_tgv2_deshade", "amb_tgv2_denoise",
    "amb_divergence", "amb_cosine_transform", "amb_inverse_cosine_transform",
    "amb_solve_poisson_in_cosine_domain", "amb_integrate_gradients_poisson",
    "amb_tgv2_deshade_pipeline", "amb_tgv2_energy", "amb_tgv2_begin", "amb_tgv2_set_state",
    "amb_tgv2_get_state", "amb_tgv2_iterate", "amb_tgv2_sweep", "amb_profile_enable",
    "amb_profile_read", "amb_profile_reset", "amb_tgv2_denoise_adjoint",
    "amb_tgv2_deshade_masked", "amb_deshade_epilogue", "amb_tgvk_deshade",
    "amb_tgvk_deshade_pipeline", "amb_tgvk_deshade_masked", "amb_resize_bspline3",
    "amb_resize_extents", "amb_tgvk_deshade_downsampled", "amb_tgv2_set_energy_interval",
    "amb_tgv2_energy_trace", "amb_context_set_plan_override", "amb_context_clear_plan_overrides",
    "amb_tgv2_energy_noalloc", "amb_profile_count", "amb_profile_fused_tiles",
]


class AmbError(RuntimeError):
    def __init__(self, status, message):
        super().__init__(f"amb status {status}: {message}")
        self.status = status


class Params(ctypes.Structure):
    _fields_ = [("lambda_", c_float), ("alpha0", c_float), ("alpha1", c_float),
                ("tau", c_float), ("sigma", c_float), ("theta", c_float),
                ("iterations", c_uint32), ("data_term", c_uint32), ("storage", c_uint32),
                ("dct_precision", c_uint32), ("callback_interval", c_uint32),
                ("flags", c_uint32)]


CALLBACK = ctypes.CFUNCTYPE(c_int, c_uint32, c_uint32, POINTER(c_float), POINTER(c_float),
                            POINTER(c_float), POINTER(c_float), c_void_p)

_lib = None


def lib():
    """Load the HIP library; raises if it has not been built (no fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                          "(hipcc --offload-arch=gfx950); there is no CPU fallback")
    L = ctypes.CDLL(LIB_PATH)
    vp, u32, mem = c_void_p, c_uint32, c_int
    fp = c_void_p   # float* passed as void*
    sig = {
        "amb_context_create": [c_int, POINTER(c_void_p)],
        "amb_context_destroy": [vp],
        "amb_tgv_params_default": [POINTER(Params)],
        "amb_device_alloc": [vp, c_size_t, POINTER(c_void_p)],
        "amb_device_free": [vp, vp],
        "amb_memcpy": [vp, vp, vp, c_size_t, c_int],
        "amb_synchronize": [vp],
        "amb_context_stream": [vp],
        "amb_tgv2_deshade": [vp, mem, fp, u32, u32, u32, POINTER(Params), fp, fp, fp, fp,
                             CALLBACK, vp],
        "amb_tgv2_denoise": [vp, mem, fp, u32, u32, u32, POINTER(Params), fp, CALLBACK, vp],
        "amb_deshade_epilogue": [vp, mem, fp, fp, fp, u32, u32, u32, u32, fp, fp],
        "amb_tgv2_deshade_masked": [vp, mem, fp, u32, u32, u32, POINTER(Params), vp, fp, fp, fp,
                                    fp, CALLBACK, vp],
        "amb_divergence": [vp, mem, fp, fp, fp, u32, u32, u32, fp],
        "amb_cosine_transform": [vp, mem, fp, u32, u32, u32, u32, fp],
        "amb_inverse_cosine_transform": [vp, mem, fp, u32, u32, u32, u32, fp],
        "amb_solve_poisson_in_cosine_domain": [vp, mem, fp, u32, u32, u32, u32, fp],
        "amb_integrate_gradients_poisson": [vp, mem, fp, fp, fp, u32, u32, u32, u32, u32, fp],
        "amb_tgv2_deshade_pipeline": [vp, mem, fp, u32, u32, u32, POINTER(Params), fp, fp, fp,
                                      fp, CALLBACK, vp],
        "amb_tgvk_deshade": [vp, mem, fp, u32, u32, u32, POINTER(Params), u32, fp, fp, fp, fp,
                             fp, CALLBACK, vp],
        "amb_tgvk_deshade_pipeline": [vp, mem, fp, u32, u32, u32, POINTER(Params), u32, fp, fp,
                                      fp, fp, fp, CALLBACK, vp],
        "amb_tgvk_deshade_masked": [vp, mem, fp, u32, u32, u32, POINTER(Params), u32, fp, vp, fp,
                                    fp, fp, fp, CALLBACK, vp],
        "amb_resize_bspline3": [vp, mem, fp, u32, u32, u32, fp, u32, u32, u32],
        "amb_resize_extents": [u32, u32, u32, c_float, POINTER(c_uint32)],
        "amb_tgvk_deshade_downsampled": [vp, mem, fp, u32, u32, u32, POINTER(Params), c_float,
                                         u32, fp, fp, fp, fp, fp, CALLBACK, vp],
        "amb_tgv2_energy": [vp, POINTER(c_double)],
        "amb_tgv2_set_energy_interval": [vp, u32],
        "amb_context_set_plan_override": [vp, ctypes.c_char_p, ctypes.c_int32],
        "amb_context_clear_plan_overrides": [vp],
        "amb_tgv2_energy_trace": [vp, vp, u32, POINTER(c_uint32)],
        "amb_tgv2_energy_noalloc": [vp, POINTER(c_void_p), u32, u32, u32, POINTER(Params), vp, vp],
        "amb_tgv2_begin": [vp, mem, fp, u32, u32, u32, POINTER(Params)],
        "amb_tgv2_set_state": [vp, mem, POINTER(c_void_p)],
        "amb_tgv2_get_state": [vp, mem, POINTER(c_void_p)],
        "amb_tgv2_iterate": [vp, u32],
        "amb_tgv2_sweep": [vp, c_int],
        "amb_profile_enable": [vp, c_int],
        "amb_profile_read": [vp, c_int, POINTER(c_double), POINTER(c_uint64)],
        "amb_profile_reset": [vp],
        "amb_profile_count": [vp, c_int, POINTER(c_uint64)],
        "amb_profile_fused_tiles": [vp, POINTER(c_uint64), POINTER(c_uint64)],
        "amb_tgv2_denoise_adjoint": [vp, mem, fp, u32, u32, u32, POINTER(Params), fp, fp, fp,
                                     POINTER(c_double), u32],
    }
    for name, args in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = c_void_p if name == "amb_context_stream" else (
            None if name in ("amb_tgv_params_default", "amb_resize_extents") else c_int)
    L.amb_last_error.restype = ctypes.c_char_p
    L.amb_version.restype = ctypes.c_char_p
    _lib = L
    return L


def _check(status):
    if status != 0:
        raise AmbError(status, lib().amb_last_error().decode())


def default_params(**kw) -> Params:
    p = Params()
    lib().amb_tgv_params_default(ctypes.byref(p))
    for k, v in kw.items():
        setattr(p, "lambda_" if k in ("lam", "lambda") else k, v)
    return p


def resize_extents(shape, factor):
    """amb_resize_extents for a (D, H, W) shape -> (D', H', W')."""
    d, h, w = (int(v) for v in shape)
    out = (c_uint32 * 3)()
    lib().amb_resize_extents(w, h, d, float(factor), out)
    return int(out[2]), int(out[1]), int(out[0])


def _f32(a):
    a = np.ascontiguousarray(a, dtype=np.float32)
    return a


def _ptr(a):
    if a is None:
        return None
    if isinstance(a, int):
        return c_void_p(a)
    return a.ctypes.data_as(c_void_p)


def _dims(shape):
    d, h, w = shape
    return c_uint32(w), c_uint32(h), c_uint32(d)


class Context:
    """An amb_context: device state + stream on one GPU."""

    def __init__(self, device: int = 0):
        self._h = c_void_p()
        _check(lib().amb_context_create(device, ctypes.byref(self._h)))

    def close(self):
        if self._h:
            lib().amb_context_destroy(self._h)
            self._h = c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- device memory ----------------------------------------------------
    def alloc(self, nbytes: int) -> int:
        p = c_void_p()
        _check(lib().amb_device_alloc(self._h, nbytes, ctypes.byref(p)))
        return p.value

    def free(self, ptr: int):
        _check(lib().amb_device_free(self._h, c_void_p(ptr)))

    def upload(self, ptr: int, a: np.ndarray):
        a = np.ascontiguousarray(a)
        _check(lib().amb_memcpy(self._h, c_void_p(ptr), _ptr(a), a.nbytes, 0))

    def download(self, ptr: int, shape, dtype=np.float32) -> np.ndarray:
        out = np.empty(shape, dtype=dtype)
        _check(lib().amb_memcpy(self._h, _ptr(out), c_void_p(ptr), out.nbytes, 1))
        return out

    def synchronize(self):
        _check(lib().amb_synchronize(self._h))

    # ---- TGV2 iteration ---------------------------------------------------
    @staticmethod
    def _callback(fn, shape, with_v):
        if fn is None:
            return CALLBACK(), None
        n = int(np.prod(shape))

        def tramp(it, count, u, vx, vy, vz, user):
            U = np.ctypeslib.as_array(u, (n,)).reshape(shape).copy()
            if with_v:
                V = [np.ctypeslib.as_array(p, (n,)).reshape(shape).copy() for p in (vx, vy, vz)]
                return int(bool(fn(it, count, U, *V)))
            return int(bool(fn(it, count, U)))
        cb = CALLBACK(tramp)
        return cb, cb

    def tgv2_deshade(self, f, params: Params, callback=None):
        """amb_tgv2_deshade on host arrays -> (u, vx, vy, vz)."""
        f = _f32(f)
        self.shape = f.shape
        outs = [np.empty_like(f) for _ in range(4)]
        cb, keep = self._callback(callback, f.shape, True)
        _check(lib().amb_tgv2_deshade(self._h, MEM_HOST, _ptr(f), *_dims(f.shape),
                                      ctypes.byref(params), *[_ptr(o) for o in outs], cb, None))
        del keep
        return tuple(outs)

    def tgv2_deshade_masked(self, f, params: Params, active, callback=None):
        """amb_tgv2_deshade_masked on host arrays; active: nonzero = update
        -> (u, vx, vy, vz)."""
        f = _f32(f)
        act = np.ascontiguousarray(np.asarray(active) != 0, dtype=np.uint8)
        if act.shape != f.shape:
            raise ValueError("active must have the shape of f")
        self.shape = f.shape
        outs = [np.empty_like(f) for _ in range(4)]
        cb, keep = self._callback(callback, f.shape, True)
        _check(lib().amb_tgv2_deshade_masked(self._h, MEM_HOST, _ptr(f), *_dims(f.shape),
                                             ctypes.byref(params), _ptr(act),
                                             *[_ptr(o) for o in outs], cb, None))
        del keep
        return tuple(outs)

    def tgvk_deshade(self, f, params: Params, alpha, callback=None):
        """amb_tgvk_deshade (order = len(alpha) >= 3) on host arrays -> (u, vx, vy, vz)."""
        f = _f32(f)
        al = np.ascontiguousarray(alpha, dtype=np.float32)
        self.shape = f.shape
        outs = [np.empty_like(f) for _ in range(4)]
        cb, keep = self._callback(callback, f.shape, True)
        _check(lib().amb_tgvk_deshade(self._h, MEM_HOST, _ptr(f), *_dims(f.shape),
                                      ctypes.byref(params), al.size, _ptr(al),
                                      *[_ptr(o) for o in outs], cb, None))
        del keep
        return tuple(outs)

    def tgvk_deshade_masked(self, f, params: Params, alpha, active, callback=None):
        """amb_tgvk_deshade_masked on host arrays; active: nonzero = update
        -> (u, vx, vy, vz)."""
        f = _f32(f)
        al = np.ascontiguousarray(alpha, dtype=np.float32)
        act = np.ascontiguousarray(np.asarray(active) != 0, dtype=np.uint8)
        if act.shape != f.shape:
            raise ValueError("active must have the shape of f")
        self.shape = f.shape
        outs = [np.empty_like(f) for _ in range(4)]
        cb, keep = self._callback(callback, f.shape, True)
        _check(lib().amb_tgvk_deshade_masked(self._h, MEM_HOST, _ptr(f), *_dims(f.shape),
                                             ctypes.byref(params), al.size, _ptr(al), _ptr(act),
                                             *[_ptr(o) for o in outs], cb, None))
        del keep
        return tuple(outs)

    def resize_bspline3(self, a, out_shape):
        """amb_resize_bspline3 on host arrays: (D, H, W) -> out_shape (D', H', W')."""
        a = _f32(a)
        d, h, w = (int(v) for v in out_shape)
        out = np.empty((d, h, w), np.float32)
        _check(lib().amb_resize_bspline3(self._h, MEM_HOST, _ptr(a), *_dims(a.shape), _ptr(out),
                                         w, h, d))
        return out

    def tgvk_deshade_downsampled(self, f, params: Params, factor, alpha, mask=None,
                                 callback=None):
        """amb_tgvk_deshade_downsampled on host arrays -> (u, shading, deshaded)."""
        f = _f32(f)
        al = np.ascontiguousarray(alpha, dtype=np.float32)
        m = _f32(mask) if mask is not None else None
        u, l, r = (np.empty_like(f) for _ in range(3))
        small = resize_extents(f.shape, factor) if factor != 1 else f.shape
        self.shape = small   # the iteration state lives on the resized volume
        cb, keep = self._callback(callback, small, True)
        _check(lib().amb_tgvk_deshade_downsampled(self._h, MEM_HOST, _ptr(f), *_dims(f.shape),
                                                  ctypes.byref(params), float(factor), al.size,
                                                  _ptr(al), _ptr(m), _ptr(u), _ptr(l), _ptr(r),
                                                  cb, None))
        del keep
        return u, l, r

    def tgvk_deshade_downsampled_device(self, f_dev: int, shape, params: Params, factor, alpha,
                                        mask_dev=None, u_dev=None, l_dev=None, r_dev=None):
        al = np.ascontiguousarray(alpha, dtype=np.float32)
        self.shape = resize_extents(shape, factor) if factor != 1 else tuple(shape)
        _check(lib().amb_tgvk_deshade_downsampled(self._h, MEM_DEVICE, c_void_p(f_dev),
                                                  *_dims(shape), ctypes.byref(params),
                                                  float(factor), al.size, _ptr(al),
                                                  _ptr(mask_dev), _ptr(u_dev), _ptr(l_dev),
                                                  _ptr(r_dev), CALLBACK(), None))

    def tgvk_deshade_pipeline(self, f, params: Params, alpha, mask=None, callback=None):
        """amb_tgvk_deshade_pipeline (order = len(alpha) >= 2) on host arrays
        -> (u, shading, deshaded)."""
        f = _f32(f)
        al = np.ascontiguousarray(alpha, dtype=np.float32)
        self.shape = f.shape
        m = _f32(mask) if mask is not None else None
        u, l, r = (np.empty_like(f) for _ in range(3))
        cb, keep = self._callback(callback, f.shape, True)
        _check(lib().amb_tgvk_deshade_pipeline(self._h, MEM_HOST, _ptr(f), *_dims(f.shape),
                                               ctypes.byref(params), al.size, _ptr(al), _ptr(m),
                                               _ptr(u), _ptr(l), _ptr(r), cb, None))
        del keep
        return u, l, r

    def tgv2_denoise(self, f, params: Params, callback=None):
        f = _f32(f)
        u = np.empty_like(f)
        cb, keep = self._callback(callback, f.shape, False)
        _check(lib().amb_tgv2_denoise(self._h, MEM_HOST, _ptr(f), *_dims(f.shape),
                                      ctypes.byref(params), _ptr(u), cb, None))
        del keep
        return u

    def deshade_pipeline(self, f, params: Params, mask=None, callback=None):
        """amb_tgv2_deshade_pipeline on host arrays -> (u, shading, deshaded)."""
        f = _f32(f)
        self.shape = f.shape
        m = _f32(mask) if mask is not None else None
        u, l, r = (np.empty_like(f) for _ in range(3))
        cb, keep = self._callback(callback, f.shape, True)
        _check(lib().amb_tgv2_deshade_pipeline(self._h, MEM_HOST, _ptr(f), *_dims(f.shape),
                                               ctypes.byref(params), _ptr(m), _ptr(u), _ptr(l),
                                               _ptr(r), cb, None))
        del keep
        return u, l, r

    def deshade_epilogue(self, u, l, mask=None, flags=FLAG_SET_NEGATIVE_ZERO):
        """amb_deshade_epilogue on host arrays -> (r, l_out)."""
        u, l = _f32(u), _f32(l)
        m = _f32(mask) if mask is not None else None
        r, lo = np.empty_like(u), np.empty_like(u)
        _check(lib().amb_deshade_epilogue(self._h, MEM_HOST, _ptr(u), _ptr(l), _ptr(m),
                                          *_dims(u.shape), int(flags), _ptr(lo), _ptr(r)))
        return r, lo

    def deshade_pipeline_device(self, f_dev: int, shape, params: Params, u_dev=None, l_dev=None,
                                r_dev=None, mask_dev=None):
        _check(lib().amb_tgv2_deshade_pipeline(self._h, MEM_DEVICE, c_void_p(f_dev), *_dims(shape),
                                               ctypes.byref(params), _ptr(mask_dev), _ptr(u_dev),
                                               _ptr(l_dev), _ptr(r_dev), CALLBACK(), None))

    def tgvk_deshade_pipeline_device(self, f_dev: int, shape, params: Params, alpha, u_dev=None,
                                     l_dev=None, r_dev=None, mask_dev=None):
        al = np.ascontiguousarray(alpha, dtype=np.float32)
        _check(lib().amb_tgvk_deshade_pipeline(self._h, MEM_DEVICE, c_void_p(f_dev), *_dims(shape),
                                               ctypes.byref(params), al.size, _ptr(al),
                                               _ptr(mask_dev), _ptr(u_dev), _ptr(l_dev),
                                               _ptr(r_dev), CALLBACK(), None))

    def begin(self, f, params: Params):
        f = _f32(f)
        self.shape = f.shape
        _check(lib().amb_tgv2_begin(self._h, MEM_HOST, _ptr(f), *_dims(f.shape), ctypes.byref(params)))

    def begin_device(self, f_dev: int, shape, params: Params):
        self.shape = tuple(shape)
        _check(lib().amb_tgv2_begin(self._h, MEM_DEVICE, c_void_p(f_dev), *_dims(shape),
                                    ctypes.byref(params)))

    def set_state(self, fields: np.ndarray):
        fields = np.ascontiguousarray(fields, dtype=np.float32)
        assert fields.shape[0] == N_FIELDS
        arr = (c_void_p * N_FIELDS)(*[fields[k].ctypes.data for k in range(N_FIELDS)])
        _check(lib().amb_tgv2_set_state(self._h, MEM_HOST, arr))

    def get_state(self) -> np.ndarray:
        out = np.empty((N_FIELDS,) + tuple(self.shape), dtype=np.float32)
        arr = (c_void_p * N_FIELDS)(*[out[k].ctypes.data for k in range(N_FIELDS)])
        _check(lib().amb_tgv2_get_state(self._h, MEM_HOST, arr))
        return out

    def iterate(self, n: int):
        _check(lib().amb_tgv2_iterate(self._h, n))

    def sweep(self, which: int):
        _check(lib().amb_tgv2_sweep(self._h, which))

    def energy(self) -> np.ndarray:
        out = (c_double * 10)()
        _check(lib().amb_tgv2_energy(self._h, out))
        return np.array(out[:])

    def set_override(self, name: str, value: int):
        """amb_context_set_plan_override (tests, tuning): value < 0 unsets; applies
        from the next call that begins an iteration."""
        _check(lib().amb_context_set_plan_override(self._h, name.encode(), int(value)))

    def clear_overrides(self):
        _check(lib().amb_context_clear_plan_overrides(self._h))

    def set_energy_interval(self, interval: int):
        """amb_tgv2_set_energy_interval: energy/gap every `interval` iterations of
        each iteration call (and after its last), reduced inside the fused sweep
        where it runs; 0 = off."""
        _check(lib().amb_tgv2_set_energy_interval(self._h, c_uint32(int(interval))))

    def energy_trace(self) -> np.ndarray:
        """Rows of the last iteration call: [iteration count, the ten amb_tgv2_energy
        values] -> array (rows, 11)."""
        rows = c_uint32(0)
        _check(lib().amb_tgv2_energy_trace(self._h, None, c_uint32(0), ctypes.byref(rows)))
        out = np.zeros((rows.value, 11), dtype=np.float64)
        if rows.value:
            _check(lib().amb_tgv2_energy_trace(self._h, out.ctypes.data_as(c_void_p),
                                               c_uint32(rows.value), ctypes.byref(rows)))
        return out

    # ---- DCT Poisson ------------------------------------------------------
    def divergence(self, vx, vy, vz):
        vx, vy, vz = _f32(vx), _f32(vy), _f32(vz)
        out = np.empty_like(vx)
        _check(lib().amb_divergence(self._h, MEM_HOST, _ptr(vx), _ptr(vy), _ptr(vz),
                                    *_dims(vx.shape), _ptr(out)))
        return out

    def cosine_transform(self, a, precision=DCT_REFERENCE):
        a = _f32(a)
        out = np.empty_like(a)
        _check(lib().amb_cosine_transform(self._h, MEM_HOST, _ptr(a), *_dims(a.shape), precision,
                                          _ptr(out)))
        return out

    def inverse_cosine_transform(self, a, precision=DCT_REFERENCE):
        a = _f32(a)
        out = np.empty_like(a)
        _check(lib().amb_inverse_cosine_transform(self._h, MEM_HOST, _ptr(a), *_dims(a.shape),
                                                  precision, _ptr(out)))
        return out

    def solve_poisson(self, a, index_quirk=False):
        a = _f32(a)
        out = np.empty_like(a)
        _check(lib().amb_solve_poisson_in_cosine_domain(
            self._h, MEM_HOST, _ptr(a), *_dims(a.shape),
            FLAG_POISSON_INDEX_QUIRK if index_quirk else 0, _ptr(out)))
        return out

    def integrate_gradients_poisson(self, vx, vy, vz, precision=DCT_REFERENCE, flags=0):
        vx, vy, vz = _f32(vx), _f32(vy), _f32(vz)
        out = np.empty_like(vx)
        _check(lib().amb_integrate_gradients_poisson(self._h, MEM_HOST, _ptr(vx), _ptr(vy),
                                                     _ptr(vz), *_dims(vx.shape), precision, flags,
                                                     _ptr(out)))
        return out

    def integrate_gradients_poisson_device(self, vx, vy, vz, shape, out, precision=DCT_REFERENCE,
                                           flags=0):
        _check(lib().amb_integrate_gradients_poisson(self._h, MEM_DEVICE, c_void_p(vx),
                                                     c_void_p(vy), c_void_p(vz), *_dims(shape),
                                                     precision, flags, c_void_p(out)))

    # ---- adjoint ------------------------------------------------------------
    def tgv2_denoise_adjoint(self, f, params: Params, dL_du, checkpoint_every=20):
        """amb_tgv2_denoise_adjoint on host arrays -> (u, dL/df, (dL/dalpha0, dL/dalpha1,
        dL/dlambda))."""
        f, g = _f32(f), _f32(dL_du)
        u, gf = np.empty_like(f), np.empty_like(f)
        gp = (c_double * 3)()
        _check(lib().amb_tgv2_denoise_adjoint(self._h, MEM_HOST, _ptr(f), *_dims(f.shape),
                                              ctypes.byref(params), _ptr(g), _ptr(u), _ptr(gf),
                                              gp, checkpoint_every))
        return u, gf, np.array(gp[:])

    def tgv2_denoise_adjoint_device(self, f_dev, shape, params: Params, g_dev, u_dev, gf_dev,
                                    checkpoint_every=20):
        gp = (c_double * 3)()
        _check(lib().amb_tgv2_denoise_adjoint(self._h, MEM_DEVICE, c_void_p(f_dev), *_dims(shape),
                                              ctypes.byref(params), c_void_p(g_dev),
                                              _ptr(u_dev), c_void_p(gf_dev), gp, checkpoint_every))
        return np.array(gp[:])

    # ---- profiling ----------------------------------------------------------
    def profile(self, enable=True, every=1):
        """Per-launch HIP-event timing; every = k >= 2 times one launch in k per class."""
        _check(lib().amb_profile_enable(self._h, (int(every) if every > 1 else 1) if enable else 0))

    def profile_reset(self):
        _check(lib().amb_profile_reset(self._h))

    def profile_read(self):
        """{class: (ms of the timed launches, timed launches)}"""
        res = {}
        for k, name in enumerate(PROFILE_CLASSES):
            ms, n = c_double(), c_uint64()
            _check(lib().amb_profile_read(self._h, k, ctypes.byref(ms), ctypes.byref(n)))
            res[name] = (ms.value, n.value)
        return res

    def profile_counts(self):
        """{class: launches since the last reset, timed or not}"""
        res = {}
        for k, name in enumerate(PROFILE_CLASSES):
            n = c_uint64()
            _check(lib().amb_profile_count(self._h, k, ctypes.byref(n)))
            res[name] = n.value
        return res

    def profile_fused_tiles(self):
        """(tiles swept, tiles in the grids) of the 3-D fused sweep since the last reset"""
        a, b = c_uint64(), c_uint64()
        _check(lib().amb_profile_fused_tiles(self._h, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value
