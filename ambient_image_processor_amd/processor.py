"""Host-side mirror of the reference's launcher and processor interface for the
TGV-DCT deshade path, on top of the gfx950 C ABI.

Same names, argument order and meaning as the reference's C++ templates
(float instantiation).  Differences: volumes are numpy float32 arrays (flat
``W*H*D`` like the reference's ``Pixel*``, or shaped (D, H, W)); results are
returned instead of ``new Pixel[]``; a failing HIP call raises AmbError
instead of calling exit() (reference source/modules/cuda/cuda_helper.in:11-22).
tau = sigma = 1/sqrt(20) and theta = 1 as the reference hard-codes
(tgv_deshade/tgv2_l1_deshade.cu:148-150), overridable by keyword.
"""
from __future__ import annotations

import threading

import numpy as np

from . import native

_tls = threading.local()


def _ctx() -> native.Context:
    c = getattr(_tls, "ctx", None)
    if c is None:
        c = native.Context(0)
        _tls.ctx = c
    return c


def _vol(a, width, height, depth):
    a = np.ascontiguousarray(a, dtype=np.float32)
    return a.reshape(depth, height, width)


def _params(lambda_, iteration_count, paint_iteration_interval, alpha0, alpha1, data_term, **kw):
    p = native.default_params(lam=lambda_, alpha0=alpha0, alpha1=alpha1,
                              iterations=iteration_count, data_term=data_term,
                              callback_interval=paint_iteration_interval or 0)
    for k, v in kw.items():
        setattr(p, "lambda_" if k == "lam" else k, v)
    return p


def tgv2_l1_deshade_launch(f_host, width, height, depth, lambda_, iteration_count,
                           paint_iteration_interval, cuda_block_dimension,
                           iteration_finished_callback, alpha0, alpha1, **kw):
    """Reference tgv_deshade/tgv2_l1_deshade.cu:33-41 -> (u, v_x, v_y, v_z).
    cuda_block_dimension is accepted and ignored (launch shapes are fixed per
    kernel on gfx950).  The callback gets (iteration_index, iteration_count,
    u, v_x, v_y, v_z) and returns True to stop.  ``data_term=native.DATA_L2``
    gives the L2 deshade loop (configurations that the reference composes
    from tgv2_l2.cu's u-update)."""
    del cuda_block_dimension
    data_term = kw.pop("data_term", native.DATA_L1)
    f = _vol(f_host, width, height, depth)
    p = _params(lambda_, iteration_count, paint_iteration_interval, alpha0, alpha1, data_term, **kw)
    return _ctx().tgv2_deshade(f, p, iteration_finished_callback)


def tgv2_l1_deshade_masked_launch(f_host, width, height, depth, lambda_, iteration_count,
                                  paint_iteration_interval, cuda_block_dimension,
                                  iteration_finished_callback, alpha0, alpha1,
                                  masked_pixel_indices, *edge_pixel_indices, **kw):
    """Reference tgv_deshade_masked/tgv2_l1_deshade_masked.cu:27-305 -> (u, v_x, v_y, v_z).
    Updates run only at the linear indices in masked_pixel_indices.  The twelve
    edge lists that follow (left, not_left, right, ..., back, not_back) are
    accepted for signature parity; the reference builds them from the volume
    faces (TGVDeshadeMaskedProcessor.cpp:355-383), which the sweeps handle
    themselves, so they carry no extra information."""
    del cuda_block_dimension, edge_pixel_indices
    f = _vol(f_host, width, height, depth)
    idx = np.asarray(masked_pixel_indices, dtype=np.int64).ravel()
    if idx.size and (idx.min() < 0 or idx.max() >= f.size):
        raise IndexError("masked pixel index outside the volume")
    active = np.zeros(f.size, np.uint8)
    active[idx] = 1
    p = _params(lambda_, iteration_count, paint_iteration_interval, alpha0, alpha1,
                native.DATA_L1, **kw)
    return _ctx().tgv2_deshade_masked(f, p, active.reshape(f.shape), iteration_finished_callback)


def _planar_params(p):
    p.flags |= native.FLAG_PLANAR
    return p


def _drop_vz(cb):
    """The 2-D callback signature (iteration_index, iteration_count, u, v_x, v_y)."""
    if cb is None:
        return None
    return lambda it, count, u, vx, vy, vz: cb(it, count, u, vx, vy)


def tgv2_l1_deshade_launch_2d(f_host, width, height, lambda_, iteration_count,
                              paint_iteration_interval, iteration_finished_callback, alpha0,
                              alpha1, **kw):
    """Reference tgv_deshade/tgv2_l1_deshade_2d.cu:27-215 -> (u, v_x, v_y), the
    2-D loop (no z backward differences in eps_B(v_bar)).  The callback gets
    (iteration_index, iteration_count, u, v_x, v_y)."""
    data_term = kw.pop("data_term", native.DATA_L1)
    f = _vol(f_host, width, height, 1)
    p = _planar_params(_params(lambda_, iteration_count, paint_iteration_interval, alpha0, alpha1,
                               data_term, **kw))
    u, vx, vy, _ = _ctx().tgv2_deshade(f, p, _drop_vz(iteration_finished_callback))
    return u, vx, vy


def tgv2_l1_deshade_masked_2d_launch(f_host, width, height, lambda_, iteration_count,
                                     paint_iteration_interval, cuda_block_dimension,
                                     iteration_finished_callback, alpha0, alpha1,
                                     masked_pixel_indices, *edge_pixel_indices, **kw):
    """Reference tgv_deshade_masked/tgv2_l1_deshade_masked_2d.cu (declared
    TGVDeshadeMaskedProcessor.cpp:53-68) -> (u, v_x, v_y): the masked 2-D
    loop.  The eight edge lists are accepted for signature parity, as in
    the 3-D launcher."""
    del cuda_block_dimension, edge_pixel_indices
    f = _vol(f_host, width, height, 1)
    idx = np.asarray(masked_pixel_indices, dtype=np.int64).ravel()
    if idx.size and (idx.min() < 0 or idx.max() >= f.size):
        raise IndexError("masked pixel index outside the volume")
    active = np.zeros(f.size, np.uint8)
    active[idx] = 1
    p = _planar_params(_params(lambda_, iteration_count, paint_iteration_interval, alpha0, alpha1,
                               native.DATA_L1, **kw))
    u, vx, vy, _ = _ctx().tgv2_deshade_masked(f, p, active.reshape(f.shape),
                                              _drop_vz(iteration_finished_callback))
    return u, vx, vy


def tgvk_l1_deshade_launch(f_host, width, height, depth, lambda_, iteration_count,
                           paint_iteration_interval, iteration_finished_callback, order, alpha,
                           **kw):
    """Reference tgvk_deshade/tgvk_l1_deshade.cu:35-45 (order >= 3) -> (u, v_x,
    v_y, v_z).  alpha: order weights (alpha[0] for p, alpha[1] for q, alpha[i+2]
    for the dual of level i).  The callback gets (iteration_index,
    iteration_count, u, v_x, v_y, v_z)."""
    if len(alpha) != order:
        raise ValueError("alpha needs one weight per order")
    f = _vol(f_host, width, height, depth)
    p = _params(lambda_, iteration_count, paint_iteration_interval, 1.0, 1.0, native.DATA_L1,
                **kw)
    return _ctx().tgvk_deshade(f, p, alpha, iteration_finished_callback)


def tgvk_l1_deshade_launch_2d(f_host, width, height, lambda_, iteration_count,
                              paint_iteration_interval, iteration_finished_callback, order, alpha,
                              **kw):
    """Reference tgvk_deshade/tgvk_l1_deshade_2d.cu (declared
    TGVKDeshadeProcessor.cpp:49-58) -> (u, v_x, v_y); the callback gets
    (iteration_index, iteration_count, u, v_x, v_y)."""
    if len(alpha) != order:
        raise ValueError("alpha needs one weight per order")
    f = _vol(f_host, width, height, 1)
    p = _planar_params(_params(lambda_, iteration_count, paint_iteration_interval, 1.0, 1.0,
                               native.DATA_L1, **kw))
    u, vx, vy, _ = _ctx().tgvk_deshade(f, p, alpha, _drop_vz(iteration_finished_callback))
    return u, vx, vy


def _active_from_indices(masked_pixel_indices, n):
    idx = np.asarray(masked_pixel_indices, dtype=np.int64).ravel()
    if idx.size and (idx.min() < 0 or idx.max() >= n):
        raise IndexError("masked pixel index outside the volume")
    active = np.zeros(n, np.uint8)
    active[idx] = 1
    return active


def tgvk_l1_deshade_masked_launch(f_host, width, height, depth, lambda_, iteration_count,
                                  paint_iteration_interval, cuda_block_dimension,
                                  iteration_finished_callback, order, alpha,
                                  masked_pixel_indices, *edge_pixel_indices, **kw):
    """Reference tgvk_deshade_masked/tgvk_l1_deshade_masked.cu:29-… (declared
    TGVKDeshadeMaskedProcessor.cpp:30-50) -> (u, v_x, v_y, v_z): the TGV-k loop
    updating only the voxels in masked_pixel_indices; the twelve edge lists are
    accepted for signature parity."""
    del cuda_block_dimension, edge_pixel_indices
    if len(alpha) != order:
        raise ValueError("alpha needs one weight per order")
    f = _vol(f_host, width, height, depth)
    active = _active_from_indices(masked_pixel_indices, f.size)
    p = _params(lambda_, iteration_count, paint_iteration_interval, 1.0, 1.0, native.DATA_L1,
                **kw)
    return _ctx().tgvk_deshade_masked(f, p, alpha, active.reshape(f.shape),
                                      iteration_finished_callback)


def tgvk_l1_deshade_masked_2d_launch(f_host, width, height, lambda_, iteration_count,
                                     paint_iteration_interval, cuda_block_dimension,
                                     iteration_finished_callback, order, alpha,
                                     masked_pixel_indices, *edge_pixel_indices, **kw):
    """Reference tgvk_deshade_masked/tgvk_l1_deshade_masked_2d.cu (declared
    TGVKDeshadeMaskedProcessor.cpp:53-68) -> (u, v_x, v_y)."""
    del cuda_block_dimension, edge_pixel_indices
    if len(alpha) != order:
        raise ValueError("alpha needs one weight per order")
    f = _vol(f_host, width, height, 1)
    active = _active_from_indices(masked_pixel_indices, f.size)
    p = _planar_params(_params(lambda_, iteration_count, paint_iteration_interval, 1.0, 1.0,
                               native.DATA_L1, **kw))
    u, vx, vy, _ = _ctx().tgvk_deshade_masked(f, p, alpha, active.reshape(f.shape),
                                              _drop_vz(iteration_finished_callback))
    return u, vx, vy


def _denoise(data_term, f_host, width, height, depth, lambda_, iteration_count,
             paint_iteration_interval, iteration_finished_callback, alpha0, alpha1, **kw):
    f = _vol(f_host, width, height, depth)
    p = _params(lambda_, iteration_count, paint_iteration_interval, alpha0, alpha1, data_term, **kw)
    return _ctx().tgv2_denoise(f, p, iteration_finished_callback)


def tgv2_l1_launch(f_host, width, height, depth, lambda_, iteration_count,
                   paint_iteration_interval, iteration_finished_callback, alpha0, alpha1, **kw):
    """Reference tgv_cuda/tgv2_l1.cu:64-71 -> u."""
    return _denoise(native.DATA_L1, f_host, width, height, depth, lambda_, iteration_count,
                    paint_iteration_interval, iteration_finished_callback, alpha0, alpha1, **kw)


def tgv2_l2_launch(f_host, width, height, depth, lambda_, iteration_count,
                   paint_iteration_interval, iteration_finished_callback, alpha0, alpha1, **kw):
    """Reference tgv_cuda/tgv2_l2.cu:60-67 -> u."""
    return _denoise(native.DATA_L2, f_host, width, height, depth, lambda_, iteration_count,
                    paint_iteration_interval, iteration_finished_callback, alpha0, alpha1, **kw)


def divergence_kernel_launch(dx, dy, dz, width, height, depth, is_host_data=False):
    """Reference cuda_image_operations/divergence.cu:59-141 (host arrays either way)."""
    del is_host_data
    return _ctx().divergence(_vol(dx, width, height, depth), _vol(dy, width, height, depth),
                             _vol(dz, width, height, depth))


def cosine_transform_kernel_launch(image_host, width, height, depth,
                                   precision=native.DCT_REFERENCE):
    """Reference cuda_image_operations/cosine_transform.cu:226-293."""
    return _ctx().cosine_transform(_vol(image_host, width, height, depth), precision)


def inverse_cosine_transform_kernel_launch(image_host, width, height, depth,
                                           precision=native.DCT_REFERENCE):
    """Reference cuda_image_operations/inverse_cosine_transform.cu:167-227."""
    return _ctx().inverse_cosine_transform(_vol(image_host, width, height, depth), precision)


def solve_poisson_in_cosine_domain_kernel_launch(image_host, width, height, depth,
                                                 index_quirk=False):
    """Reference cuda_image_operations/solve_poisson_in_cosine_domain.cu:86-130."""
    return _ctx().solve_poisson(_vol(image_host, width, height, depth), index_quirk)


class TGVDeshadeProcessor:
    """Reference tgv_deshade/TGVDeshadeProcessor.{h,cpp} (the methods on the hot path)."""

    @staticmethod
    def integrate_image_gradients_poisson_cosine_transform(gradient_x, gradient_y, gradient_z,
                                                           width, height, depth,
                                                           is_host_data=False,
                                                           precision=native.DCT_REFERENCE):
        """TGVDeshadeProcessor.cpp:386-408."""
        del is_host_data
        return _ctx().integrate_gradients_poisson(_vol(gradient_x, width, height, depth),
                                                  _vol(gradient_y, width, height, depth),
                                                  _vol(gradient_z, width, height, depth),
                                                  precision)

    @staticmethod
    def processTGV2L1GPUCuda(input_image, lambda_, alpha0, alpha1, iteration_count,
                             cuda_block_dimension=-1, mask_image=None,
                             set_negative_values_to_zero=True, **kw):
        """TGVDeshadeProcessor.cpp:122-166 (CLI overload): returns
        (denoised_image, shading_image, deshaded_image) as (D, H, W) arrays.
        The mask multiplies the deshaded image only, as in the reference."""
        del cuda_block_dimension
        f = np.ascontiguousarray(input_image, dtype=np.float32)
        flags = native.FLAG_SET_NEGATIVE_ZERO if set_negative_values_to_zero else 0
        p = _params(lambda_, iteration_count, 0, alpha0, alpha1, kw.pop("data_term", native.DATA_L1),
                    flags=flags, **kw)
        return _ctx().deshade_pipeline(f, p, mask_image)

    @staticmethod
    def processTGV2L1GPUCuda_callback(input_image, lambda_, alpha0, alpha1, iteration_count,
                                      paint_iteration_interval, iteration_finished_callback,
                                      mask=None, set_negative_values_to_zero=True,
                                      add_background_back=False, **kw):
        """TGVDeshadeProcessor.cpp:285-314 -> processTVGPUCuda :224-282 (mask also
        multiplies the shading; optional add-back of u*(1-mask)).  The callback
        receives (iteration_index, iteration_count, u, v_x, v_y, v_z)."""
        f = np.ascontiguousarray(input_image, dtype=np.float32)
        flags = native.FLAG_MASK_SHADING
        if set_negative_values_to_zero:
            flags |= native.FLAG_SET_NEGATIVE_ZERO
        if add_background_back:
            flags |= native.FLAG_ADD_BACKGROUND_BACK
        p = _params(lambda_, iteration_count, paint_iteration_interval, alpha0, alpha1,
                    kw.pop("data_term", native.DATA_L1), flags=flags, **kw)
        return _ctx().deshade_pipeline(f, p, mask, iteration_finished_callback)


    @staticmethod
    def processTGV2L1GPUCuda2D(input_image, lambda_, alpha0, alpha1, iteration_count,
                               paint_iteration_interval=0, iteration_finished_callback=None,
                               mask=None, set_negative_values_to_zero=True,
                               add_background_back=False, **kw):
        """TGVDeshadeProcessor.cpp:316-378, the widget's path for depth-1 images
        (TGVDeshadeWidget.cpp:100-127): the 2-D loop, then l from the 2-D
        divergence of v, r = u - l, mask multiplying l and r, clamp, optional
        add-back of u*(1-mask).  Returns (deshaded_image, denoised_image,
        shading_image) as (1, H, W) arrays; the callback receives
        (iteration_index, iteration_count, u, l, r) with the same composition."""
        f = np.ascontiguousarray(input_image, dtype=np.float32)
        if f.ndim == 2:
            f = f[None]
        if f.shape[0] != 1:
            raise ValueError("processTGV2L1GPUCuda2D needs a depth-1 image")
        m = None if mask is None else np.ascontiguousarray(mask, dtype=np.float32).reshape(f.shape)
        flags = native.FLAG_MASK_SHADING | native.FLAG_PLANAR
        if set_negative_values_to_zero:
            flags |= native.FLAG_SET_NEGATIVE_ZERO
        if add_background_back:
            flags |= native.FLAG_ADD_BACKGROUND_BACK
        p = _params(lambda_, iteration_count, paint_iteration_interval, alpha0, alpha1,
                    kw.pop("data_term", native.DATA_L1), flags=flags, **kw)
        cb = None
        if iteration_finished_callback is not None:
            def cb(it, count, u, vx, vy, vz):
                aux = _aux_ctx()
                l = aux.integrate_gradients_poisson(vx, vy, vz, p.dct_precision)
                r, l = aux.deshade_epilogue(u, l, m, flags)
                return iteration_finished_callback(it, count, u, l, r)
        u, l, r = _ctx().deshade_pipeline(f, p, m, cb)
        return r, u, l


class TGVProcessor:
    """Reference tgv/TGVProcessor.cpp:140-180 (TGV2 denoise entry points)."""

    @staticmethod
    def processTGV2L2GPUCuda(input_image, lambda_, alpha0, alpha1, iteration_count,
                             paint_iteration_interval=0, iteration_finished_callback=None, **kw):
        f = np.ascontiguousarray(input_image, dtype=np.float32)
        d, h, w = f.shape
        return tgv2_l2_launch(f, w, h, d, lambda_, iteration_count, paint_iteration_interval,
                              iteration_finished_callback, alpha0, alpha1, **kw)

    @staticmethod
    def processTGV2L1GPUCuda(input_image, lambda_, alpha0, alpha1, iteration_count,
                             paint_iteration_interval=0, iteration_finished_callback=None, **kw):
        f = np.ascontiguousarray(input_image, dtype=np.float32)
        d, h, w = f.shape
        return tgv2_l1_launch(f, w, h, d, lambda_, iteration_count, paint_iteration_interval,
                              iteration_finished_callback, alpha0, alpha1, **kw)


class TGVKDeshadeProcessor:
    """Reference tgvk_deshade/TGVKDeshadeProcessor.{h,cpp}."""

    @staticmethod
    def processTGVKL1Cuda(input_image, lambda_, order, alpha, iteration_count, mask=None,
                          set_negative_values_to_zero=True, add_background_back=False,
                          paint_iteration_interval=0, iteration_finished_callback=None,
                          calculate_div_v=False, **kw):
        """TGVKDeshadeProcessor.cpp:76-174 (and :177-270 at depth 1) ->
        (denoised_image, shading_image, deshaded_image, div_v_image or None).
        Order 2 runs the TGV2 loop with alpha0 = alpha[1], alpha1 = alpha[0];
        order >= 3 the TGV-k loop; depth 1 the 2-D loops.  The mask multiplies
        l and r, r is clamped, u*(1 - mask) optionally added back.  The callback
        receives (iteration_index, iteration_count, u, l, r) computed with the
        same composition."""
        if len(alpha) != order:
            raise ValueError("alpha needs one weight per order")
        f = np.ascontiguousarray(input_image, dtype=np.float32)
        m = None if mask is None else np.ascontiguousarray(mask, dtype=np.float32).reshape(f.shape)
        flags = native.FLAG_MASK_SHADING
        if set_negative_values_to_zero:
            flags |= native.FLAG_SET_NEGATIVE_ZERO
        if add_background_back:
            flags |= native.FLAG_ADD_BACKGROUND_BACK
        p = _params(lambda_, iteration_count, paint_iteration_interval, 1.0, 1.0, native.DATA_L1,
                    flags=flags, **kw)
        cb = None
        if iteration_finished_callback is not None:
            def cb(it, count, u, vx, vy, vz):
                aux = _aux_ctx()
                l = aux.integrate_gradients_poisson(vx, vy, vz, p.dct_precision)
                r, l = aux.deshade_epilogue(u, l, m, flags)
                return iteration_finished_callback(it, count, u, l, r)
        u, l, r = _ctx().tgvk_deshade_pipeline(f, p, alpha, m, cb)
        div_v = None
        if calculate_div_v:
            st = _ctx().get_state()
            div_v = _ctx().divergence(st[2], st[3], st[4])
        return u, l, r, div_v

    @staticmethod
    def updateAlpha(order, updateElement):
        """TGVKDeshadeProcessor.cpp:272-281: the widget's default weights
        1, 4, 7, ... for alpha[0 .. order-1]."""
        for i in range(order):
            updateElement(1 + 3 * i)


class TGVKDeshadeMaskedProcessor:
    """Reference tgvk_deshade_masked/TGVKDeshadeMaskedProcessor.{h,cpp}."""

    @staticmethod
    def processTGVKL1Cuda(input_image, lambda_, order, alpha, iteration_count,
                          cuda_block_dimension=-1, mask=None, set_negative_values_to_zero=True,
                          add_background_back=False, paint_iteration_interval=0,
                          iteration_finished_callback=None, calculate_div_v=False, **kw):
        """TGVKDeshadeMaskedProcessor.cpp:76-226 -> (denoised_image, shading_image,
        deshaded_image, div_v_image or None).  Order 2 is the masked TGV2
        processor (alpha0 = alpha[1], alpha1 = alpha[0]); order >= 3 the masked
        TGV-k loop on the unmasked input (2-D at depth 1), updates where
        mask >= 1; l and r times the mask, clamp, then the add-back of
        u*(1 - mask) or r times the mask again.  No mask: all ones.  The
        callback receives (iteration_index, iteration_count, u, l, r)."""
        del cuda_block_dimension
        if len(alpha) != order:
            raise ValueError("alpha needs one weight per order")
        f = np.ascontiguousarray(input_image, dtype=np.float32)
        m = (np.ones_like(f) if mask is None
             else np.ascontiguousarray(mask, dtype=np.float32).reshape(f.shape))
        flags = native.FLAG_MASKED_ITERATION | native.FLAG_MASK_SHADING
        if set_negative_values_to_zero:
            flags |= native.FLAG_SET_NEGATIVE_ZERO
        if add_background_back:
            flags |= native.FLAG_ADD_BACKGROUND_BACK
        p = _params(lambda_, iteration_count, paint_iteration_interval, 1.0, 1.0, native.DATA_L1,
                    flags=flags, **kw)
        cb = None
        if iteration_finished_callback is not None:
            def cb(it, count, u, vx, vy, vz):
                aux = _aux_ctx()
                l = aux.integrate_gradients_poisson(vx, vy, vz, p.dct_precision)
                r, l = aux.deshade_epilogue(u, l, m, flags)
                return iteration_finished_callback(it, count, u, l, r)
        u, l, r = _ctx().tgvk_deshade_pipeline(f, p, alpha, m, cb)
        div_v = None
        if calculate_div_v:
            st = _ctx().get_state()
            div_v = _ctx().divergence(st[2], st[3], st[4])
        return u, l, r, div_v


class ResizeProcessor:
    """Reference resize/ResizeProcessor.{h,cpp}: the cubic B-spline resampling
    (InterpolationMethod::BSpline3) the downsampled deshade uses."""

    @staticmethod
    def process(image, *args):
        """process(image, size_factor) or process(image, width, height, depth)
        (ResizeProcessor.cpp:41-152, BSpline3) -> the resized (D', H', W') array."""
        a = np.ascontiguousarray(image, dtype=np.float32)
        if len(args) == 1:
            out_shape = native.resize_extents(a.shape, args[0])
        elif len(args) == 3:
            w, h, d = args
            out_shape = (int(d), int(h), int(w))
        else:
            raise TypeError("process(image, size_factor) or process(image, width, height, depth)")
        return _ctx().resize_bspline3(a, out_shape)


def _downsampled(masked_variant, input_image, downsampling_factor, lambda_, order, alpha,
                 iteration_count, mask, set_negative_values_to_zero, add_background_back,
                 paint_iteration_interval, iteration_finished_callback, calculate_div_v, kw):
    if len(alpha) != order:
        raise ValueError("alpha needs one weight per order")
    f = np.ascontiguousarray(input_image, dtype=np.float32)
    m = None if mask is None else np.ascontiguousarray(mask, dtype=np.float32).reshape(f.shape)
    flags = native.FLAG_MASKED_ITERATION if masked_variant else 0
    if set_negative_values_to_zero:
        flags |= native.FLAG_SET_NEGATIVE_ZERO
    if add_background_back:
        flags |= native.FLAG_ADD_BACKGROUND_BACK
    p = _params(lambda_, iteration_count, paint_iteration_interval, 1.0, 1.0, native.DATA_L1,
                flags=flags, **kw)
    cb = None
    if iteration_finished_callback is not None:
        small = (f.shape if downsampling_factor == 1
                 else native.resize_extents(f.shape, downsampling_factor))
        ms = None
        if m is not None:
            ms = m if downsampling_factor == 1 else _aux_ctx().resize_bspline3(m, small)
        cflags = native.FLAG_MASK_SHADING | (flags & native.FLAG_SET_NEGATIVE_ZERO)
        if masked_variant:
            cflags |= native.FLAG_MASKED_ITERATION | (flags & native.FLAG_ADD_BACKGROUND_BACK)
            if ms is None:
                ms = np.ones(small, np.float32)
        elif ms is not None:
            cflags |= flags & native.FLAG_ADD_BACKGROUND_BACK

        def cb(it, count, u, vx, vy, vz):
            aux = _aux_ctx()
            l = aux.integrate_gradients_poisson(vx, vy, vz, p.dct_precision)
            r, l = aux.deshade_epilogue(u, l, ms, cflags)
            return iteration_finished_callback(it, count, u, l, r)
    u, l, r = _ctx().tgvk_deshade_downsampled(f, p, downsampling_factor, alpha, m, cb)
    div_v = None
    if calculate_div_v:
        st = _ctx().get_state()
        div_v = _ctx().divergence(st[2], st[3], st[4])
        if downsampling_factor != 1:
            div_v = _ctx().resize_bspline3(div_v, f.shape)
    return u, l, r, div_v


class TGVKDeshadeDownsampledProcessor:
    """Reference tgvk_deshade_downsampled/TGVKDeshadeDownsampledProcessor.{h,cpp}."""

    @staticmethod
    def processTGVKL1Cuda(input_image, downsampling_factor, lambda_, order, alpha,
                          iteration_count, mask=None, set_negative_values_to_zero=True,
                          add_background_back=False, paint_iteration_interval=0,
                          iteration_finished_callback=None, calculate_div_v=False, **kw):
        """TGVKDeshadeDownsampledProcessor.cpp:29-129 -> (denoised_image,
        shading_image, deshaded_image, div_v_image or None): the masked input
        and the mask resized by the factor (cubic B-splines), the masked TGV-k
        processor there, u and l resized back, r = f*mask - l with the mask,
        add-back and clamp.  The callback sees the resized volume's (u, l, r)."""
        return _downsampled(True, input_image, downsampling_factor, lambda_, order, alpha,
                            iteration_count, mask, set_negative_values_to_zero,
                            add_background_back, paint_iteration_interval,
                            iteration_finished_callback, calculate_div_v, kw)


class TGVKDeshadeDownsampledNotMaskedProcessor:
    """Reference tgvk_deshade_downsampled/TGVKDeshadeDownsampledNotMaskedProcessor.{h,cpp}."""

    @staticmethod
    def processTGVKL1Cuda(input_image, downsampling_factor, lambda_, order, alpha,
                          iteration_count, mask=None, set_negative_values_to_zero=True,
                          add_background_back=False, paint_iteration_interval=0,
                          iteration_finished_callback=None, calculate_div_v=False, **kw):
        """TGVKDeshadeDownsampledNotMaskedProcessor.cpp:29-120: as the masked
        form, with the input itself resized, the unmasked TGV-k processor and
        r = f - l (times the mask)."""
        return _downsampled(False, input_image, downsampling_factor, lambda_, order, alpha,
                            iteration_count, mask, set_negative_values_to_zero,
                            add_background_back, paint_iteration_interval,
                            iteration_finished_callback, calculate_div_v, kw)


def _aux_ctx() -> native.Context:
    """Second context for work issued from inside an iteration callback."""
    c = getattr(_tls, "aux", None)
    if c is None:
        c = native.Context(0)
        _tls.aux = c
    return c


class TGVDeshadeMaskedProcessor:
    """Reference tgv_deshade_masked/TGVDeshadeMaskedProcessor.{h,cpp}."""

    @staticmethod
    def buildMaskIndices(mask):
        """TGVDeshadeMaskedProcessor.cpp:337-385: linear indices of the voxels
        with mask >= 1, and of those on / not on each volume face.  Returns
        (pixel_indices_inside_mask, left, not_left, right, not_right, top,
        not_top, bottom, not_bottom, front, not_front, back, not_back)."""
        m = np.asarray(mask, dtype=np.float32)
        d, h, w = m.shape
        inside = np.flatnonzero(m.ravel() >= 1).astype(np.uint32)
        z, rem = np.divmod(inside, w * h)
        y, x = np.divmod(rem, w)
        out = [inside]
        for coord, last in ((x, w - 1), (y, h - 1), (z, d - 1)):
            for edge in (coord == 0, coord == last):
                out += [inside[edge], inside[~edge]]
        return tuple(out)

    @staticmethod
    def buildMaskIndices2D(mask):
        """TGVDeshadeMaskedProcessor.cpp:388-421: linear indices of the pixels
        with mask >= 1, and of those on / not on each image edge.  Returns
        (pixel_indices_inside_mask, left, not_left, right, not_right, top,
        not_top, bottom, not_bottom)."""
        m = np.asarray(mask, dtype=np.float32)
        m = m.reshape((-1,) + m.shape[-2:])
        if m.shape[0] != 1:
            raise ValueError("buildMaskIndices2D needs a depth-1 mask")
        return TGVDeshadeMaskedProcessor.buildMaskIndices(m)[:9]

    @staticmethod
    def processTGV2L1GPUCuda(input_image, lambda_, alpha0, alpha1, iteration_count,
                             cuda_block_dimension=-1, paint_iteration_interval=0,
                             iteration_finished_callback=None, mask=None,
                             set_negative_values_to_zero=True, add_background_back=False,
                             calculate_div_v=False, **kw):
        """TGVDeshadeMaskedProcessor.cpp:76-210 -> (deshaded_image, denoised_image,
        shading_image, div_v_image or None).  f is the unmasked input; the
        iteration updates only voxels with mask >= 1; l and r are multiplied by
        the mask, r clamped, then either u*(1 - mask) is added back or r is
        multiplied by the mask again.  Without a mask the mask is all ones.
        The callback receives (iteration_index, iteration_count, u, l, r)
        computed with the same composition, as in the reference.  A depth-1
        image takes the 2-D loop (processTGV2L1GPUCuda2D, :93-98 and :215-335;
        div_v from the 2-D divergence)."""
        del cuda_block_dimension
        f = np.ascontiguousarray(input_image, dtype=np.float32)
        m = (np.ones_like(f) if mask is None
             else np.ascontiguousarray(mask, dtype=np.float32).reshape(f.shape))
        flags = native.FLAG_MASKED_ITERATION | native.FLAG_MASK_SHADING
        if f.shape[0] == 1:
            flags |= native.FLAG_PLANAR
        if set_negative_values_to_zero:
            flags |= native.FLAG_SET_NEGATIVE_ZERO
        if add_background_back:
            flags |= native.FLAG_ADD_BACKGROUND_BACK
        p = _params(lambda_, iteration_count, paint_iteration_interval, alpha0, alpha1,
                    native.DATA_L1, flags=flags, **kw)
        cb = None
        if iteration_finished_callback is not None:
            def cb(it, count, u, vx, vy, vz):
                aux = _aux_ctx()
                l = aux.integrate_gradients_poisson(vx, vy, vz, p.dct_precision)
                r, l = aux.deshade_epilogue(u, l, m, flags)
                return iteration_finished_callback(it, count, u, l, r)
        u, l, r = _ctx().deshade_pipeline(f, p, m, cb)
        div_v = None
        if calculate_div_v:
            st = _ctx().get_state()
            div_v = _ctx().divergence(st[2], st[3], st[4])
        return r, u, l, div_v
