// ambient_amd_reference_compat.hpp - the reference's C++ launcher templates
// (josefkoller/ambient_image_processor), re-exposed with their exact
// signatures on top of the C ABI in ambient_amd.h, so that processor code
// written against them (TGVDeshadeProcessor.cpp, TGVProcessor.cpp,
// CudaImageOperationsProcessor.cpp) links unchanged against libambient_amd.so.
//
// Semantics kept from the reference:
//   - host arrays in, `new Pixel[voxel_count]` out (the caller delete[]s);
//   - tau = sigma = 1/sqrt(20), theta = 1 (tgv_deshade/tgv2_l1_deshade.cu:148-150);
//   - a failing call prints a message and exit()s with the status, as
//     cudaCheckError does (cuda/cuda_helper.in:11-22);
//   - the iteration callback gets host copies of u (and v) every
//     paint_iteration_interval iterations and stops the loop by returning true.
// Only the float instantiation exists (ITKImage::PixelType is float,
// itk/ITKImage.h:32); cuda_block_dimension is accepted and ignored.
// One amb_context per calling thread, on device 0 (the reference uses device 0).
#ifndef AMBIENT_AMD_REFERENCE_COMPAT_HPP
#define AMBIENT_AMD_REFERENCE_COMPAT_HPP

#include "ambient_amd.h"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <functional>
#include <type_traits>
#include <vector>

namespace ambient_amd_compat {

typedef unsigned int uint;

inline void check(amb_status s, const char* what) {
    if (s != AMB_OK) {
        std::printf("ambient_amd error: %s (%s)\n", amb_last_error(), what);
        std::exit(int(s));
    }
}

inline amb_context* context() {
    struct Holder {
        amb_context* ctx = nullptr;
        ~Holder() { if (ctx) amb_context_destroy(ctx); }
    };
    thread_local Holder h;
    if (!h.ctx) check(amb_context_create(0, &h.ctx), "amb_context_create");
    return h.ctx;
}

inline amb_tgv_params reference_params(float lambda, uint iterations, uint interval,
                                       float alpha0, float alpha1, uint32_t data_term) {
    amb_tgv_params p;
    amb_tgv_params_default(&p);
    p.lambda = lambda;
    p.alpha0 = alpha0;
    p.alpha1 = alpha1;
    p.iterations = iterations;
    p.callback_interval = interval;
    p.data_term = data_term;
    return p;
}

template <class Fn>
struct Trampoline {
    static int deshade(uint32_t it, uint32_t count, const float* u, const float* vx,
                       const float* vy, const float* vz, void* user) {
        Fn& fn = *static_cast<Fn*>(user);
        return fn(it, count, const_cast<float*>(u), const_cast<float*>(vx),
                  const_cast<float*>(vy), const_cast<float*>(vz)) ? 1 : 0;
    }
    static int deshade_2d(uint32_t it, uint32_t count, const float* u, const float* vx,
                          const float* vy, const float*, void* user) {
        Fn& fn = *static_cast<Fn*>(user);
        return fn(it, count, const_cast<float*>(u), const_cast<float*>(vx),
                  const_cast<float*>(vy)) ? 1 : 0;
    }
    static int denoise(uint32_t it, uint32_t count, const float* u, const float*, const float*,
                       const float*, void* user) {
        Fn& fn = *static_cast<Fn*>(user);
        return fn(it, count, const_cast<float*>(u)) ? 1 : 0;
    }
};

template <typename Pixel>
Pixel* image_op(Pixel* image_host, uint width, uint height, uint depth,
                amb_status (*op)(amb_context*, amb_memory, const float*, uint32_t, uint32_t,
                                 uint32_t, uint32_t, float*),
                uint32_t arg, const char* what) {
    static_assert(std::is_same<Pixel, float>::value, "only the float instantiation exists");
    Pixel* out = new Pixel[size_t(width) * height * depth];
    check(op(context(), AMB_MEM_HOST, image_host, width, height, depth, arg, out), what);
    return out;
}

}  // namespace ambient_amd_compat

// ---- reference tgv_deshade/tgv2_l1_deshade.cu:33-41 ------------------------
template <typename Pixel>
Pixel* tgv2_l1_deshade_launch(Pixel* f_host, ambient_amd_compat::uint width,
                              ambient_amd_compat::uint height, ambient_amd_compat::uint depth,
                              Pixel lambda, ambient_amd_compat::uint iteration_count,
                              ambient_amd_compat::uint paint_iteration_interval,
                              const int cuda_block_dimension,
                              std::function<bool(ambient_amd_compat::uint, ambient_amd_compat::uint,
                                                 Pixel*, Pixel*, Pixel*, Pixel*)>
                                  iteration_finished_callback,
                              Pixel alpha0, Pixel alpha1, Pixel** v_x_host, Pixel** v_y_host,
                              Pixel** v_z_host) {
    using namespace ambient_amd_compat;
    static_assert(std::is_same<Pixel, float>::value, "only the float instantiation exists");
    (void)cuda_block_dimension;
    typedef std::function<bool(uint, uint, Pixel*, Pixel*, Pixel*, Pixel*)> Fn;
    const size_t n = size_t(width) * height * depth;
    amb_tgv_params p = reference_params(lambda, iteration_count, paint_iteration_interval,
                                        alpha0, alpha1, AMB_DATA_L1);
    Pixel* u = new Pixel[n];
    *v_x_host = new Pixel[n];
    *v_y_host = new Pixel[n];
    *v_z_host = new Pixel[n];
    amb_iteration_callback cb = iteration_finished_callback ? &Trampoline<Fn>::deshade : nullptr;
    check(amb_tgv2_deshade(context(), AMB_MEM_HOST, f_host, width, height, depth, &p, u,
                           *v_x_host, *v_y_host, *v_z_host, cb, &iteration_finished_callback),
          "tgv2_l1_deshade_launch");
    return u;
}

// ---- reference tgv_deshade_masked/tgv2_l1_deshade_masked.cu:27-44 ---------
// IndexVector is const std::vector<uint>& (tgv_masked_common.cu:41).  The
// iteration updates only the voxels in masked_pixel_indices; the twelve edge
// lists are built by the reference from the volume faces
// (TGVDeshadeMaskedProcessor.cpp:355-383), which the sweeps know themselves,
// so they are accepted and not read.
template <typename Pixel>
Pixel* tgv2_l1_deshade_masked_launch(
    Pixel* f_host, ambient_amd_compat::uint width, ambient_amd_compat::uint height,
    ambient_amd_compat::uint depth, Pixel lambda, ambient_amd_compat::uint iteration_count,
    ambient_amd_compat::uint paint_iteration_interval, const int cuda_block_dimension,
    std::function<bool(ambient_amd_compat::uint, ambient_amd_compat::uint, Pixel*, Pixel*, Pixel*,
                       Pixel*)>
        iteration_finished_callback,
    Pixel alpha0, Pixel alpha1, Pixel** v_x_host, Pixel** v_y_host, Pixel** v_z_host,
    const std::vector<ambient_amd_compat::uint>& masked_pixel_indices,
    const std::vector<ambient_amd_compat::uint>& /*left_edge_pixel_indices*/,
    const std::vector<ambient_amd_compat::uint>& /*not_left_edge_pixel_indices*/,
    const std::vector<ambient_amd_compat::uint>& /*right_edge_pixel_indices*/,
    const std::vector<ambient_amd_compat::uint>& /*not_right_edge_pixel_indices*/,
    const std::vector<ambient_amd_compat::uint>& /*top_edge_pixel_indices*/,
    const std::vector<ambient_amd_compat::uint>& /*not_top_edge_pixel_indices*/,
    const std::vector<ambient_amd_compat::uint>& /*bottom_edge_pixel_indices*/,
    const std::vector<ambient_amd_compat::uint>& /*not_bottom_edge_pixel_indices*/,
    const std::vector<ambient_amd_compat::uint>& /*front_edge_pixel_indices*/,
    const std::vector<ambient_amd_compat::uint>& /*not_front_edge_pixel_indices*/,
    const std::vector<ambient_amd_compat::uint>& /*back_edge_pixel_indices*/,
    const std::vector<ambient_amd_compat::uint>& /*not_back_edge_pixel_indices*/) {
    using namespace ambient_amd_compat;
    static_assert(std::is_same<Pixel, float>::value, "only the float instantiation exists");
    (void)cuda_block_dimension;
    typedef std::function<bool(uint, uint, Pixel*, Pixel*, Pixel*, Pixel*)> Fn;
    const size_t n = size_t(width) * height * depth;
    std::vector<uint8_t> active(n, 0);
    for (uint i : masked_pixel_indices) {
        if (size_t(i) >= n) {
            std::printf("ambient_amd error: masked pixel index %u outside the volume\n", i);
            std::exit(int(AMB_ERR_INVALID_ARGUMENT));
        }
        active[i] = 1;
    }
    amb_tgv_params p = reference_params(lambda, iteration_count, paint_iteration_interval,
                                        alpha0, alpha1, AMB_DATA_L1);
    Pixel* u = new Pixel[n];
    *v_x_host = new Pixel[n];
    *v_y_host = new Pixel[n];
    *v_z_host = new Pixel[n];
    amb_iteration_callback cb = iteration_finished_callback ? &Trampoline<Fn>::deshade : nullptr;
    check(amb_tgv2_deshade_masked(context(), AMB_MEM_HOST, f_host, width, height, depth, &p,
                                  active.data(), u, *v_x_host, *v_y_host, *v_z_host, cb,
                                  &iteration_finished_callback),
          "tgv2_l1_deshade_masked_launch");
    return u;
}

// ---- reference tgv_deshade/tgv2_l1_deshade_2d.cu:27-36 (declared
// TGVDeshadeProcessor.cpp:41-50) -------------------------------------------
// The 2-D loop: AMB_FLAG_PLANAR at depth 1; v_z is computed as 0 and dropped.
template <typename Pixel>
Pixel* tgv2_l1_deshade_launch_2d(Pixel* f_host, ambient_amd_compat::uint width,
                                 ambient_amd_compat::uint height, Pixel lambda,
                                 ambient_amd_compat::uint iteration_count,
                                 ambient_amd_compat::uint paint_iteration_interval,
                                 std::function<bool(ambient_amd_compat::uint,
                                                    ambient_amd_compat::uint, Pixel*, Pixel*,
                                                    Pixel*)>
                                     iteration_finished_callback,
                                 Pixel alpha0, Pixel alpha1, Pixel** v_x_host, Pixel** v_y_host) {
    using namespace ambient_amd_compat;
    static_assert(std::is_same<Pixel, float>::value, "only the float instantiation exists");
    typedef std::function<bool(uint, uint, Pixel*, Pixel*, Pixel*)> Fn;
    const size_t n = size_t(width) * height;
    amb_tgv_params p = reference_params(lambda, iteration_count, paint_iteration_interval,
                                        alpha0, alpha1, AMB_DATA_L1);
    p.flags |= AMB_FLAG_PLANAR;
    Pixel* u = new Pixel[n];
    *v_x_host = new Pixel[n];
    *v_y_host = new Pixel[n];
    amb_iteration_callback cb = iteration_finished_callback ? &Trampoline<Fn>::deshade_2d : nullptr;
    check(amb_tgv2_deshade(context(), AMB_MEM_HOST, f_host, width, height, 1, &p, u, *v_x_host,
                           *v_y_host, nullptr, cb, &iteration_finished_callback),
          "tgv2_l1_deshade_launch_2d");
    return u;
}

// ---- reference tgv_deshade_masked/tgv2_l1_deshade_masked_2d.cu (declared
// TGVDeshadeMaskedProcessor.cpp:53-68) ---------------------------------------
// The masked 2-D loop; the eight edge lists are accepted and not read.
template <typename Pixel>
Pixel* tgv2_l1_deshade_masked_2d_launch(
    Pixel* f_host, ambient_amd_compat::uint width, ambient_amd_compat::uint height, Pixel lambda,
    ambient_amd_compat::uint iteration_count, ambient_amd_compat::uint paint_iteration_interval,
    const int cuda_block_dimension,
    std::function<bool(ambient_amd_compat::uint, ambient_amd_compat::uint, Pixel*, Pixel*, Pixel*)>
        iteration_finished_callback,
    Pixel alpha0, Pixel alpha1, Pixel** v_x_host, Pixel** v_y_host,
    const std::vector<ambient_amd_compat::uint>& masked_pixel_indices,
    const std::vector<ambient_amd_compat::uint>& /*left_edge_pixel_indices*/,
    const std::vector<ambient_amd_compat::uint>& /*not_left_edge_pixel_indices*/,
    const std::vector<ambient_amd_compat::uint>& /*right_edge_pixel_indices*/,
    const std::vector<ambient_amd_compat::uint>& /*not_right_edge_pixel_indices*/,
    const std::vector<ambient_amd_compat::uint>& /*top_edge_pixel_indices*/,
    const std::vector<ambient_amd_compat::uint>& /*not_top_edge_pixel_indices*/,
    const std::vector<ambient_amd_compat::uint>& /*bottom_edge_pixel_indices*/,
    const std::vector<ambient_amd_compat::uint>& /*not_bottom_edge_pixel_indices*/) {
    using namespace ambient_amd_compat;
    static_assert(std::is_same<Pixel, float>::value, "only the float instantiation exists");
    (void)cuda_block_dimension;
    typedef std::function<bool(uint, uint, Pixel*, Pixel*, Pixel*)> Fn;
    const size_t n = size_t(width) * height;
    std::vector<uint8_t> active(n, 0);
    for (uint i : masked_pixel_indices) {
        if (size_t(i) >= n) {
            std::printf("ambient_amd error: masked pixel index %u outside the image\n", i);
            std::exit(int(AMB_ERR_INVALID_ARGUMENT));
        }
        active[i] = 1;
    }
    amb_tgv_params p = reference_params(lambda, iteration_count, paint_iteration_interval,
                                        alpha0, alpha1, AMB_DATA_L1);
    p.flags |= AMB_FLAG_PLANAR;
    Pixel* u = new Pixel[n];
    *v_x_host = new Pixel[n];
    *v_y_host = new Pixel[n];
    amb_iteration_callback cb = iteration_finished_callback ? &Trampoline<Fn>::deshade_2d : nullptr;
    check(amb_tgv2_deshade_masked(context(), AMB_MEM_HOST, f_host, width, height, 1, &p,
                                  active.data(), u, *v_x_host, *v_y_host, nullptr, cb,
                                  &iteration_finished_callback),
          "tgv2_l1_deshade_masked_2d_launch");
    return u;
}

// ---- reference tgvk_deshade/tgvk_l1_deshade.cu:35-45 (declared
// TGVKDeshadeProcessor.cpp:27-36) -------------------------------------------
// order >= 3; alpha holds `order` weights.
template <typename Pixel>
Pixel* tgvk_l1_deshade_launch(Pixel* f_host, ambient_amd_compat::uint width,
                              ambient_amd_compat::uint height, ambient_amd_compat::uint depth,
                              Pixel lambda, ambient_amd_compat::uint iteration_count,
                              ambient_amd_compat::uint paint_iteration_interval,
                              std::function<bool(ambient_amd_compat::uint, ambient_amd_compat::uint,
                                                 Pixel*, Pixel*, Pixel*, Pixel*)>
                                  iteration_finished_callback,
                              const ambient_amd_compat::uint order, const Pixel* alpha,
                              Pixel** v_x_host, Pixel** v_y_host, Pixel** v_z_host) {
    using namespace ambient_amd_compat;
    static_assert(std::is_same<Pixel, float>::value, "only the float instantiation exists");
    typedef std::function<bool(uint, uint, Pixel*, Pixel*, Pixel*, Pixel*)> Fn;
    const size_t n = size_t(width) * height * depth;
    amb_tgv_params p = reference_params(lambda, iteration_count, paint_iteration_interval,
                                        alpha[1], alpha[0], AMB_DATA_L1);
    Pixel* u = new Pixel[n];
    *v_x_host = new Pixel[n];
    *v_y_host = new Pixel[n];
    *v_z_host = new Pixel[n];
    amb_iteration_callback cb = iteration_finished_callback ? &Trampoline<Fn>::deshade : nullptr;
    check(amb_tgvk_deshade(context(), AMB_MEM_HOST, f_host, width, height, depth, &p, order,
                           alpha, u, *v_x_host, *v_y_host, *v_z_host, cb,
                           &iteration_finished_callback),
          "tgvk_l1_deshade_launch");
    return u;
}

// ---- reference tgvk_deshade/tgvk_l1_deshade_2d.cu (declared
// TGVKDeshadeProcessor.cpp:49-58) -------------------------------------------
template <typename Pixel>
Pixel* tgvk_l1_deshade_launch_2d(Pixel* f_host, ambient_amd_compat::uint width,
                                 ambient_amd_compat::uint height, Pixel lambda,
                                 ambient_amd_compat::uint iteration_count,
                                 ambient_amd_compat::uint paint_iteration_interval,
                                 std::function<bool(ambient_amd_compat::uint,
                                                    ambient_amd_compat::uint, Pixel*, Pixel*,
                                                    Pixel*)>
                                     iteration_finished_callback,
                                 const ambient_amd_compat::uint order, const Pixel* alpha,
                                 Pixel** v_x_host, Pixel** v_y_host) {
    using namespace ambient_amd_compat;
    static_assert(std::is_same<Pixel, float>::value, "only the float instantiation exists");
    typedef std::function<bool(uint, uint, Pixel*, Pixel*, Pixel*)> Fn;
    const size_t n = size_t(width) * height;
    amb_tgv_params p = reference_params(lambda, iteration_count, paint_iteration_interval,
                                        alpha[1], alpha[0], AMB_DATA_L1);
    p.flags |= AMB_FLAG_PLANAR;
    Pixel* u = new Pixel[n];
    *v_x_host = new Pixel[n];
    *v_y_host = new Pixel[n];
    amb_iteration_callback cb = iteration_finished_callback ? &Trampoline<Fn>::deshade_2d : nullptr;
    check(amb_tgvk_deshade(context(), AMB_MEM_HOST, f_host, width, height, 1, &p, order, alpha,
                           u, *v_x_host, *v_y_host, nullptr, cb, &iteration_finished_callback),
          "tgvk_l1_deshade_launch_2d");
    return u;
}

namespace ambient_amd_compat {
// active set from the reference's index list of in-mask voxels
inline std::vector<uint8_t> active_from_indices(const std::vector<uint>& indices, size_t n) {
    std::vector<uint8_t> active(n, 0);
    for (uint i : indices) {
        if (size_t(i) >= n) {
            std::printf("ambient_amd error: masked pixel index %u outside the volume\n", i);
            std::exit(int(AMB_ERR_INVALID_ARGUMENT));
        }
        active[i] = 1;
    }
    return active;
}
}  // namespace ambient_amd_compat

// ---- reference tgvk_deshade_masked/tgvk_l1_deshade_masked.cu:29-… (declared
// TGVKDeshadeMaskedProcessor.cpp:30-50) --------------------------------------
// The twelve edge lists are accepted and not read (the sweeps know the faces).
template <typename Pixel>
Pixel* tgvk_l1_deshade_masked_launch(
    Pixel* f_host, const ambient_amd_compat::uint width, const ambient_amd_compat::uint height,
    const ambient_amd_compat::uint depth, const Pixel lambda,
    const ambient_amd_compat::uint iteration_count,
    const ambient_amd_compat::uint paint_iteration_interval, const int cuda_block_dimension,
    std::function<bool(ambient_amd_compat::uint, ambient_amd_compat::uint, Pixel*, Pixel*, Pixel*,
                       Pixel*)>
        iteration_finished_callback,
    const ambient_amd_compat::uint order, const Pixel* alpha, Pixel** v_x_host, Pixel** v_y_host,
    Pixel** v_z_host, const std::vector<ambient_amd_compat::uint>& masked_pixel_indices,
    const std::vector<ambient_amd_compat::uint>& /*left*/,
    const std::vector<ambient_amd_compat::uint>& /*not_left*/,
    const std::vector<ambient_amd_compat::uint>& /*right*/,
    const std::vector<ambient_amd_compat::uint>& /*not_right*/,
    const std::vector<ambient_amd_compat::uint>& /*top*/,
    const std::vector<ambient_amd_compat::uint>& /*not_top*/,
    const std::vector<ambient_amd_compat::uint>& /*bottom*/,
    const std::vector<ambient_amd_compat::uint>& /*not_bottom*/,
    const std::vector<ambient_amd_compat::uint>& /*front*/,
    const std::vector<ambient_amd_compat::uint>& /*not_front*/,
    const std::vector<ambient_amd_compat::uint>& /*back*/,
    const std::vector<ambient_amd_compat::uint>& /*not_back*/) {
    using namespace ambient_amd_compat;
    static_assert(std::is_same<Pixel, float>::value, "only the float instantiation exists");
    (void)cuda_block_dimension;
    typedef std::function<bool(uint, uint, Pixel*, Pixel*, Pixel*, Pixel*)> Fn;
    const size_t n = size_t(width) * height * depth;
    const std::vector<uint8_t> active = active_from_indices(masked_pixel_indices, n);
    amb_tgv_params p = reference_params(lambda, iteration_count, paint_iteration_interval,
                                        alpha[1], alpha[0], AMB_DATA_L1);
    Pixel* u = new Pixel[n];
    *v_x_host = new Pixel[n];
    *v_y_host = new Pixel[n];
    *v_z_host = new Pixel[n];
    amb_iteration_callback cb = iteration_finished_callback ? &Trampoline<Fn>::deshade : nullptr;
    check(amb_tgvk_deshade_masked(context(), AMB_MEM_HOST, f_host, width, height, depth, &p,
                                  order, alpha, active.data(), u, *v_x_host, *v_y_host, *v_z_host,
                                  cb, &iteration_finished_callback),
          "tgvk_l1_deshade_masked_launch");
    return u;
}

// ---- reference tgvk_deshade_masked/tgvk_l1_deshade_masked_2d.cu (declared
// TGVKDeshadeMaskedProcessor.cpp:53-68) --------------------------------------
template <typename Pixel>
Pixel* tgvk_l1_deshade_masked_2d_launch(
    Pixel* f_host, const ambient_amd_compat::uint width, const ambient_amd_compat::uint height,
    const Pixel lambda, const ambient_amd_compat::uint iteration_count,
    const ambient_amd_compat::uint paint_iteration_interval, const int cuda_block_dimension,
    std::function<bool(ambient_amd_compat::uint, ambient_amd_compat::uint, Pixel*, Pixel*, Pixel*)>
        iteration_finished_callback,
    const ambient_amd_compat::uint order, const Pixel* alpha, Pixel** v_x_host, Pixel** v_y_host,
    const std::vector<ambient_amd_compat::uint>& masked_pixel_indices,
    const std::vector<ambient_amd_compat::uint>& /*left*/,
    const std::vector<ambient_amd_compat::uint>& /*not_left*/,
    const std::vector<ambient_amd_compat::uint>& /*right*/,
    const std::vector<ambient_amd_compat::uint>& /*not_right*/,
    const std::vector<ambient_amd_compat::uint>& /*top*/,
    const std::vector<ambient_amd_compat::uint>& /*not_top*/,
    const std::vector<ambient_amd_compat::uint>& /*bottom*/,
    const std::vector<ambient_amd_compat::uint>& /*not_bottom*/) {
    using namespace ambient_amd_compat;
    static_assert(std::is_same<Pixel, float>::value, "only the float instantiation exists");
    (void)cuda_block_dimension;
    typedef std::function<bool(uint, uint, Pixel*, Pixel*, Pixel*)> Fn;
    const size_t n = size_t(width) * height;
    const std::vector<uint8_t> active = active_from_indices(masked_pixel_indices, n);
    amb_tgv_params p = reference_params(lambda, iteration_count, paint_iteration_interval,
                                        alpha[1], alpha[0], AMB_DATA_L1);
    p.flags |= AMB_FLAG_PLANAR;
    Pixel* u = new Pixel[n];
    *v_x_host = new Pixel[n];
    *v_y_host = new Pixel[n];
    amb_iteration_callback cb = iteration_finished_callback ? &Trampoline<Fn>::deshade_2d : nullptr;
    check(amb_tgvk_deshade_masked(context(), AMB_MEM_HOST, f_host, width, height, 1, &p, order,
                                  alpha, active.data(), u, *v_x_host, *v_y_host, nullptr, cb,
                                  &iteration_finished_callback),
          "tgvk_l1_deshade_masked_2d_launch");
    return u;
}

// ---- reference tgv_cuda/tgv2_l1.cu:64-71 and tgv_cuda/tgv2_l2.cu:60-67 -------
template <typename Pixel>
Pixel* tgv2_denoise_launch_impl(uint32_t data_term, Pixel* f_host, ambient_amd_compat::uint width,
                                ambient_amd_compat::uint height, ambient_amd_compat::uint depth,
                                Pixel lambda, ambient_amd_compat::uint iteration_count,
                                ambient_amd_compat::uint paint_iteration_interval,
                                std::function<bool(ambient_amd_compat::uint,
                                                   ambient_amd_compat::uint, Pixel*)>
                                    iteration_finished_callback,
                                Pixel alpha0, Pixel alpha1) {
    using namespace ambient_amd_compat;
    static_assert(std::is_same<Pixel, float>::value, "only the float instantiation exists");
    typedef std::function<bool(uint, uint, Pixel*)> Fn;
    amb_tgv_params p = reference_params(lambda, iteration_count, paint_iteration_interval,
                                        alpha0, alpha1, data_term);
    Pixel* u = new Pixel[size_t(width) * height * depth];
    amb_iteration_callback cb = iteration_finished_callback ? &Trampoline<Fn>::denoise : nullptr;
    check(amb_tgv2_denoise(context(), AMB_MEM_HOST, f_host, width, height, depth, &p, u, cb,
                           &iteration_finished_callback),
          data_term == AMB_DATA_L1 ? "tgv2_l1_launch" : "tgv2_l2_launch");
    return u;
}

template <typename Pixel>
Pixel* tgv2_l1_launch(Pixel* f_host, ambient_amd_compat::uint width, ambient_amd_compat::uint height,
                      ambient_amd_compat::uint depth, Pixel lambda,
                      ambient_amd_compat::uint iteration_count,
                      ambient_amd_compat::uint paint_iteration_interval,
                      std::function<bool(ambient_amd_compat::uint, ambient_amd_compat::uint, Pixel*)>
                          iteration_finished_callback,
                      Pixel alpha0, Pixel alpha1) {
    return tgv2_denoise_launch_impl<Pixel>(AMB_DATA_L1, f_host, width, height, depth, lambda,
                                           iteration_count, paint_iteration_interval,
                                           iteration_finished_callback, alpha0, alpha1);
}

template <typename Pixel>
Pixel* tgv2_l2_launch(Pixel* f_host, ambient_amd_compat::uint width, ambient_amd_compat::uint height,
                      ambient_amd_compat::uint depth, Pixel lambda,
                      ambient_amd_compat::uint iteration_count,
                      ambient_amd_compat::uint paint_iteration_interval,
                      std::function<bool(ambient_amd_compat::uint, ambient_amd_compat::uint, Pixel*)>
                          iteration_finished_callback,
                      Pixel alpha0, Pixel alpha1) {
    return tgv2_denoise_launch_impl<Pixel>(AMB_DATA_L2, f_host, width, height, depth, lambda,
                                           iteration_count, paint_iteration_interval,
                                           iteration_finished_callback, alpha0, alpha1);
}

// ---- reference cuda_image_operations/*.cu launchers --------------------------
template <typename Pixel>
Pixel* divergence_kernel_launch(Pixel* dx, Pixel* dy, Pixel* dz, const ambient_amd_compat::uint width,
                                const ambient_amd_compat::uint height,
                                const ambient_amd_compat::uint depth, bool is_host_data = false) {
    using namespace ambient_amd_compat;
    static_assert(std::is_same<Pixel, float>::value, "only the float instantiation exists");
    (void)is_host_data;   // the reference's device-pointer mode is an internal detail of its launcher
    Pixel* out = new Pixel[size_t(width) * height * depth];
    check(amb_divergence(context(), AMB_MEM_HOST, dx, dy, dz, width, height, depth, out),
          "divergence_kernel_launch");
    return out;
}

// ---- reference cuda_image_operations/divergence_2d.cu:59-117 ---------------
// B_x dx + B_y dy: the 3-D divergence at depth 1 with dz = 0 (B_z of zeros adds 0).
template <typename Pixel>
Pixel* divergence_2d_kernel_launch(Pixel* dx, Pixel* dy, const ambient_amd_compat::uint width,
                                   const ambient_amd_compat::uint height, bool is_host_data = false) {
    using namespace ambient_amd_compat;
    static_assert(std::is_same<Pixel, float>::value, "only the float instantiation exists");
    (void)is_host_data;
    const size_t n = size_t(width) * height;
    std::vector<float> zeros(n, 0.0f);
    Pixel* out = new Pixel[n];
    check(amb_divergence(context(), AMB_MEM_HOST, dx, dy, zeros.data(), width, height, 1, out),
          "divergence_2d_kernel_launch");
    return out;
}

template <typename Pixel>
Pixel* cosine_transform_kernel_launch(Pixel* image_host, ambient_amd_compat::uint width,
                                      ambient_amd_compat::uint height, ambient_amd_compat::uint depth) {
    return ambient_amd_compat::image_op<Pixel>(image_host, width, height, depth,
                                               &amb_cosine_transform, AMB_DCT_REFERENCE,
                                               "cosine_transform_kernel_launch");
}

template <typename Pixel>
Pixel* inverse_cosine_transform_kernel_launch(Pixel* image_host, ambient_amd_compat::uint width,
                                              ambient_amd_compat::uint height,
                                              ambient_amd_compat::uint depth) {
    return ambient_amd_compat::image_op<Pixel>(image_host, width, height, depth,
                                               &amb_inverse_cosine_transform, AMB_DCT_REFERENCE,
                                               "inverse_cosine_transform_kernel_launch");
}

template <typename Pixel>
Pixel* solve_poisson_in_cosine_domain_kernel_launch(Pixel* image_host, ambient_amd_compat::uint width,
                                                    ambient_amd_compat::uint height,
                                                    ambient_amd_compat::uint depth) {
    return ambient_amd_compat::image_op<Pixel>(image_host, width, height, depth,
                                               &amb_solve_poisson_in_cosine_domain, 0u,
                                               "solve_poisson_in_cosine_domain_kernel_launch");
}

#endif  // AMBIENT_AMD_REFERENCE_COMPAT_HPP
