// tgv2_deshade_application: the reference's deshade CLI contract on the MI355X path.
//
//   tgv2_deshade_application input_image_path lambda alpha0 alpha1 iteration_count
//                            cuda_block_dimension [mask] output_denoised_path
//                            output_shading_path output_deshaded_path
//
// Same positional arguments, usage text, exit codes and output trio as the
// reference CLI (terminal_programs/tgv2_deshade_main.cpp:33-181): TGV2-L1
// deshade with tau = sigma = 1/sqrt(20), theta = 1, then the Poisson
// integration of v and r = u - l with negative values set to zero
// (TGVDeshadeProcessor.cpp:122-166).  Images are MetaImage files (.mha/.mhd)
// instead of any ITK-readable format; outputs keep the input's spacing and
// origin.  With a mask, as the reference CLI (:69-84), the masked deshade
// TGVDeshadeMaskedProcessor::processTGV2L1GPUCuda with negative values set to
// zero and the background added back: the iteration runs on the unmasked
// input and updates only voxels with mask >= 1, l and r are multiplied by
// the mask, r clamped, and u*(1 - mask) added to r
// (TGVDeshadeMaskedProcessor.cpp:76-210).
#include "ambient_amd.h"
#include "metaimage.hpp"

#include <cstdio>
#include <iostream>
#include <string>

static void print_usage() {
    std::cout << "parameter: input_image_path lambda alpha0 alpha1 iteration_count cuda_block_dimension [mask] "
              << "output_denoised_path output_shading_path output_deshaded_path";
}

int main(int argc, char* argv[]) {
    if (argc < 9) {
        print_usage();
        return 1;
    }
    const std::string input_path = argv[1];
    const float lambda = std::stof(argv[2]);
    const float alpha0 = std::stof(argv[3]);
    const float alpha1 = std::stof(argv[4]);
    const int iteration_count = std::stoi(argv[5]);
    (void)std::stoi(argv[6]);   // cuda_block_dimension: accepted, launch shapes are per kernel
    const bool mask_given = argc == 11;
    const std::string mask_path = mask_given ? argv[7] : "";
    const std::string out_u = argv[mask_given ? 8 : 7];
    const std::string out_l = argv[mask_given ? 9 : 8];
    const std::string out_r = argv[mask_given ? 10 : 9];

    metaimage::Volume in;
    try {
        in = metaimage::read(input_path);
    } catch (const std::exception& e) {
        std::cerr << "input image file does not exist: " << input_path << " (" << e.what() << ")"
                  << std::endl;
        return 2;
    }
    metaimage::Volume mask;
    if (mask_given) {
        try {
            mask = metaimage::read(mask_path);
        } catch (const std::exception& e) {
            std::cerr << "mask image file does not exist: " << mask_path << std::endl;
            return 3;
        }
        if (mask.w != in.w || mask.h != in.h || mask.d != in.d) {
            std::cerr << "mask image size differs from the input image size" << std::endl;
            return 4;
        }
    }
    if (iteration_count < 0) {
        std::cerr << "iteration_count must be >= 0" << std::endl;
        return 1;
    }

    amb_context* ctx = nullptr;
    if (amb_context_create(0, &ctx) != AMB_OK) {
        std::printf("ambient_amd error: %s\n", amb_last_error());
        return 5;
    }
    amb_tgv_params p;
    amb_tgv_params_default(&p);   // tau = sigma = 1/sqrt(20), theta = 1, L1, FP32
    p.lambda = lambda;
    p.alpha0 = alpha0;
    p.alpha1 = alpha1;
    p.iterations = uint32_t(iteration_count);
    p.flags = AMB_FLAG_SET_NEGATIVE_ZERO;
    if (mask_given)
        p.flags |= AMB_FLAG_MASKED_ITERATION | AMB_FLAG_MASK_SHADING | AMB_FLAG_ADD_BACKGROUND_BACK;
    metaimage::Volume u = in, l = in, r = in;
    const amb_status st = amb_tgv2_deshade_pipeline(
        ctx, AMB_MEM_HOST, in.data.data(), in.w, in.h, in.d, &p,
        mask_given ? mask.data.data() : nullptr, u.data.data(), l.data.data(), r.data.data(),
        nullptr, nullptr);
    if (st != AMB_OK) {
        std::printf("ambient_amd error: %s\n", amb_last_error());
        amb_context_destroy(ctx);
        return 5;
    }
    amb_context_destroy(ctx);
    try {
        metaimage::write(out_u, u);
        metaimage::write(out_l, l);
        metaimage::write(out_r, r);
    } catch (const std::exception& e) {
        std::cerr << e.what() << std::endl;
        return 6;
    }
    return 0;
}
