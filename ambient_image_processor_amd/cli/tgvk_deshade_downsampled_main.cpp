// tgvk_deshade_downsampled_application: the reference's downsampled TGV-k
// deshade CLI contract on the MI355X path.
//
//   tgvk_deshade_downsampled_application input_image_path f_downsampling order lambda
//                                        iteration_count [mask] output_image_prefix
//
// Same positional arguments, usage text and outputs as the reference program
// (terminal_programs/tgvk_deshade_downsampled_main.cpp:33-172): weights
// alpha[i] = i + 1, negative values set to zero, background added back, then
// TGVKDeshadeDownsampledProcessor::processTGVKL1Cuda
// (tgvk_deshade_downsampled/TGVKDeshadeDownsampledProcessor.cpp:29-129) - the
// masked input resized by f_downsampling with cubic B-splines, the masked TGV-k
// processor there, the shading resized back and subtracted.  Writes
// <prefix>_shading.mha and <prefix>_deshaded.mha.  Images are MetaImage files;
// outputs keep the input's spacing and origin.
#include "ambient_amd.h"
#include "metaimage.hpp"

#include <chrono>
#include <cstdio>
#include <iostream>
#include <string>
#include <vector>

static void print_usage() {
    std::cout << "parameter: input_image_path f_downsampling order lambda iteration_count [mask] "
              << "output_image_prefix";
}

int main(int argc, char* argv[]) {
    std::cout << "started program: " << argv[0] << std::endl;
    if (argc < 7) {
        print_usage();
        return 1;
    }
    const std::string input_path = argv[1];
    const float f_downsampling = std::stof(argv[2]);
    const int order = std::stoi(argv[3]);
    const float lambda = std::stof(argv[4]);
    const int iteration_count = std::stoi(argv[5]);
    const bool mask_given = argc == 8;
    const std::string mask_path = mask_given ? argv[6] : "";
    const std::string prefix = mask_given ? argv[7] : argv[6];
    std::cout << "input_image: " << input_path << std::endl;
    std::cout << "f_downsampling: " << f_downsampling << std::endl;
    std::cout << "order: " << order << std::endl;
    std::cout << "lambda: " << lambda << std::endl;
    std::cout << "iteration_count: " << iteration_count << std::endl;
    std::cout << "mask_path: " << mask_path << std::endl;
    std::cout << "output_image_prefix: " << prefix << std::endl;

    metaimage::Volume in;
    try {
        in = metaimage::read(input_path);
    } catch (const std::exception& e) {
        std::cerr << "input image file does not exist: " << input_path << std::endl;
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
    if (order < 2 || iteration_count < 0 || !(f_downsampling > 0.f)) {
        std::cerr << "order must be >= 2, iteration_count >= 0, f_downsampling > 0" << std::endl;
        return 1;
    }
    std::cout << "mask image is null: " << !mask_given << std::endl;

    std::vector<float> alpha(static_cast<size_t>(order), 0.0f);
    for (int i = 0; i < order; ++i) alpha[size_t(i)] = float(i + 1);

    amb_context* ctx = nullptr;
    if (amb_context_create(0, &ctx) != AMB_OK) {
        std::printf("ambient_amd error: %s\n", amb_last_error());
        return 5;
    }
    amb_tgv_params p;
    amb_tgv_params_default(&p);   // tau = sigma = 1/sqrt(20), theta = 1, L1, FP32
    p.lambda = lambda;
    p.iterations = uint32_t(iteration_count);
    p.flags = AMB_FLAG_MASKED_ITERATION | AMB_FLAG_SET_NEGATIVE_ZERO | AMB_FLAG_ADD_BACKGROUND_BACK;
    metaimage::Volume l = in, r = in;
    const auto start = std::chrono::system_clock::now();
    const amb_status st = amb_tgvk_deshade_downsampled(
        ctx, AMB_MEM_HOST, in.data.data(), in.w, in.h, in.d, &p, f_downsampling, uint32_t(order),
        alpha.data(), mask_given ? mask.data.data() : nullptr, nullptr, l.data.data(),
        r.data.data(), nullptr, nullptr);
    const std::chrono::duration<double> elapsed = std::chrono::system_clock::now() - start;
    if (st != AMB_OK) {
        std::printf("ambient_amd error: %s\n", amb_last_error());
        amb_context_destroy(ctx);
        return 5;
    }
    amb_context_destroy(ctx);
    std::cout << "elapsed time: " << elapsed.count() << "s\n";
    try {
        metaimage::write(prefix + "_shading.mha", l);
        metaimage::write(prefix + "_deshaded.mha", r);
    } catch (const std::exception& e) {
        std::cerr << e.what() << std::endl;
        return 6;
    }
    return 0;
}
