// A caller written against the reference's launcher templates only: it
// composes TGVDeshadeProcessor::processTGV2L1GPUCuda (CLI overload, reference
// tgv_deshade/TGVDeshadeProcessor.cpp:122-166) from tgv2_l1_deshade_launch,
// divergence, cosine transform, Poisson divide and inverse transform, each a
// host-array launcher as in the reference, then r = u - l and the clamp.
// Compiled against include/ambient_amd_reference_compat.hpp + libambient_amd.so.
//
// With a mask it composes TGVDeshadeMaskedProcessor::processTGV2L1GPUCuda
// (reference tgv_deshade_masked/TGVDeshadeMaskedProcessor.cpp:76-210) from
// buildMaskIndices-style index lists and tgv2_l1_deshade_masked_launch.
//
// --2d (depth 1): the 2-D launchers instead - tgv2_l1_deshade_launch_2d
// (the widget's depth-1 path, TGVDeshadeProcessor.cpp:316-378) or, with a
// mask, buildMaskIndices2D-style lists into tgv2_l1_deshade_masked_2d_launch
// (TGVDeshadeMaskedProcessor.cpp:215-335), and divergence_2d_kernel_launch.
//
// --tgvk a0,a1,...,a(k-1) (k >= 3, no mask): TGVKDeshadeProcessor::processTGVKL1Cuda
// (TGVKDeshadeProcessor.cpp:76-174) from tgvk_l1_deshade_launch (with --2d:
// tgvk_l1_deshade_launch_2d); alpha0/alpha1 are then ignored.
//
//   processor_demo [--2d] [--tgvk a0,a1,...] in.raw W H D lambda alpha0 alpha1 iterations
//                  out_u.raw out_l.raw out_r.raw [mask.raw]
#include "ambient_amd_reference_compat.hpp"

#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

static std::vector<float> read_raw(const char* path, size_t n) {
    std::vector<float> v(n);
    FILE* fp = std::fopen(path, "rb");
    if (!fp || std::fread(v.data(), sizeof(float), n, fp) != n) {
        std::fprintf(stderr, "cannot read %s\n", path);
        std::exit(2);
    }
    std::fclose(fp);
    return v;
}

static void write_raw(const char* path, const float* p, size_t n) {
    FILE* fp = std::fopen(path, "wb");
    if (!fp || std::fwrite(p, sizeof(float), n, fp) != n) {
        std::fprintf(stderr, "cannot write %s\n", path);
        std::exit(3);
    }
    std::fclose(fp);
}

int main(int argc, char** argv) {
    bool two_d = false;
    std::vector<float> kalpha;
    for (;;) {
        if (argc > 1 && std::string(argv[1]) == "--2d") {
            two_d = true;
            ++argv; --argc;
        } else if (argc > 2 && std::string(argv[1]) == "--tgvk") {
            const std::string list = argv[2];
            size_t pos = 0;
            while (pos <= list.size()) {
                const size_t e = list.find(',', pos);
                kalpha.push_back(float(std::atof(list.substr(pos, e - pos).c_str())));
                if (e == std::string::npos) break;
                pos = e + 1;
            }
            argv += 2; argc -= 2;
        } else {
            break;
        }
    }
    if (argc != 12 && argc != 13) {
        std::fprintf(stderr, "usage: %s in.raw W H D lambda alpha0 alpha1 iterations u l r [mask]\n",
                     argv[0]);
        return 1;
    }
    const unsigned W = unsigned(std::atoi(argv[2])), H = unsigned(std::atoi(argv[3])),
                   D = unsigned(std::atoi(argv[4]));
    const float lambda = float(std::atof(argv[5])), alpha0 = float(std::atof(argv[6])),
                alpha1 = float(std::atof(argv[7]));
    const unsigned iterations = unsigned(std::atoi(argv[8]));
    const size_t n = size_t(W) * H * D;
    if (two_d && D != 1) {
        std::fprintf(stderr, "--2d needs D = 1\n");
        return 1;
    }
    std::vector<float> f = read_raw(argv[1], n);

    const bool masked = argc == 13;
    if (!kalpha.empty() && (masked || kalpha.size() < 3)) {
        std::fprintf(stderr, "--tgvk needs at least three weights and no mask\n");
        return 1;
    }
    std::vector<float> mask = masked ? read_raw(argv[12], n) : std::vector<float>();
    float *v_x, *v_y, *v_z = nullptr;
    float* u = nullptr;
    const unsigned order = unsigned(kalpha.size());
    if (order && two_d) {
        u = tgvk_l1_deshade_launch_2d<float>(f.data(), W, H, lambda, iterations, iterations + 1,
                                             nullptr, order, kalpha.data(), &v_x, &v_y);
    } else if (order) {
        u = tgvk_l1_deshade_launch<float>(f.data(), W, H, D, lambda, iterations, iterations + 1,
                                          nullptr, order, kalpha.data(), &v_x, &v_y, &v_z);
    } else if (two_d && !masked) {
        u = tgv2_l1_deshade_launch_2d<float>(f.data(), W, H, lambda, iterations, iterations + 1,
                                             nullptr, alpha0, alpha1, &v_x, &v_y);
    } else if (two_d) {
        // TGVDeshadeMaskedProcessor.cpp:388-421
        typedef std::vector<unsigned> IV;
        IV inside, e[8];
        for (unsigned y = 0; y < H; ++y)
            for (unsigned x = 0; x < W; ++x) {
                const unsigned i = y * W + x;
                if (mask[i] < 1) continue;
                inside.push_back(i);
                e[x == 0 ? 0 : 1].push_back(i);
                e[x == W - 1 ? 2 : 3].push_back(i);
                e[y == 0 ? 4 : 5].push_back(i);
                e[y == H - 1 ? 6 : 7].push_back(i);
            }
        u = tgv2_l1_deshade_masked_2d_launch<float>(f.data(), W, H, lambda, iterations,
                                                    iterations + 1, -1, nullptr, alpha0, alpha1,
                                                    &v_x, &v_y, inside, e[0], e[1], e[2], e[3],
                                                    e[4], e[5], e[6], e[7]);
    } else if (!masked) {
        // TGVDeshadeProcessor.cpp:131-141: the iteration, paint interval = iterations + 1
        u = tgv2_l1_deshade_launch<float>(f.data(), W, H, D, lambda, iterations, iterations + 1,
                                          -1, nullptr, alpha0, alpha1, &v_x, &v_y, &v_z);
    } else {
        // TGVDeshadeMaskedProcessor.cpp:337-385: index lists of the voxels with mask >= 1
        typedef std::vector<unsigned> IV;
        IV inside, e[12];
        for (unsigned z = 0; z < D; ++z)
            for (unsigned y = 0; y < H; ++y)
                for (unsigned x = 0; x < W; ++x) {
                    const unsigned i = (z * H + y) * W + x;
                    if (mask[i] < 1) continue;
                    inside.push_back(i);
                    e[x == 0 ? 0 : 1].push_back(i);
                    e[x == W - 1 ? 2 : 3].push_back(i);
                    e[y == 0 ? 4 : 5].push_back(i);
                    e[y == H - 1 ? 6 : 7].push_back(i);
                    e[z == 0 ? 8 : 9].push_back(i);
                    e[z == D - 1 ? 10 : 11].push_back(i);
                }
        // the unmasked f goes in (TGVDeshadeMaskedProcessor.cpp:99)
        u = tgv2_l1_deshade_masked_launch<float>(f.data(), W, H, D, lambda, iterations,
                                                 iterations + 1, -1, nullptr, alpha0, alpha1, &v_x,
                                                 &v_y, &v_z, inside, e[0], e[1], e[2], e[3], e[4],
                                                 e[5], e[6], e[7], e[8], e[9], e[10], e[11]);
    }
    // TGVDeshadeProcessor.cpp:386-408: Poisson integration of v
    float* g = two_d ? divergence_2d_kernel_launch<float>(v_x, v_y, W, H, true)
                     : divergence_kernel_launch<float>(v_x, v_y, v_z, W, H, D, true);
    float* G = cosine_transform_kernel_launch<float>(g, W, H, D);
    float* P = solve_poisson_in_cosine_domain_kernel_launch<float>(G, W, H, D);
    float* l = inverse_cosine_transform_kernel_launch<float>(P, W, H, D);
    // TGVDeshadeProcessor.cpp:159-165 (and :83-93 with the mask): r = u - l,
    // l and r times the mask, negatives to zero; masked processor: r times the
    // mask again (TGVDeshadeMaskedProcessor.cpp:205-209)
    std::vector<float> r(n);
    for (size_t i = 0; i < n; ++i) {
        float ri = u[i] - l[i];
        if (masked) {
            l[i] = l[i] * mask[i];
            ri = ri * mask[i];
        }
        if (ri < 0) ri = 0;
        if (masked) ri = ri * mask[i];
        r[i] = ri;
    }
    write_raw(argv[9], u, n);
    write_raw(argv[10], l, n);
    write_raw(argv[11], r.data(), n);
    delete[] u; delete[] v_x; delete[] v_y; delete[] v_z;
    delete[] g; delete[] G; delete[] P; delete[] l;
    std::printf("processor_demo: %ux%ux%u, %u iterations done\n", W, H, D, iterations);
    return 0;
}
