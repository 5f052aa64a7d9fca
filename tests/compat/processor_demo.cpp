// A caller written against the reference's launcher templates only: it
// composes TGVDeshadeProcessor::processTGV2L1GPUCuda (CLI overload, reference
// tgv_deshade/TGVDeshadeProcessor.cpp:122-166) from tgv2_l1_deshade_launch,
// divergence, cosine transform, Poisson divide and inverse transform, each a
// host-array launcher as in the reference, then r = u - l and the clamp.
// Compiled against include/ambient_amd_reference_compat.hpp + libambient_amd.so.
//
//   processor_demo in.raw W H D lambda alpha0 alpha1 iterations out_u.raw out_l.raw out_r.raw
#include "ambient_amd_reference_compat.hpp"

#include <cstdio>
#include <cstdlib>
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
    if (argc != 12) {
        std::fprintf(stderr, "usage: %s in.raw W H D lambda alpha0 alpha1 iterations u l r\n", argv[0]);
        return 1;
    }
    const unsigned W = unsigned(std::atoi(argv[2])), H = unsigned(std::atoi(argv[3])),
                   D = unsigned(std::atoi(argv[4]));
    const float lambda = float(std::atof(argv[5])), alpha0 = float(std::atof(argv[6])),
                alpha1 = float(std::atof(argv[7]));
    const unsigned iterations = unsigned(std::atoi(argv[8]));
    const size_t n = size_t(W) * H * D;
    std::vector<float> f = read_raw(argv[1], n);

    // TGVDeshadeProcessor.cpp:131-141: the iteration, paint interval = iterations + 1
    float *v_x, *v_y, *v_z;
    float* u = tgv2_l1_deshade_launch<float>(f.data(), W, H, D, lambda, iterations,
                                             iterations + 1, -1, nullptr, alpha0, alpha1, &v_x,
                                             &v_y, &v_z);
    // TGVDeshadeProcessor.cpp:386-408: Poisson integration of v
    float* g = divergence_kernel_launch<float>(v_x, v_y, v_z, W, H, D, true);
    float* G = cosine_transform_kernel_launch<float>(g, W, H, D);
    float* P = solve_poisson_in_cosine_domain_kernel_launch<float>(G, W, H, D);
    float* l = inverse_cosine_transform_kernel_launch<float>(P, W, H, D);
    // TGVDeshadeProcessor.cpp:159-165: r = u - l, negatives to zero (no mask)
    std::vector<float> r(n);
    for (size_t i = 0; i < n; ++i) {
        float ri = u[i] - l[i];
        if (ri < 0) ri = 0;
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
