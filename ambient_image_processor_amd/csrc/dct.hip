// DCT Poisson bias solve for gfx950: B-divergence, separable DCT-II/III
// passes, Laplacian-eigenvalue divide, and the deshade epilogue.
//
// Reference: cuda_image_operations/divergence.cu:26-56, cosine_transform.cu:142-223,
// solve_poisson_in_cosine_domain.cu:27-54, inverse_cosine_transform.cu:79-164,
// subtract.cu:27, multiply.cu:25, clamp_negative_values.cu:27, add.cu:27,
// invert.cu:26; composition tgv_deshade/TGVDeshadeProcessor.cpp:71-95,122-166,386-408.
//
// Built with -ffp-contract=off: every fused multiply-add is written as fma.
#include "amb_internal.hpp"

#include <cmath>

namespace amb {

double host_cospi(double t) {
    t = std::fabs(t);
    double r = std::fmod(t, 2.0);
    if (r > 1.0) r = 2.0 - r;
    double sgn = 1.0;
    if (r > 0.5) { r = 1.0 - r; sgn = -1.0; }
    if (r == 0.5) return 0.0;
    const double v = (r <= 0.25) ? std::cos(M_PI * r) : std::sin(M_PI * (0.5 - r));
    return sgn * v;
}

namespace {

__device__ __forceinline__ void axis_info(const Geometry& g, int axis, size_t& stride, int& n) {
    stride = axis == 0 ? 1 : (axis == 1 ? size_t(g.w) : size_t(g.w) * g.h);
    n = axis == 0 ? g.w : (axis == 1 ? g.h : g.d);
}

// A14: g = (B_x v_x + B_y v_y) + B_z v_z (B_z only when depth > 1),
// B a = -a[i] + a[i-1], -a[0] at i = 0 (tgv_cuda/tgv_common.cu:133-197).
__global__ void divergence_kernel(const float* __restrict__ vx, const float* __restrict__ vy,
                                  const float* __restrict__ vz, float* __restrict__ out,
                                  Geometry g) {
    const size_t n = g.n();
    const size_t plane = size_t(g.w) * g.h;
    const size_t stride = size_t(gridDim.x) * blockDim.x;
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
        const int x = int(i % g.w);
        const int y = int((i / g.w) % g.h);
        const int z = int(i / plane);
        const float bx = x > 0 ? -vx[i] + vx[i - 1] : -vx[i];
        const float by = y > 0 ? -vy[i] + vy[i - g.w] : -vy[i];
        float r = bx + by;
        if (g.d > 1) {
            const float bz = z > 0 ? -vz[i] + vz[i - plane] : -vz[i];
            r = r + bz;
        }
        out[i] = r;
    }
}

// A15/A17 in the reference's numerics: every output is its own sequence of
// terms (double)x[m] * cos64 added into an FP32 accumulator (fma in double,
// rounded to float each term), in the reference's term order.
//   forward:  acc = sum_{m=0}^{n-1} in[m] * T[m][k]; last pass: acc *= 8
//   inverse:  acc = sum_{m=1}^{n-1} in[m] * T[m][k]; acc *= 2; acc += in[0];
//             last pass: acc = acc / (W*H*D*8.0) in double
// T[m][k] = cospi((m + 0.5) k / n) forward, cospi((k + 0.5) m / n) inverse.
// A block owns 64 lines x 64 outputs; chunks of 16 terms of the input lines
// (FP32) and of the table (FP64) are staged in LDS and every thread carries a
// 4 x 4 block of accumulators through them.  The per-output sequence is
// unchanged, so the result does not depend on the tiling.  LINE_CONTIG: the x
// axis, whose lines are contiguous (threads then walk the outputs along k).
constexpr int DT_L = 64, DT_K = 64, DT_M = 16;

__device__ __forceinline__ size_t line_base(const Geometry& g, int axis, long l, int n) {
    if (axis == 0) return size_t(l) * size_t(n);
    if (axis == 1) return size_t(l / g.w) * size_t(g.w) * size_t(g.h) + size_t(l % g.w);
    return size_t(l);
}

template <bool INVERSE, bool LINE_CONTIG>
__global__ __launch_bounds__(256) void dct_reference_kernel(
    const float* __restrict__ in, float* __restrict__ out, const double* __restrict__ table,
    Geometry g, int axis, int last, long lines) {
    __shared__ float s_in[DT_M][DT_L];
    __shared__ double s_t[DT_M][DT_K];
    size_t st;
    int n;
    axis_info(g, axis, st, n);
    const int nkt = (n + DT_K - 1) / DT_K;
    // the nkt blocks of one line panel are adjacent tiles on one XCD
    const unsigned tile = xcd_contiguous(blockIdx.x, gridDim.x);
    const long l0 = long(tile / unsigned(nkt)) * DT_L;
    const int k0 = int(tile % unsigned(nkt)) * DT_K;
    const int t = int(threadIdx.x);
    const int tl = LINE_CONTIG ? (t / 16) * 4 : (t % 16) * 4;
    const int tk = LINE_CONTIG ? (t % 16) * 4 : (t / 16) * 4;
    float acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] = 0.0f;

    for (int m0 = INVERSE ? 1 : 0; m0 < n; m0 += DT_M) {
        const int mc = min(DT_M, n - m0);
        __syncthreads();
        for (int e = t; e < DT_M * DT_L; e += 256) {
            const int m = LINE_CONTIG ? e % DT_M : e / DT_L;
            const int l = LINE_CONTIG ? e / DT_M : e % DT_L;
            float v = 0.0f;
            if (m < mc && l0 + l < lines) v = in[line_base(g, axis, l0 + l, n) + size_t(m0 + m) * st];
            s_in[m][l] = v;
        }
        for (int e = t; e < DT_M * DT_K; e += 256) {
            const int m = e / DT_K, k = e % DT_K;
            double v = 0.0;
            if (m < mc && k0 + k < n) v = table[size_t(m0 + m) * size_t(n) + size_t(k0 + k)];
            s_t[m][k] = v;
        }
        __syncthreads();
        for (int m = 0; m < mc; ++m) {
            double xin[4], tv[4];
#pragma unroll
            for (int a = 0; a < 4; ++a) xin[a] = double(s_in[m][tl + a]);
#pragma unroll
            for (int b = 0; b < 4; ++b) tv[b] = s_t[m][tk + b];
#pragma unroll
            for (int a = 0; a < 4; ++a)
#pragma unroll
                for (int b = 0; b < 4; ++b)
                    acc[a][b] = float(fma(xin[a], tv[b], double(acc[a][b])));
        }
    }

    const double vol8 = double(size_t(g.w) * g.h * g.d) * 8.0;
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        const long l = l0 + tl + a;
        if (l >= lines) continue;
        const size_t base = line_base(g, axis, l, n);
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const int k = k0 + tk + b;
            if (k >= n) continue;
            float r = acc[a][b];
            if (!INVERSE) {
                if (last) r = r * 8.0f;
            } else {
                r = r * 2.0f;
                r = r + in[base];
                if (last) r = float(double(r) / vol8);
            }
            out[base + size_t(k) * st] = r;
        }
    }
}

// Voxel coordinates for the Poisson divide.  quirk != 0 reproduces the
// reference's float arithmetic (solve_poisson_in_cosine_domain.cu:36-40),
// which differs from the exact coordinates only for linear index >= 2^24
// (y may then reach H; the cy table has H + 1 entries).
__device__ __forceinline__ void poisson_coords(size_t i, const Geometry& g, int quirk, int& x,
                                               int& y, int& z) {
    const unsigned wh = unsigned(g.w) * unsigned(g.h);
    if (!quirk) {
        x = int(i % g.w); y = int((i / g.w) % g.h); z = int(i / wh);
        return;
    }
    const float zf = floorf(float(unsigned(i) / wh));
    int rest = int(float(int(i)) - zf * float(wh));
    const float yf = floorf(float(unsigned(rest) / unsigned(g.w)));
    rest = int(float(rest) - yf * float(unsigned(g.w)));
    x = rest; y = int(yf); z = int(zf);
}

// A16: X / (6.0 - 2.0*cx[x] - 2.0*cy[y] - 2.0*cz[z]), (0,0,0) -> 0; cx etc. are the
// reference's FP32 cospif(x / W) values, the denominator and quotient are double.
// Depth 1: the reference's 2-D kernel, 4.0 - 2.0*cx[x] - 2.0*cy[y]
// (solve_poisson_in_cosine_domain.cu:57-84, chosen at :107).
__global__ void poisson_kernel(const float* __restrict__ in, float* __restrict__ out,
                               const float* __restrict__ cx, const float* __restrict__ cy,
                               const float* __restrict__ cz, Geometry g, int quirk) {
    const size_t total = g.n();
    const size_t stride = size_t(gridDim.x) * blockDim.x;
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < total; i += stride) {
        int x, y, z;
        poisson_coords(i, g, quirk, x, y, z);
        if (x == 0 && y == 0 && z == 0) { out[i] = 0.0f; continue; }
        const double den =
            g.d == 1 ? 4.0 - 2.0 * double(cx[x]) - 2.0 * double(cy[y])
                     : 6.0 - 2.0 * double(cx[x]) - 2.0 * double(cy[y]) - 2.0 * double(cz[z]);
        out[i] = float(double(in[i]) / den);
    }
}

// A19 epilogue: r = u - l ; mask ; clamp ; add-back (see ambient_amd.h flags).
// mask_again: the masked processor's final r * mask when no add-back
// (tgv_deshade_masked/TGVDeshadeMaskedProcessor.cpp:201-209).
__global__ void epilogue_kernel(const float* __restrict__ u, float* __restrict__ l,
                                const float* __restrict__ mask, float* __restrict__ r, size_t n,
                                int mask_shading, int set_negative_zero, int add_background_back,
                                int mask_again) {
    const size_t stride = size_t(gridDim.x) * blockDim.x;
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
        const float uv = u[i];
        float ri = uv - l[i];
        if (mask) {
            const float m = mask[i];
            if (mask_shading) l[i] = l[i] * m;
            ri = ri * m;
        }
        if (set_negative_zero && ri < 0.0f) ri = 0.0f;
        if (add_background_back && mask) {
            const float bg = float(1.0 - double(mask[i]));
            ri = ri + uv * bg;
        } else if (mask_again && mask) {
            ri = ri * mask[i];
        }
        r[i] = ri;
    }
}

// Active set of the masked iteration: mask >= 1
// (TGVDeshadeMaskedProcessor.cpp:348-353).
__global__ void active_kernel(const float* __restrict__ mask, uint8_t* __restrict__ act, size_t n) {
    const size_t stride = size_t(gridDim.x) * blockDim.x;
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride)
        act[i] = mask[i] >= 1.0f ? uint8_t(1) : uint8_t(0);
}

inline int elem_blocks(size_t n) {
    size_t b = (n + 255) / 256;
    if (b > 16384) b = 16384;
    return int(b ? b : 1);
}

}  // namespace

hipError_t launch_divergence(hipStream_t s, const float* vx, const float* vy, const float* vz,
                             float* out, Geometry g) {
    hipLaunchKernelGGL(divergence_kernel, dim3(elem_blocks(g.n())), dim3(256), 0, s, vx, vy, vz,
                       out, g);
    return hipGetLastError();
}

hipError_t launch_dct_reference_pass(hipStream_t s, const float* in, float* out,
                                     const double* table, Geometry g, int axis, int inverse,
                                     int last) {
    const int n = axis == 0 ? g.w : (axis == 1 ? g.h : g.d);
    const long lines = long(g.n() / size_t(n));
    const long blocks = ((lines + DT_L - 1) / DT_L) * long((n + DT_K - 1) / DT_K);
    const dim3 grid{static_cast<unsigned>(blocks), 1, 1};
    if (axis == 0) {
        if (inverse)
            hipLaunchKernelGGL((dct_reference_kernel<true, true>), grid, dim3(256), 0, s, in, out,
                               table, g, axis, last, lines);
        else
            hipLaunchKernelGGL((dct_reference_kernel<false, true>), grid, dim3(256), 0, s, in, out,
                               table, g, axis, last, lines);
    } else {
        if (inverse)
            hipLaunchKernelGGL((dct_reference_kernel<true, false>), grid, dim3(256), 0, s, in, out,
                               table, g, axis, last, lines);
        else
            hipLaunchKernelGGL((dct_reference_kernel<false, false>), grid, dim3(256), 0, s, in, out,
                               table, g, axis, last, lines);
    }
    return hipGetLastError();
}

hipError_t launch_poisson_divide(hipStream_t s, const float* in, float* out, const float* cx,
                                 const float* cy, const float* cz, Geometry g, int quirk) {
    hipLaunchKernelGGL(poisson_kernel, dim3(elem_blocks(g.n())), dim3(256), 0, s, in, out, cx, cy,
                       cz, g, quirk);
    return hipGetLastError();
}

hipError_t launch_epilogue(hipStream_t s, const float* u, float* l, const float* mask, float* r,
                           size_t n, int mask_shading, int set_negative_zero,
                           int add_background_back, int mask_again) {
    hipLaunchKernelGGL(epilogue_kernel, dim3(elem_blocks(n)), dim3(256), 0, s, u, l, mask, r, n,
                       mask_shading, set_negative_zero, add_background_back, mask_again);
    return hipGetLastError();
}

hipError_t launch_active_from_mask(hipStream_t s, const float* mask, uint8_t* act, size_t n) {
    hipLaunchKernelGGL(active_kernel, dim3(elem_blocks(n)), dim3(256), 0, s, mask, act, n);
    return hipGetLastError();
}

}  // namespace amb
