// Cubic B-spline resampling on gfx950, as ResizeProcessor::process(image, w, h,
// d, BSpline3) (reference resize/ResizeProcessor.cpp:41-152) computes it through
// ITK 4.8.1: itk::BSplineDecompositionImageFilter coefficients, then
// itk::ResampleImageFilter with itk::BSplineInterpolateImageFunction (order 3).
// The algorithm (M. Unser, "Splines: a perfect fit for signal and image
// processing", IEEE SPM 1999, Box 2) in FP64:
//   coefficients - per axis x, y, z (skipped for length 1), one thread per line
//     (x: lines staged through LDS in chunks):
//     gain (1 - z)(1 - 1/z), pole z = sqrt(3) - 2, mirror-boundary causal start
//     (18-term truncated sum for lines longer than 18, exact otherwise), causal
//     and anticausal recursions;
//   resampling  - one thread per output voxel: continuous input index
//     j * spacing per axis (spacing = 1 / (float(new) / float(old)), with the
//     +2 geometry of ResizeProcessor's upsampling branch), 0 outside
//     [-0.5, N - 0.5), the 4 x 4 x 4 cubic B-spline weights and mirrored
//     support, products summed x fastest, rounded to float.
// The operations are those of the host restatement in the same order (built
// with -ffp-contract=off), so the results agree bit for bit; pow(z, N - 1) and
// the recursion constants are computed on the host.
// Also the elementwise tail of TGVKDeshadeDownsampledProcessor.cpp:107-126
// (and its not-masked form).
#include "amb_internal.hpp"

#include <cmath>

namespace amb {
namespace {

__global__ void to_double_kernel(const float* __restrict__ in, double* __restrict__ c, size_t n) {
    const size_t stride = size_t(gridDim.x) * blockDim.x;
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride)
        c[i] = double(in[i]);
}

struct Spline3Line {
    double z, gain, pole_factor, z2n0;   // z2n0 = z^(N-1) for the exact start
    long horizon;
};

__global__ void spline3_decompose_kernel(double* __restrict__ c, Geometry g, int axis, long lines,
                                         Spline3Line sp) {
    const long l = long(blockIdx.x) * blockDim.x + threadIdx.x;
    if (l >= lines) return;
    const long W = g.w, H = g.h;
    const long N = axis == 0 ? W : (axis == 1 ? H : long(g.d));
    const long st = axis == 0 ? 1 : (axis == 1 ? W : W * H);
    long base;
    if (axis == 0) base = l * W;
    else if (axis == 1) base = (l / W) * W * H + (l % W);
    else base = l;
    double* a = c + base;
    const double z = sp.z, gain = sp.gain;
    // the gain is applied where each sample is first read (the same products as
    // a separate scaling pass, one memory pass fewer)
    double prev;
    if (sp.horizon < N) {
        double zn = z, sum = a[0] * gain;
        for (long n = 1; n < sp.horizon; ++n) {
            sum += zn * (a[n * st] * gain);
            zn *= z;
        }
        prev = sum;
    } else {
        double zn = z;
        const double iz = 1.0 / z;
        double z2n = sp.z2n0;
        double sum = a[0] * gain + z2n * (a[(N - 1) * st] * gain);
        z2n *= z2n * iz;
        for (long n = 1; n <= N - 2; ++n) {
            sum += (zn + z2n) * (a[n * st] * gain);
            zn *= z;
            z2n *= iz;
        }
        prev = sum / (1.0 - zn * zn);
    }
    a[0] = prev;
    double last = 0.0;
    for (long n = 1; n < N; ++n) {
        const double v = a[n * st] * gain + z * prev;
        a[n * st] = v;
        last = prev;
        prev = v;
    }
    double next = sp.pole_factor * (z * last + prev);
    a[(N - 1) * st] = next;
    for (long n = N - 2; n >= 0; --n) {
        next = z * (next - a[n * st]);
        a[n * st] = next;
    }
}

// The x prefilter (lines contiguous in memory): one wave per 64 lines walks
// along x in chunks of 32 samples staged through LDS, so global loads and
// stores are coalesced, and each lane runs its line's recursions from LDS with
// the operations of spline3_decompose_kernel in the same order: a forward
// pass over the chunks (gain, causal start in chunk 0 - the 18-term start and
// the exact start of lines up to 18 samples both lie there - and the causal
// recursion), then a backward pass (anticausal start from the last two causal
// values, kept in registers, and the anticausal recursion).
constexpr int XD_LINES = 64, XD_CHUNK = 32, XD_PITCH = XD_CHUNK + 1;

__global__ __launch_bounds__(XD_LINES) void spline3_decompose_x_kernel(double* __restrict__ c,
                                                                        long N, long lines,
                                                                        Spline3Line sp) {
    __shared__ double tile[XD_LINES * XD_PITCH];
    const int lane = int(threadIdx.x);
    const long l0 = long(blockIdx.x) * XD_LINES;
    const bool has = l0 + lane < lines;
    const long nch = (N + XD_CHUNK - 1) / XD_CHUNK;
    const double z = sp.z;
    double* const a = tile + lane * XD_PITCH;
    // cooperative chunk copies: lanes 0-31 one line, 32-63 the next
    const int cl = lane / XD_CHUNK, cx = lane % XD_CHUNK;
    auto load = [&](long x0, int cw) {
        for (int r = 0; r < XD_LINES; r += 2) {
            const long gl = l0 + r + cl;
            if (gl < lines && cx < cw) tile[(r + cl) * XD_PITCH + cx] = c[gl * N + x0 + cx];
        }
        __syncthreads();
    };
    auto store = [&](long x0, int cw) {
        __syncthreads();
        for (int r = 0; r < XD_LINES; r += 2) {
            const long gl = l0 + r + cl;
            if (gl < lines && cx < cw) c[gl * N + x0 + cx] = tile[(r + cl) * XD_PITCH + cx];
        }
        __syncthreads();
    };

    double prev = 0.0, last = 0.0;   // causal values at n - 1 (and the last two at the end)
    for (long ch = 0; ch < nch; ++ch) {
        const long x0 = ch * XD_CHUNK;
        const int cw = int(min(long(XD_CHUNK), N - x0));
        load(x0, cw);
        if (has) {
            for (int n = 0; n < cw; ++n) a[n] *= sp.gain;
            int n0 = 0;
            if (ch == 0) {
                if (sp.horizon < N) {
                    double zn = z, sum = a[0];
                    for (long n = 1; n < sp.horizon; ++n) {
                        sum += zn * a[n];
                        zn *= z;
                    }
                    a[0] = sum;
                } else {   // N <= 18 <= XD_CHUNK: the whole line is in this chunk
                    double zn = z;
                    const double iz = 1.0 / z;
                    double z2n = sp.z2n0;
                    double sum = a[0] + z2n * a[N - 1];
                    z2n *= z2n * iz;
                    for (long n = 1; n <= N - 2; ++n) {
                        sum += (zn + z2n) * a[n];
                        zn *= z;
                        z2n *= iz;
                    }
                    a[0] = sum / (1.0 - zn * zn);
                }
                prev = a[0];
                n0 = 1;
            }
            for (int n = n0; n < cw; ++n) {
                a[n] += z * prev;
                last = prev;
                prev = a[n];
            }
        }
        store(x0, cw);
    }
    // prev = causal c[N - 1], last = causal c[N - 2]
    double next = 0.0;
    for (long ch = nch - 1; ch >= 0; --ch) {
        const long x0 = ch * XD_CHUNK;
        const int cw = int(min(long(XD_CHUNK), N - x0));
        load(x0, cw);
        if (has) {
            int n = cw - 1;
            if (ch == nch - 1) {
                a[n] = sp.pole_factor * (z * last + prev);
                next = a[n];
                --n;
            }
            for (; n >= 0; --n) {
                a[n] = z * (next - a[n]);
                next = a[n];
            }
        }
        store(x0, cw);
    }
}

// Whole-sample mirror about 0 and N - 1 (period 2N - 2); all 0 for N = 1.
// (Volumes hold fewer than 2^31 voxels, so indices fit in int.)
__device__ __forceinline__ int mirror_index(int k, int N) {
    if (k >= 0 && k < N) return k;   // interior taps: no reflection
    if (N == 1) return 0;
    const int period = 2 * (N - 1);
    k %= period;
    if (k < 0) k += period;
    return k >= N ? period - k : k;
}

__global__ void spline3_resample_kernel(const double* __restrict__ c, Geometry gi,
                                        float* __restrict__ out, Geometry go, double sp0,
                                        double sp1, double sp2) {
    const uint32_t n = uint32_t(go.n());
    const uint32_t stride = gridDim.x * blockDim.x;
    const uint32_t OW = uint32_t(go.w), OH = uint32_t(go.h);
    const int ext[3] = {gi.w, gi.h, gi.d};
    const double spacing[3] = {sp0, sp1, sp2};
    for (uint32_t o = blockIdx.x * blockDim.x + threadIdx.x; o < n; o += stride) {
        const uint32_t oyz = o / OW;
        const uint32_t oc[3] = {o - oyz * OW, oyz % OH, oyz / OH};
        double x[3];
        bool inside = true;
        for (int a = 0; a < 3; ++a) {
            x[a] = double(oc[a]) * spacing[a];
            if (!(x[a] >= -0.5 && x[a] < double(ext[a]) - 0.5)) inside = false;
        }
        if (!inside) {
            out[o] = 0.0f;
            continue;
        }
        int idx[3][4];
        double wt[3][4];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const int i1 = int(floor(x[a]));
            const double w = x[a] - double(i1);
            wt[a][3] = (1.0 / 6.0) * w * w * w;
            wt[a][0] = (1.0 / 6.0) + 0.5 * w * (w - 1.0) - wt[a][3];
            wt[a][2] = w + wt[a][0] - 2.0 * wt[a][3];
            wt[a][1] = 1.0 - wt[a][0] - wt[a][2] - wt[a][3];
#pragma unroll
            for (int k = 0; k < 4; ++k) idx[a][k] = mirror_index(i1 - 1 + k, ext[a]);
        }
        double acc = 0.0;
#pragma unroll 4
        for (int p = 0; p < 64; ++p) {
            const int k0 = p & 3, k1 = (p >> 2) & 3, k2 = p >> 4;
            double w = 1.0;
            w *= wt[0][k0];
            w *= wt[1][k1];
            w *= wt[2][k2];
            acc += w * c[(idx[2][k2] * ext[1] + idx[1][k1]) * ext[0] + idx[0][k0]];
        }
        out[o] = float(acc);
    }
}

// TGVKDeshadeDownsampledProcessor.cpp:107-126 (masked_variant) and
// TGVKDeshadeDownsampledNotMaskedProcessor.cpp:103-116: r = f' - l with f' the
// masked input (masked variant) or the input; with a mask l (masked variant)
// and r times the mask; the add-back of f * (1 - mask); then the clamp.
__global__ void downsampled_tail_kernel(const float* __restrict__ f, const float* __restrict__ mask,
                                        float* __restrict__ l, float* __restrict__ r, size_t n,
                                        int masked_variant, int add_background_back,
                                        int set_negative_zero) {
    const size_t stride = size_t(gridDim.x) * blockDim.x;
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
        const float fv = f[i];
        float ri;
        if (mask) {
            const float m = mask[i];
            const float fin = masked_variant ? fv * m : fv;
            ri = fin - l[i];
            if (masked_variant) l[i] = l[i] * m;
            ri = ri * m;
            if (add_background_back) ri = ri + fv * float(1.0 - double(m));
        } else {
            ri = fv - l[i];
        }
        if (set_negative_zero && ri < 0.0f) ri = 0.0f;
        r[i] = ri;
    }
}

__global__ void multiply_kernel(const float* __restrict__ a, const float* __restrict__ b,
                                float* __restrict__ out, size_t n) {
    const size_t stride = size_t(gridDim.x) * blockDim.x;
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride)
        out[i] = a[i] * b[i];
}

inline int grid_for(size_t n) {
    size_t b = (n + 255) / 256;
    if (b > 16384) b = 16384;
    return int(b ? b : 1);
}

}  // namespace

// Resample in (gi) to out (go, the kept voxels), coef: gi.n() doubles of scratch.
hipError_t launch_resize_spline3(hipStream_t s, const float* in, Geometry gi, double* coef,
                                 float* out, Geometry go) {
    const size_t n = gi.n();
    hipLaunchKernelGGL(to_double_kernel, dim3(grid_for(n)), dim3(256), 0, s, in, coef, n);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    Spline3Line sp;
    sp.z = std::sqrt(3.0) - 2.0;
    sp.gain = (1.0 - sp.z) * (1.0 - 1.0 / sp.z);
    sp.pole_factor = sp.z / (sp.z * sp.z - 1.0);
    sp.horizon = long(std::ceil(std::log(1e-10) / std::log(std::fabs(sp.z))));
    const long ext[3] = {gi.w, gi.h, gi.d};
    for (int axis = 0; axis < 3; ++axis) {
        const long N = ext[axis];
        if (N == 1) continue;
        sp.z2n0 = std::pow(sp.z, double(N - 1));
        const long lines = long(n) / N;
        if (axis == 0)
            hipLaunchKernelGGL(spline3_decompose_x_kernel,
                               dim3(unsigned((lines + XD_LINES - 1) / XD_LINES)), dim3(XD_LINES), 0,
                               s, coef, N, lines, sp);
        else
            hipLaunchKernelGGL(spline3_decompose_kernel, dim3(unsigned((lines + 255) / 256)),
                               dim3(256), 0, s, coef, gi, axis, lines, sp);
        e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    const bool up = go.w > gi.w || go.h > gi.h || go.d > gi.d;
    const int pad = up ? 2 : 0;
    const long geo[3] = {long(go.w) + pad, long(go.h) + pad, long(go.d) + pad};
    double spacing[3];
    for (int a = 0; a < 3; ++a) spacing[a] = 1.0 / double(float(geo[a]) / float(ext[a]));
    hipLaunchKernelGGL(spline3_resample_kernel, dim3(grid_for(go.n())), dim3(256), 0, s, coef, gi,
                       out, go, spacing[0], spacing[1], spacing[2]);
    return hipGetLastError();
}

hipError_t launch_downsampled_tail(hipStream_t s, const float* f, const float* mask, float* l,
                                   float* r, size_t n, int masked_variant, int add_background_back,
                                   int set_negative_zero) {
    hipLaunchKernelGGL(downsampled_tail_kernel, dim3(grid_for(n)), dim3(256), 0, s, f, mask, l, r,
                       n, masked_variant, add_background_back, set_negative_zero);
    return hipGetLastError();
}

// out = a * b (the reference's multiply.cu, FP32)
hipError_t launch_multiply(hipStream_t s, const float* a, const float* b, float* out, size_t n) {
    hipLaunchKernelGGL(multiply_kernel, dim3(grid_for(n)), dim3(256), 0, s, a, b, out, n);
    return hipGetLastError();
}

}  // namespace amb
