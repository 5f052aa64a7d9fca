// FP64 energy and primal-dual gap of the TGV2 saddle problem (no counterpart
// in the reference, which records no energy; its closest quantity is the host
// NCC/TV of l in tgv_deshade/TGVDeshadeProcessor.cpp:509-517).
//
// Operators are the ones the iteration applies: grad = D+, eps = eps_B, with
// their true transposes for the dual terms (B is not the exact transpose of D+:
// reference tgv_cuda/tgv_common.cu:146-151).  Per voxel, in FP64:
//   E_data  = lambda|u-f| (L1) or lambda/2 (u-f)^2 (L2)
//   E_first = alpha1 |D+u - v|_2 ,  E_second = alpha0 |eps_B v|_w
//   s = -(D+)^T p ;  b1 = lambda|u-f| - (u-f)s (L1), (lambda(u-f) - s)^2 / (2 lambda) (L2)
//   b2 = E_first - <D+u - v, p> ,  b3 = E_second - <eps_B v, q>_w
//   infeasibility = max(0, |s| - lambda) (L1) ; (eps_B^T q - p)^2 summed -> v residual
// Wave64 shuffle reduction of FP64 partials, one partial row per block, then
// a single-block FP64 sum of the rows (deterministic: fixed order).
#include "amb_internal.hpp"

namespace amb {
namespace {

constexpr int EB = 256;       // threads per block
constexpr int NQ = 10;        // reduced quantities (partials slots 0..9)
constexpr int MAX_BLOCKS = 4096;

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    return v;
}

// the 17 state field pointers, passed by value
struct EnergyFields {
    const void* p[N_FIELDS];
};

template <class S>
__global__ __launch_bounds__(EB) void energy_kernel(
    EnergyFields EF, const float* __restrict__ f, Geometry g,
    double lam, double a0, double a1, int l2, int planar, double* __restrict__ partials) {
    typedef typename S::T T;
    const T* const* F = reinterpret_cast<const T* const*>(EF.p);
    const size_t n = g.n();
    const size_t plane = size_t(g.w) * g.h;
    const size_t st[3] = {1, size_t(g.w), plane};
    const int ext[3] = {g.w, g.h, g.d};
    // planar (depth 1): the 2-D second-order term has no z backward differences
    const int na = planar ? 2 : 3;
    double acc[NQ];
#pragma unroll
    for (int k = 0; k < NQ; ++k) acc[k] = 0.0;
    const typename S::T* u = F[F_U];
    const typename S::T* v[3] = {F[F_VX], F[F_VY], F[F_VZ]};
    const typename S::T* p[3] = {F[F_PX], F[F_PY], F[F_PZ]};
    const typename S::T* q[6] = {F[F_QXX], F[F_QYY], F[F_QZZ], F[F_QXY], F[F_QXZ], F[F_QYZ]};
    // symmetric tensor component (a, c) -> q slot
    const int qi[3][3] = {{0, 3, 4}, {3, 1, 5}, {4, 5, 2}};

    const size_t stride = size_t(gridDim.x) * blockDim.x;
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
        const int xc[3] = {int(i % g.w), int((i / g.w) % g.h), int(i / plane)};
        const double uc = S::ld(u, i);
        double du[3], bv[3][3];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            du[a] = (xc[a] < ext[a] - 1) ? double(S::ld(u, i + st[a])) - uc : 0.0;
#pragma unroll
            for (int cc = 0; cc < 3; ++cc) {
                const double vc = S::ld(v[cc], i);
                bv[a][cc] = a >= na ? 0.0 : (xc[a] > 0) ? double(S::ld(v[cc], i - st[a])) - vc : -vc;
            }
        }
        double sdiv = 0.0;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            double t = 0.0;
            if (xc[a] >= 1) t += double(S::ld(p[a], i - st[a]));
            if (xc[a] < ext[a] - 1) t -= double(S::ld(p[a], i));
            sdiv -= t;
        }
        const double umf = uc - double(f[i]);
        if (!l2) {
            acc[1] += lam * fabs(umf);
            acc[5] += lam * fabs(umf) - umf * sdiv;
            acc[8] += fmax(0.0, fabs(sdiv) - lam);
        } else {
            acc[1] += 0.5 * lam * umf * umf;
            const double r = lam * umf - sdiv;
            acc[5] += r * r / (2.0 * lam);
        }
        double gn = 0.0, gp = 0.0;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const double ga = du[a] - double(S::ld(v[a], i));
            gn += ga * ga;
            gp += ga * double(S::ld(p[a], i));
        }
        gn = sqrt(gn);
        acc[2] += a1 * gn;
        acc[6] += a1 * gn - gp;
        const double e[6] = {bv[0][0], bv[1][1], bv[2][2], 0.5 * (bv[0][1] + bv[1][0]),
                             0.5 * (bv[0][2] + bv[2][0]), 0.5 * (bv[1][2] + bv[2][1])};
        double en = 0.0, eq = 0.0;
#pragma unroll
        for (int k = 0; k < 6; ++k) {
            const double wgt = k < 3 ? 1.0 : 2.0;
            en += wgt * e[k] * e[k];
            eq += wgt * e[k] * double(S::ld(q[k], i));
        }
        en = sqrt(en);
        acc[3] += a0 * en;
        acc[7] += a0 * en - eq;
        for (int cc = 0; cc < na; ++cc) {
            double t = 0.0;
            for (int a = 0; a < na; ++a) {
                const typename S::T* qq = q[qi[a][cc]];
                double bt = -double(S::ld(qq, i));
                if (xc[a] < ext[a] - 1) bt += double(S::ld(qq, i + st[a]));
                t += bt;
            }
            const double r = t - double(S::ld(p[cc], i));
            acc[9] += r * r;
        }
    }
    acc[0] = acc[1] + acc[2] + acc[3];
    acc[4] = acc[5] + acc[6] + acc[7];

    __shared__ double red[EB / 64][NQ];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < NQ; ++k) {
        const double s = wave_sum(acc[k]);
        if (lane == 0) red[wv][k] = s;
    }
    __syncthreads();
    if (threadIdx.x < NQ) {
        double s = 0.0;
        for (int w = 0; w < EB / 64; ++w) s += red[w][threadIdx.x];
        partials[size_t(blockIdx.x) * NQ + threadIdx.x] = s;
    }
}

__global__ void energy_finish_kernel(const double* __restrict__ partials, int nblocks,
                                     double* __restrict__ out) {
    // one wave per quantity, fixed summation order
    const int k = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (k >= NQ) return;
    double s = 0.0;
    for (int b = lane; b < nblocks; b += 64) s += partials[size_t(b) * NQ + k];
    s = wave_sum(s);
    if (lane == 0) out[k] = (k == 9) ? sqrt(s) : s;
}

inline int energy_blocks(const Geometry& g) {
    size_t b = (g.n() + EB - 1) / EB;
    if (b > MAX_BLOCKS) b = MAX_BLOCKS;
    return int(b ? b : 1);
}

template <class S>
hipError_t energy_impl(hipStream_t s, const void* const* fields, const float* f, Geometry g,
                       double lam, double a0, double a1, int l2, int planar, double* partials,
                       double* out) {
    const int nb = energy_blocks(g);
    EnergyFields EF;
    for (int k = 0; k < N_FIELDS; ++k) EF.p[k] = fields[k];
    hipLaunchKernelGGL(energy_kernel<S>, dim3(nb), dim3(EB), 0, s, EF, f, g, lam, a0, a1, l2,
                       planar, partials);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(energy_finish_kernel, dim3(1), dim3(64 * NQ), 0, s, partials, nb, out);
    return hipGetLastError();
}

}  // namespace

size_t energy_partials_size(Geometry g) { return size_t(energy_blocks(g)) * NQ + 16; }

hipError_t launch_energy_finish(hipStream_t s, const double* partials, int rows, double* out10) {
    hipLaunchKernelGGL(energy_finish_kernel, dim3(1), dim3(64 * NQ), 0, s, partials, rows, out10);
    return hipGetLastError();
}

// `fields`: host array of the 17 device field pointers.
hipError_t launch_energy(hipStream_t s, int storage, const void* const* fields, const float* f,
                         Geometry g, double lambda, double alpha0, double alpha1, int data_term,
                         int planar, double* partials, double* out10) {
    if (storage == 1)
        return energy_impl<StoreBF16>(s, fields, f, g, lambda, alpha0, alpha1, data_term == 1,
                                      planar, partials, out10);
    return energy_impl<StoreF32>(s, fields, f, g, lambda, alpha0, alpha1, data_term == 1,
                                 planar, partials, out10);
}

}  // namespace amb
