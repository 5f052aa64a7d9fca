// Hand-written adjoint of the unrolled TGV2-L2 iteration (no counterpart in the
// reference).  The forward map is the loop of tgv_deshade/tgv2_l1_deshade.cu:169-264
// with the L2 u-update of tgv_cuda/tgv2_l2.cu:27-56, exactly as the sweeps in
// tgv_sweeps.hip compute it; the reverse map uses the true transposes of the
// operators as implemented (D+, the reference's B, eps_B, div2).
//
// One reverse iteration k, given x_k (S), x_{k+1} (T) and the adjoint A of
// x_{k+1} (17 fields, the state layout), in three phases:
//   phase A (pointwise): u/v over-relaxation and L2 prox adjoints; writes the new
//            u and v adjoints in place, A* = U'/(1 + tau lambda) and
//            W* = -tau V' into scratch, and the p adjoint plus tau V' into the
//            p buffer (3 fields); FP64 partial of dL/dlambda =
//            sum U' tau (f - u') / (1 + tau lambda).
//   phase B (stencil): P' = (p buffer) - tau B^T A*, Q' = A.q + div2^T W*; projection
//            Jacobians at p~ = p + sigma (D+ u_bar - v_bar), q~ = q + sigma eps_B(v_bar)
//            recomputed from S with the forward's FP32 arithmetic (same active
//            set); writes the p/q adjoints; FP64 partials of dL/dalpha1, dL/dalpha0.
//   phase C (stencil): u_bar adjoint = sigma (D+)^T P~, v_bar adjoint =
//            sigma eps_B^T Q~ - sigma P~.
// Phase A of iteration k - 1 reads only its own voxel's u_bar / v_bar adjoints,
// so it runs in the launch of phase C of iteration k (adjoint_phase_ca) on the
// values in registers: those adjoints are stored only by the last phase C.
// C reads the p adjoint at neighbouring voxels, hence A's separate p buffer.
// The kernels run one block per row piece (x fastest; row index -> y, z once
// per block, 32-bit voxel indices, no per-voxel division).  Scalar gradients:
// each block reduces its voxels' terms (wave64 shuffles) and adds the sum to
// its own FP64 slot, which accumulates over all steps; one fixed-order sum of
// the slots at the end (deterministic).
#include "amb_internal.hpp"

#include <algorithm>

namespace amb {
namespace {

constexpr int AB = 256;   // largest block: one row piece of up to 256 voxels

struct Fields17 {
    float* f[N_FIELDS];
};

__device__ __forceinline__ double wsum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    return v;
}

// Block sums of the terms in acc (blockDim.x a multiple of 64) added to this
// block's slots: slot[blockIdx.x * 3 + OFF + k] += sum (each block owns its
// slots; the phases run in stream order, so no two blocks touch one slot).
template <int NQ, int OFF>
__device__ __forceinline__ void block_accumulate(double (&acc)[NQ], double* slots) {
    __shared__ double red[AB / 64][NQ];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = int(blockDim.x) >> 6;
#pragma unroll
    for (int k = 0; k < NQ; ++k) {
        const double s = wsum(acc[k]);
        if (lane == 0) red[w][k] = s;
    }
    __syncthreads();
    if (threadIdx.x < NQ) {
        double s = 0.0;
        for (int i = 0; i < nw; ++i) s += red[i][threadIdx.x];
        slots[size_t(blockIdx.x) * 3 + OFF + threadIdx.x] += s;
    }
}

// Voxel of this thread: row piece b = xcd_contiguous(blockIdx.x) covers row
// r = b / nbx (y = r % H, z = r / H), columns (b % nbx) * blockDim.x + threadIdx.x.
// The remap gives each XCD a contiguous range of rows, so the y and z neighbour
// rows a block reads were mostly fetched into the same XCD's L2 by its
// neighbouring blocks (round-robin placement fetched them again from HBM).
struct RowVoxel {
    int x, y, z;
    uint32_t i;
    bool in;
    __device__ __forceinline__ RowVoxel(const Geometry& g) {
        const uint32_t nbx = (uint32_t(g.w) + blockDim.x - 1) / blockDim.x;
        const uint32_t b = xcd_contiguous(blockIdx.x, gridDim.x);
        const uint32_t r = b / nbx;
        x = int((b - r * nbx) * blockDim.x + threadIdx.x);
        y = int(r % uint32_t(g.h));
        z = int(r / uint32_t(g.h));
        in = x < g.w;
        i = (uint32_t(z) * uint32_t(g.h) + uint32_t(y)) * uint32_t(g.w) + uint32_t(x);
    }
};

// (D+)^T b [j] = b[j-1] [j >= 1] - b[j] [j < n-1]
__device__ __forceinline__ float dplus_t(const float* b, uint32_t i, uint32_t st, int c, int n) {
    float r = 0.0f;
    if (c >= 1) r += b[i - st];
    if (c < n - 1) r -= b[i];
    return r;
}
// B^T b [j] = b[j+1] [j < n-1] - b[j]
__device__ __forceinline__ float bm_t(const float* b, uint32_t i, uint32_t st, int c, int n) {
    float r = -b[i];
    if (c < n - 1) r += b[i + st];
    return r;
}

// Phase A at voxel i from the u_bar and v_bar adjoints ub, vb: the p adjoint
// plus tau V' goes to pout (3 fields, stride n); the lambda term to acc.
__device__ __forceinline__ void phase_a_voxel(const Fields17& A, uint32_t i, uint32_t n, float ub,
                                              const float (&vb)[3], const float* __restrict__ f,
                                              const float* __restrict__ u_next,
                                              float* __restrict__ ast, float* __restrict__ wst,
                                              float* __restrict__ pout,
                                              float* __restrict__ grad_f, float tau, float tl,
                                              float theta, double& acc) {
    const float un = A.f[F_U][i] + (1.0f + theta) * ub;
    const float as = un / (1.0f + tl);
    A.f[F_U][i] = -theta * ub + as;
    ast[i] = as;
    grad_f[i] += tl / (1.0f + tl) * un;
    acc += double(un) * double(tau) * (double(f[i]) - double(u_next[i])) / (1.0 + double(tl));
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const float vn = A.f[F_VX + c][i] + (1.0f + theta) * vb[c];
        A.f[F_VX + c][i] = -theta * vb[c] + vn;
        wst[size_t(c) * n + i] = -tau * vn;
        pout[size_t(c) * n + i] = A.f[F_PX + c][i] + tau * vn;
    }
}

__global__ __launch_bounds__(AB) void adjoint_phase_a(Fields17 A, const float* __restrict__ f,
                                                      const float* __restrict__ u_next,
                                                      float* __restrict__ ast,
                                                      float* __restrict__ wst,
                                                      float* __restrict__ pout,
                                                      float* __restrict__ grad_f, Geometry g,
                                                      float tau, float lambda, float theta,
                                                      double* __restrict__ slots) {
    const uint32_t n = uint32_t(g.n());
    double acc[1] = {0.0};
    const RowVoxel rv(g);
    if (rv.in) {
        const uint32_t i = rv.i;
        const float vb[3] = {A.f[F_VBX][i], A.f[F_VBY][i], A.f[F_VBZ][i]};
        phase_a_voxel(A, i, n, A.f[F_UBAR][i], vb, f, u_next, ast, wst, pout, grad_f, tau,
                      tau * lambda, theta, acc[0]);
    }
    block_accumulate<1, 2>(acc, slots);   // slot 2: lambda
}

__global__ __launch_bounds__(AB) void adjoint_phase_b(Fields17 A, Fields17 S,
                                                      const float* __restrict__ pin,
                                                      const float* __restrict__ ast,
                                                      const float* __restrict__ wst, Geometry g,
                                                      float tau, float sigma, float alpha0,
                                                      float alpha1, double* __restrict__ slots) {
    const uint32_t n = uint32_t(g.n());
    const uint32_t st[3] = {1u, uint32_t(g.w), uint32_t(g.w) * uint32_t(g.h)};
    const int ext[3] = {g.w, g.h, g.d};
    double acc[2] = {0.0, 0.0};   // alpha0, alpha1
    const RowVoxel rv(g);
    if (rv.in) {
        const uint32_t i = rv.i;
        const int c3[3] = {rv.x, rv.y, rv.z};
        // P' = A.p - tau B^T A*   (component a uses B_a)
        float P[3];
#pragma unroll
        for (int a = 0; a < 3; ++a) P[a] = pin[size_t(a) * n + i] - tau * bm_t(ast, i, st[a], c3[a], ext[a]);
        // Q' = A.q + div2^T W*
        const float* W0 = wst;
        const float* W1 = wst + n;
        const float* W2 = wst + 2 * n;
        float Q[6];
        Q[0] = A.f[F_QXX][i] + dplus_t(W0, i, st[0], c3[0], ext[0]);
        Q[1] = A.f[F_QYY][i] + dplus_t(W1, i, st[1], c3[1], ext[1]);
        Q[2] = A.f[F_QZZ][i] + dplus_t(W2, i, st[2], c3[2], ext[2]);
        Q[3] = A.f[F_QXY][i] + (dplus_t(W0, i, st[1], c3[1], ext[1]) + dplus_t(W1, i, st[0], c3[0], ext[0]));
        Q[4] = A.f[F_QXZ][i] + (dplus_t(W0, i, st[2], c3[2], ext[2]) + dplus_t(W2, i, st[0], c3[0], ext[0]));
        Q[5] = A.f[F_QYZ][i] + (dplus_t(W1, i, st[2], c3[2], ext[2]) + dplus_t(W2, i, st[1], c3[1], ext[1]));

        // p~ from x_k, forward arithmetic (tgv_sweeps.hip dual sweep)
        const float ubc = S.f[F_UBAR][i];
        float pt[3];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const float d = (c3[a] < ext[a] - 1) ? S.f[F_UBAR][i + st[a]] - ubc : 0.0f;
            pt[a] = fmaf(sigma, d - S.f[F_VBX + a][i], S.f[F_PX + a][i]);
        }
        float n2 = pt[0] * pt[0];
        n2 = fmaf(pt[1], pt[1], n2);
        n2 = fmaf(pt[2], pt[2], n2);
        const float npn = sqrtf(n2);
        if (npn / alpha1 > 1.0f) {
            const float dot = P[0] * pt[0] + P[1] * pt[1] + P[2] * pt[2];
            acc[1] += double(dot) / double(npn);
            const float k1 = alpha1 / npn, k2 = dot / (npn * npn);
#pragma unroll
            for (int a = 0; a < 3; ++a) P[a] = k1 * (P[a] - k2 * pt[a]);
        }
        // q~ from x_k
        float bxv[3], byv[3], bzv[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const float* vb = S.f[F_VBX + c];
            const float vc = vb[i];
            bxv[c] = c3[0] > 0 ? -vc + vb[i - st[0]] : -vc;
            byv[c] = c3[1] > 0 ? -vc + vb[i - st[1]] : -vc;
            bzv[c] = c3[2] > 0 ? -vc + vb[i - st[2]] : -vc;
        }
        float qt[6];
        qt[0] = fmaf(sigma, bxv[0], S.f[F_QXX][i]);
        qt[1] = fmaf(sigma, byv[1], S.f[F_QYY][i]);
        qt[2] = fmaf(sigma, bzv[2], S.f[F_QZZ][i]);
        qt[3] = fmaf(sigma, (bxv[1] + byv[0]) * 0.5f, S.f[F_QXY][i]);
        qt[4] = fmaf(sigma, (bxv[2] + bzv[0]) * 0.5f, S.f[F_QXZ][i]);
        qt[5] = fmaf(sigma, (byv[2] + bzv[1]) * 0.5f, S.f[F_QYZ][i]);
        float m2 = qt[0] * qt[0];
        m2 = fmaf(qt[1], qt[1], m2);
        m2 = fmaf(2.0f * qt[3], qt[3], m2);
        m2 = fmaf(qt[2], qt[2], m2);
        m2 = fmaf(2.0f * qt[4], qt[4], m2);
        m2 = fmaf(2.0f * qt[5], qt[5], m2);
        const float nq = sqrtf(m2);
        if (nq / alpha0 > 1.0f) {
            float dot = 0.0f;
#pragma unroll
            for (int k = 0; k < 6; ++k) dot += Q[k] * qt[k];
            acc[0] += double(dot) / double(nq);
            const float k1 = alpha0 / nq, k2 = dot / (nq * nq);
#pragma unroll
            for (int k = 0; k < 6; ++k) Q[k] = k1 * (Q[k] - k2 * (k < 3 ? 1.0f : 2.0f) * qt[k]);
        }
#pragma unroll
        for (int a = 0; a < 3; ++a) A.f[F_PX + a][i] = P[a];
#pragma unroll
        for (int k = 0; k < 6; ++k) A.f[F_QXX + k][i] = Q[k];
    }
    block_accumulate<2, 0>(acc, slots);   // slots 0, 1: alpha0, alpha1
}

// Phase C at voxel (i, c3): the u_bar and v_bar adjoints from P~, Q~ (the p
// and q adjoints after phase B).
__device__ __forceinline__ void phase_c_voxel(const Fields17& A, uint32_t i, const int (&c3)[3],
                                              const uint32_t (&st)[3], const int (&ext)[3],
                                              float sigma, float& ub_out, float (&vb)[3]) {
    // u_bar adjoint = sigma (D+)^T P~
    float ub = 0.0f;
#pragma unroll
    for (int a = 0; a < 3; ++a) ub += dplus_t(A.f[F_PX + a], i, st[a], c3[a], ext[a]);
    ub_out = sigma * ub;
    // v_bar adjoint = sigma eps_B^T Q~ - sigma P~
    //   x: B_x^T Q_xx + 1/2 B_y^T Q_xy + 1/2 B_z^T Q_xz, etc.
    const float* Qxx = A.f[F_QXX]; const float* Qyy = A.f[F_QYY]; const float* Qzz = A.f[F_QZZ];
    const float* Qxy = A.f[F_QXY]; const float* Qxz = A.f[F_QXZ]; const float* Qyz = A.f[F_QYZ];
    const float ex = bm_t(Qxx, i, st[0], c3[0], ext[0]) +
                     0.5f * bm_t(Qxy, i, st[1], c3[1], ext[1]) +
                     0.5f * bm_t(Qxz, i, st[2], c3[2], ext[2]);
    const float ey = bm_t(Qyy, i, st[1], c3[1], ext[1]) +
                     0.5f * bm_t(Qxy, i, st[0], c3[0], ext[0]) +
                     0.5f * bm_t(Qyz, i, st[2], c3[2], ext[2]);
    const float ez = bm_t(Qzz, i, st[2], c3[2], ext[2]) +
                     0.5f * bm_t(Qxz, i, st[0], c3[0], ext[0]) +
                     0.5f * bm_t(Qyz, i, st[1], c3[1], ext[1]);
    vb[0] = sigma * ex - sigma * A.f[F_PX][i];
    vb[1] = sigma * ey - sigma * A.f[F_PY][i];
    vb[2] = sigma * ez - sigma * A.f[F_PZ][i];
}

__global__ __launch_bounds__(AB) void adjoint_phase_c(Fields17 A, Geometry g, float sigma) {
    const uint32_t st[3] = {1u, uint32_t(g.w), uint32_t(g.w) * uint32_t(g.h)};
    const int ext[3] = {g.w, g.h, g.d};
    const RowVoxel rv(g);
    if (rv.in) {
        const int c3[3] = {rv.x, rv.y, rv.z};
        float ub, vb[3];
        phase_c_voxel(A, rv.i, c3, st, ext, sigma, ub, vb);
        A.f[F_UBAR][rv.i] = ub;
        A.f[F_VBX][rv.i] = vb[0];
        A.f[F_VBY][rv.i] = vb[1];
        A.f[F_VBZ][rv.i] = vb[2];
    }
}

// Phase C of iteration k and phase A of iteration k - 1 (u_next: u of x_k).
__global__ __launch_bounds__(AB) void adjoint_phase_ca(Fields17 A, const float* __restrict__ f,
                                                       const float* __restrict__ u_next,
                                                       float* __restrict__ ast,
                                                       float* __restrict__ wst,
                                                       float* __restrict__ pout,
                                                       float* __restrict__ grad_f, Geometry g,
                                                       float tau, float sigma, float lambda,
                                                       float theta, double* __restrict__ slots) {
    const uint32_t n = uint32_t(g.n());
    const uint32_t st[3] = {1u, uint32_t(g.w), uint32_t(g.w) * uint32_t(g.h)};
    const int ext[3] = {g.w, g.h, g.d};
    double acc[1] = {0.0};
    const RowVoxel rv(g);
    if (rv.in) {
        const int c3[3] = {rv.x, rv.y, rv.z};
        float ub, vb[3];
        phase_c_voxel(A, rv.i, c3, st, ext, sigma, ub, vb);
        phase_a_voxel(A, rv.i, n, ub, vb, f, u_next, ast, wst, pout, grad_f, tau, tau * lambda,
                      theta, acc[0]);
    }
    block_accumulate<1, 2>(acc, slots);   // slot 2: lambda
}

// out[k] = sum over blocks of slots[b][k], fixed order (one wave per k < 3)
__global__ void adjoint_reduce(const double* __restrict__ slots, int nblocks,
                               double* __restrict__ out) {
    const int k = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (k >= 3) return;
    double s = 0.0;
    for (int b = lane; b < nblocks; b += 64) s += slots[size_t(b) * 3 + k];
    s = wsum(s);
    if (lane == 0) out[k] = s;
}

__global__ void adjoint_finish_f(const float* __restrict__ au, const float* __restrict__ aub,
                                 float* __restrict__ grad_f, size_t n) {
    for (size_t i = size_t(blockIdx.x) * AB + threadIdx.x; i < n; i += size_t(gridDim.x) * AB)
        grad_f[i] += au[i] + aub[i];
}

// threads per block: the row length rounded up to whole waves, at most AB
inline int row_threads(const Geometry& g) { return std::min(AB, (g.w + 63) / 64 * 64); }
inline int row_blocks(const Geometry& g) {
    const int t = row_threads(g);
    return ((g.w + t - 1) / t) * g.h * g.d;
}

inline int blocks_for(size_t n) {
    size_t b = (n + AB - 1) / AB;
    if (b > 2048) b = 2048;
    return int(b ? b : 1);
}

Fields17 to_fields(void* const* p) {
    Fields17 r;
    for (int k = 0; k < N_FIELDS; ++k) r.f[k] = static_cast<float*>(p[k]);
    return r;
}

}  // namespace

size_t adjoint_slots_size(Geometry g) { return size_t(row_blocks(g)) * 3; }

// slots: adjoint_slots_size(g) doubles, zeroed before the first step; each
// phase adds every block's (alpha0, alpha1, lambda) terms to its own slots.
hipError_t launch_adjoint_a(hipStream_t s, const AdjointArgs& a, const float* u_next) {
    const Fields17 A = to_fields(a.adj);
    hipLaunchKernelGGL(adjoint_phase_a, dim3(row_blocks(a.g)), dim3(row_threads(a.g)), 0, s, A,
                       a.f, u_next, a.ast, a.wst, a.pbuf, a.grad_f, a.g, a.tau, a.lambda, a.theta,
                       a.slots);
    return hipGetLastError();
}

hipError_t launch_adjoint_b(hipStream_t s, const AdjointArgs& a, void* const* xk) {
    const Fields17 A = to_fields(a.adj), S = to_fields(xk);
    hipLaunchKernelGGL(adjoint_phase_b, dim3(row_blocks(a.g)), dim3(row_threads(a.g)), 0, s, A, S,
                       a.pbuf, a.ast, a.wst, a.g, a.tau, a.sigma, a.alpha0, a.alpha1, a.slots);
    return hipGetLastError();
}

hipError_t launch_adjoint_ca(hipStream_t s, const AdjointArgs& a, const float* u_next) {
    const Fields17 A = to_fields(a.adj);
    hipLaunchKernelGGL(adjoint_phase_ca, dim3(row_blocks(a.g)), dim3(row_threads(a.g)), 0, s, A,
                       a.f, u_next, a.ast, a.wst, a.pbuf, a.grad_f, a.g, a.tau, a.sigma, a.lambda,
                       a.theta, a.slots);
    return hipGetLastError();
}

hipError_t launch_adjoint_c(hipStream_t s, const AdjointArgs& a) {
    const Fields17 A = to_fields(a.adj);
    hipLaunchKernelGGL(adjoint_phase_c, dim3(row_blocks(a.g)), dim3(row_threads(a.g)), 0, s, A,
                       a.g, a.sigma);
    return hipGetLastError();
}

hipError_t launch_adjoint_reduce(hipStream_t s, const double* slots, Geometry g, double* out3) {
    hipLaunchKernelGGL(adjoint_reduce, dim3(1), dim3(192), 0, s, slots, row_blocks(g), out3);
    return hipGetLastError();
}

hipError_t launch_adjoint_finish(hipStream_t s, void* const* adj, float* grad_f, Geometry g) {
    hipLaunchKernelGGL(adjoint_finish_f, dim3(blocks_for(g.n())), dim3(AB), 0, s,
                       static_cast<const float*>(adj[F_U]), static_cast<const float*>(adj[F_UBAR]),
                       grad_f, g.n());
    return hipGetLastError();
}

}  // namespace amb
