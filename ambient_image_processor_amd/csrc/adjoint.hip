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

// VX consecutive voxels of one row per thread (VX = 2 needs an even width):
// lane l of row piece b = xcd_contiguous(blockIdx.x) covers x = VX l, ...,
// VX l + VX - 1 of row r = b / nbx (y = r % H, z = r / H).  Own values are
// loaded VX at a time; the x neighbours across the pair come from the other
// element or one scalar load.
template <int VX>
struct RowVox {
    int x, y, z;   // x of the first voxel
    uint32_t i;    // its index
    bool in;
    __device__ __forceinline__ RowVox(const Geometry& g) {
        const uint32_t lanes = uint32_t(g.w) / VX;
        const uint32_t nbx = (lanes + blockDim.x - 1) / blockDim.x;
        const uint32_t b = xcd_contiguous(blockIdx.x, gridDim.x);
        const uint32_t r = b / nbx;
        const uint32_t lane = (b - r * nbx) * blockDim.x + threadIdx.x;
        x = int(lane) * VX;
        y = int(r % uint32_t(g.h));
        z = int(r / uint32_t(g.h));
        in = lane < lanes;
        i = (uint32_t(z) * uint32_t(g.h) + uint32_t(y)) * uint32_t(g.w) + uint32_t(x);
    }
};

template <int VX>
__device__ __forceinline__ void ldv(const float* __restrict__ p, uint32_t i, float (&v)[VX]) {
    if constexpr (VX == 2) {
        const float2 t = *reinterpret_cast<const float2*>(p + i);
        v[0] = t.x; v[1] = t.y;
    } else {
        v[0] = p[i];
    }
}
template <int VX>
__device__ __forceinline__ void stv(float* __restrict__ p, uint32_t i, const float (&v)[VX]) {
    if constexpr (VX == 2) *reinterpret_cast<float2*>(p + i) = make_float2(v[0], v[1]);
    else p[i] = v[0];
}

// Neighbour values of the VX voxels along axis a (0 = x): at +1 (nxt) or -1
// (prv); 0 where the neighbour lies outside (its flag is false there and the
// value unused).  own: the field's values at the VX voxels.
template <int VX>
__device__ __forceinline__ void nxt(const float* __restrict__ p, const RowVox<VX>& v, int a,
                                    const float (&own)[VX], const Geometry& g, float (&out)[VX]) {
    if (a == 0) {
#pragma unroll
        for (int e = 0; e + 1 < VX; ++e) out[e] = own[e + 1];
        out[VX - 1] = v.x + VX < g.w ? p[v.i + VX] : 0.0f;
    } else if (a == 1) {
        if (v.y < g.h - 1) ldv<VX>(p, v.i + uint32_t(g.w), out);
        else for (int e = 0; e < VX; ++e) out[e] = 0.0f;
    } else {
        if (v.z < g.d - 1) ldv<VX>(p, v.i + uint32_t(g.w) * uint32_t(g.h), out);
        else for (int e = 0; e < VX; ++e) out[e] = 0.0f;
    }
}
template <int VX>
__device__ __forceinline__ void prv(const float* __restrict__ p, const RowVox<VX>& v, int a,
                                    const float (&own)[VX], const Geometry& g, float (&out)[VX]) {
    if (a == 0) {
        out[0] = v.x > 0 ? p[v.i - 1] : 0.0f;
#pragma unroll
        for (int e = 1; e < VX; ++e) out[e] = own[e - 1];
    } else if (a == 1) {
        if (v.y > 0) ldv<VX>(p, v.i - uint32_t(g.w), out);
        else for (int e = 0; e < VX; ++e) out[e] = 0.0f;
    } else {
        if (v.z > 0) ldv<VX>(p, v.i - uint32_t(g.w) * uint32_t(g.h), out);
        else for (int e = 0; e < VX; ++e) out[e] = 0.0f;
    }
}

// (D+)^T b [j] = b[j-1] [j >= 1] - b[j] [j < n-1], from own value o and
// previous value m (has_m: j >= 1, has_p: j < n-1)
__device__ __forceinline__ float dplus_tv(float o, float m, bool has_m, bool has_p) {
    float r = 0.0f;
    if (has_m) r += m;
    if (has_p) r -= o;
    return r;
}
// B^T b [j] = b[j+1] [j < n-1] - b[j], from own value o and next value q
__device__ __forceinline__ float bm_tv(float o, float q, bool has_p) {
    float r = -o;
    if (has_p) r += q;
    return r;
}

// Per-voxel flags: hp[a]: the voxel has a +a neighbour; hm[a]: a -a neighbour.
template <int VX>
__device__ __forceinline__ void flags(const RowVox<VX>& v, int e, const Geometry& g, bool (&hp)[3],
                                      bool (&hm)[3]) {
    const int c[3] = {v.x + e, v.y, v.z};
    const int ext[3] = {g.w, g.h, g.d};
#pragma unroll
    for (int a = 0; a < 3; ++a) { hp[a] = c[a] < ext[a] - 1; hm[a] = c[a] >= 1; }
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
    const RowVox<1> rv(g);
    if (rv.in) {
        const uint32_t i = rv.i;
        const float vb[3] = {A.f[F_VBX][i], A.f[F_VBY][i], A.f[F_VBZ][i]};
        phase_a_voxel(A, i, n, A.f[F_UBAR][i], vb, f, u_next, ast, wst, pout, grad_f, tau,
                      tau * lambda, theta, acc[0]);
    }
    block_accumulate<1, 2>(acc, slots);   // slot 2: lambda
}

template <int VX>
__global__ __launch_bounds__(AB) void adjoint_phase_b(Fields17 A, Fields17 S,
                                                      const float* __restrict__ pin,
                                                      const float* __restrict__ ast,
                                                      const float* __restrict__ wst, Geometry g,
                                                      float tau, float sigma, float alpha0,
                                                      float alpha1, double* __restrict__ slots) {
    const uint32_t n = uint32_t(g.n());
    double acc[2] = {0.0, 0.0};   // alpha0, alpha1
    const RowVox<VX> rv(g);
    if (rv.in) {
        const uint32_t i = rv.i;
        // own values and the neighbours the transposes and the recomputed p~, q~ read
        float pi[3][VX], as[VX], asn[3][VX], w[3][VX], wp[3][3][VX], qa[6][VX];
        float ub[VX], ubn[3][VX], vb[3][VX], vbp[3][3][VX], sp[3][VX], sq[6][VX];
#pragma unroll
        for (int a = 0; a < 3; ++a) ldv<VX>(pin, a * n + i, pi[a]);
        ldv<VX>(ast, i, as);
        ldv<VX>(S.f[F_UBAR], i, ub);
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            nxt<VX>(ast, rv, a, as, g, asn[a]);
            nxt<VX>(S.f[F_UBAR], rv, a, ub, g, ubn[a]);
        }
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            ldv<VX>(wst, k * n + i, w[k]);
            ldv<VX>(S.f[F_VBX + k], i, vb[k]);
            ldv<VX>(S.f[F_PX + k], i, sp[k]);
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                prv<VX>(wst + k * n, rv, a, w[k], g, wp[k][a]);
                prv<VX>(S.f[F_VBX + k], rv, a, vb[k], g, vbp[k][a]);
            }
        }
#pragma unroll
        for (int k = 0; k < 6; ++k) {
            ldv<VX>(A.f[F_QXX + k], i, qa[k]);
            ldv<VX>(S.f[F_QXX + k], i, sq[k]);
        }
        float Po[3][VX], Qo[6][VX];
#pragma unroll
        for (int e = 0; e < VX; ++e) {
            bool hp[3], hm[3];
            flags<VX>(rv, e, g, hp, hm);
            // P' = A.p - tau B^T A*   (component a uses B_a)
            float P[3];
#pragma unroll
            for (int a = 0; a < 3; ++a) P[a] = pi[a][e] - tau * bm_tv(as[e], asn[a][e], hp[a]);
            // Q' = A.q + div2^T W*
            auto dw = [&](int k, int a) { return dplus_tv(w[k][e], wp[k][a][e], hm[a], hp[a]); };
            float Q[6];
            Q[0] = qa[0][e] + dw(0, 0);
            Q[1] = qa[1][e] + dw(1, 1);
            Q[2] = qa[2][e] + dw(2, 2);
            Q[3] = qa[3][e] + (dw(0, 1) + dw(1, 0));
            Q[4] = qa[4][e] + (dw(0, 2) + dw(2, 0));
            Q[5] = qa[5][e] + (dw(1, 2) + dw(2, 1));

            // p~ from x_k, forward arithmetic (tgv_sweeps.hip dual sweep)
            const float ubc = ub[e];
            float pt[3];
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                const float d = hp[a] ? ubn[a][e] - ubc : 0.0f;
                pt[a] = fmaf(sigma, d - vb[a][e], sp[a][e]);
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
                const float vc = vb[c][e];
                bxv[c] = hm[0] ? -vc + vbp[c][0][e] : -vc;
                byv[c] = hm[1] ? -vc + vbp[c][1][e] : -vc;
                bzv[c] = hm[2] ? -vc + vbp[c][2][e] : -vc;
            }
            float qt[6];
            qt[0] = fmaf(sigma, bxv[0], sq[0][e]);
            qt[1] = fmaf(sigma, byv[1], sq[1][e]);
            qt[2] = fmaf(sigma, bzv[2], sq[2][e]);
            qt[3] = fmaf(sigma, (bxv[1] + byv[0]) * 0.5f, sq[3][e]);
            qt[4] = fmaf(sigma, (bxv[2] + bzv[0]) * 0.5f, sq[4][e]);
            qt[5] = fmaf(sigma, (byv[2] + bzv[1]) * 0.5f, sq[5][e]);
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
            for (int a = 0; a < 3; ++a) Po[a][e] = P[a];
#pragma unroll
            for (int k = 0; k < 6; ++k) Qo[k][e] = Q[k];
        }
#pragma unroll
        for (int a = 0; a < 3; ++a) stv<VX>(A.f[F_PX + a], i, Po[a]);
#pragma unroll
        for (int k = 0; k < 6; ++k) stv<VX>(A.f[F_QXX + k], i, Qo[k]);
    }
    block_accumulate<2, 0>(acc, slots);   // slots 0, 1: alpha0, alpha1
}

// Phase C at the VX voxels: the u_bar and v_bar adjoints from P~, Q~ (the p
// and q adjoints after phase B); P~ at the voxels also returned (pt).
template <int VX>
__device__ __forceinline__ void phase_c_vox(const Fields17& A, const RowVox<VX>& rv,
                                            const Geometry& g, float sigma, float (&ubo)[VX],
                                            float (&vbo)[3][VX], float (&pt)[3][VX]) {
    const uint32_t i = rv.i;
    float pm[3][VX], q[6][VX], qn[6][3][VX];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        ldv<VX>(A.f[F_PX + a], i, pt[a]);
        prv<VX>(A.f[F_PX + a], rv, a, pt[a], g, pm[a]);
    }
    // Q~ next values used: xx +x; yy +y; zz +z; xy +x, +y; xz +x, +z; yz +y, +z
    const int axes[6][2] = {{0, 0}, {1, 1}, {2, 2}, {0, 1}, {0, 2}, {1, 2}};
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        ldv<VX>(A.f[F_QXX + k], i, q[k]);
        nxt<VX>(A.f[F_QXX + k], rv, axes[k][0], q[k], g, qn[k][axes[k][0]]);
        if (axes[k][1] != axes[k][0]) nxt<VX>(A.f[F_QXX + k], rv, axes[k][1], q[k], g, qn[k][axes[k][1]]);
    }
#pragma unroll
    for (int e = 0; e < VX; ++e) {
        bool hp[3], hm[3];
        flags<VX>(rv, e, g, hp, hm);
        // u_bar adjoint = sigma (D+)^T P~
        float ub = 0.0f;
#pragma unroll
        for (int a = 0; a < 3; ++a) ub += dplus_tv(pt[a][e], pm[a][e], hm[a], hp[a]);
        ubo[e] = sigma * ub;
        // v_bar adjoint = sigma eps_B^T Q~ - sigma P~
        //   x: B_x^T Q_xx + 1/2 B_y^T Q_xy + 1/2 B_z^T Q_xz, etc.
        auto bq = [&](int k, int a) { return bm_tv(q[k][e], qn[k][a][e], hp[a]); };
        const float ex = bq(0, 0) + 0.5f * bq(3, 1) + 0.5f * bq(4, 2);
        const float ey = bq(1, 1) + 0.5f * bq(3, 0) + 0.5f * bq(5, 2);
        const float ez = bq(2, 2) + 0.5f * bq(4, 0) + 0.5f * bq(5, 1);
        vbo[0][e] = sigma * ex - sigma * pt[0][e];
        vbo[1][e] = sigma * ey - sigma * pt[1][e];
        vbo[2][e] = sigma * ez - sigma * pt[2][e];
    }
}

template <int VX>
__global__ __launch_bounds__(AB) void adjoint_phase_c(Fields17 A, Geometry g, float sigma) {
    const RowVox<VX> rv(g);
    if (rv.in) {
        float ub[VX], vb[3][VX], pt[3][VX];
        phase_c_vox<VX>(A, rv, g, sigma, ub, vb, pt);
        stv<VX>(A.f[F_UBAR], rv.i, ub);
#pragma unroll
        for (int c = 0; c < 3; ++c) stv<VX>(A.f[F_VBX + c], rv.i, vb[c]);
    }
}

// Phase C of iteration k and phase A of iteration k - 1 (u_next: u of x_k).
template <int VX>
__global__ __launch_bounds__(AB) void adjoint_phase_ca(Fields17 A, const float* __restrict__ f,
                                                       const float* __restrict__ u_next,
                                                       float* __restrict__ ast,
                                                       float* __restrict__ wst,
                                                       float* __restrict__ pout,
                                                       float* __restrict__ grad_f, Geometry g,
                                                       float tau, float sigma, float lambda,
                                                       float theta, double* __restrict__ slots) {
    const uint32_t n = uint32_t(g.n());
    double acc[1] = {0.0};
    const RowVox<VX> rv(g);
    if (rv.in) {
        const uint32_t i = rv.i;
        float ub[VX], vb[3][VX], pt[3][VX];
        phase_c_vox<VX>(A, rv, g, sigma, ub, vb, pt);
        // phase A of iteration k - 1 on the same voxels (phase_a_voxel's arithmetic)
        const float tl = tau * lambda;
        float au[VX], fv[VX], un_[VX], gf[VX], av[3][VX];
        ldv<VX>(A.f[F_U], i, au);
        ldv<VX>(f, i, fv);
        ldv<VX>(u_next, i, un_);
        ldv<VX>(grad_f, i, gf);
#pragma unroll
        for (int c = 0; c < 3; ++c) ldv<VX>(A.f[F_VX + c], i, av[c]);
        float auo[VX], aso[VX], gfo[VX], avo[3][VX], wso[3][VX], po[3][VX];
#pragma unroll
        for (int e = 0; e < VX; ++e) {
            const float un = au[e] + (1.0f + theta) * ub[e];
            const float as = un / (1.0f + tl);
            auo[e] = -theta * ub[e] + as;
            aso[e] = as;
            gfo[e] = gf[e] + tl / (1.0f + tl) * un;
            acc[0] += double(un) * double(tau) * (double(fv[e]) - double(un_[e])) / (1.0 + double(tl));
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const float vn = av[c][e] + (1.0f + theta) * vb[c][e];
                avo[c][e] = -theta * vb[c][e] + vn;
                wso[c][e] = -tau * vn;
                po[c][e] = pt[c][e] + tau * vn;
            }
        }
        stv<VX>(A.f[F_U], i, auo);
        stv<VX>(ast, i, aso);
        stv<VX>(grad_f, i, gfo);
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            stv<VX>(A.f[F_VX + c], i, avo[c]);
            stv<VX>(wst + c * n, i, wso[c]);
            stv<VX>(pout + c * n, i, po[c]);
        }
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

// threads per block: a row's VX-voxel lanes rounded up to whole waves, at most
// AB; blocks: the row pieces of all rows (VX = 1 gives the most)
inline int row_threads(const Geometry& g, int vx = 1) {
    return std::min(AB, (g.w / vx + 63) / 64 * 64);
}
inline int row_blocks(const Geometry& g, int vx = 1) {
    const int t = row_threads(g, vx), lanes = g.w / vx;
    return ((lanes + t - 1) / t) * g.h * g.d;
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

int adjoint_vx(Geometry g, int want) {
    if (want == 1) return 1;
    return g.w % 2 == 0 ? 2 : 1;
}

hipError_t launch_adjoint_b(hipStream_t s, const AdjointArgs& a, void* const* xk) {
    const Fields17 A = to_fields(a.adj), S = to_fields(xk);
    const dim3 grid(row_blocks(a.g, a.vx)), block(row_threads(a.g, a.vx));
    if (a.vx == 2)
        hipLaunchKernelGGL(adjoint_phase_b<2>, grid, block, 0, s, A, S, a.pbuf, a.ast, a.wst, a.g,
                           a.tau, a.sigma, a.alpha0, a.alpha1, a.slots);
    else
        hipLaunchKernelGGL(adjoint_phase_b<1>, grid, block, 0, s, A, S, a.pbuf, a.ast, a.wst, a.g,
                           a.tau, a.sigma, a.alpha0, a.alpha1, a.slots);
    return hipGetLastError();
}

hipError_t launch_adjoint_ca(hipStream_t s, const AdjointArgs& a, const float* u_next) {
    const Fields17 A = to_fields(a.adj);
    const dim3 grid(row_blocks(a.g, a.vx)), block(row_threads(a.g, a.vx));
    if (a.vx == 2)
        hipLaunchKernelGGL(adjoint_phase_ca<2>, grid, block, 0, s, A, a.f, u_next, a.ast, a.wst,
                           a.pbuf, a.grad_f, a.g, a.tau, a.sigma, a.lambda, a.theta, a.slots);
    else
        hipLaunchKernelGGL(adjoint_phase_ca<1>, grid, block, 0, s, A, a.f, u_next, a.ast, a.wst,
                           a.pbuf, a.grad_f, a.g, a.tau, a.sigma, a.lambda, a.theta, a.slots);
    return hipGetLastError();
}

hipError_t launch_adjoint_c(hipStream_t s, const AdjointArgs& a) {
    const Fields17 A = to_fields(a.adj);
    const dim3 grid(row_blocks(a.g, a.vx)), block(row_threads(a.g, a.vx));
    if (a.vx == 2)
        hipLaunchKernelGGL(adjoint_phase_c<2>, grid, block, 0, s, A, a.g, a.sigma);
    else
        hipLaunchKernelGGL(adjoint_phase_c<1>, grid, block, 0, s, A, a.g, a.sigma);
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
