// TGV-k (k >= 3) deshade iteration on gfx950 (reference
// tgvk_deshade/tgvk_l1_deshade.cu:152-351 with tgv3/tgv3_common.cu and
// tgvk/tgvk_common.cu).
//
// The reference runs, per iteration: the p-update, the u-update, the q-update
// (shifted by -w_bar_0), div2(q), the v-update, then for every level i the
// dual update of r_i (from G(w_bar_i) - w_bar_{i+1}) followed by the primal
// update of w_i (from H(r_i) - q~, q~ = q or r_{i-1}).  Every dual update reads
// only primal-bar fields of the previous iteration (u_bar, v_bar, w_bar_i,
// w_bar_{i+1} are written later), and every primal update reads only dual
// fields of this iteration and its own old value.  So one iteration is two
// launches whatever k is:
//   dual   - p, q and every r_i, in place (each voxel reads only its own p, q, r);
//   primal - u, v and every w_i with their over-relaxations, in place.
// Each thread owns one (x, y) column and marches along z over `zchunk`
// planes; neighbours are read from global memory (the x/y neighbours were just
// read by adjacent lanes and waves, the z neighbour by the same thread one
// plane earlier, so they come from L1/L2).  Level fields are reached through a
// device table of 18 pointers per level (w[6], w_bar[6], r[6], components
// x, y, z, xy, xz, yz), so the order is a run-time value.
//
// Per-voxel arithmetic keeps the reference's FP32 operation order (fmaf where
// the reference writes fmaf or nvcc contracts a multiply-add; correctly
// rounded division and sqrt; built with -ffp-contract=off).
#include "amb_internal.hpp"

namespace amb {
namespace {

constexpr int KTX = 64, KTY = 4;

// tensor component c -> its first and second axis (-1: diagonal component)
__device__ __forceinline__ int comp_axis_a(int c) { return c < 3 ? c : (c == 5 ? 1 : 0); }
__device__ __forceinline__ int comp_axis_b(int c) { return c < 3 ? -1 : (c == 3 ? 1 : 2); }

struct KCoord {
    int x, y, z0, z1;
    size_t plane, col;
    bool in;
    __device__ __forceinline__ KCoord(const Geometry& g, int zchunk) {
        x = int(blockIdx.x) * KTX + int(threadIdx.x);
        y = int(blockIdx.y) * KTY + int(threadIdx.y);
        z0 = int(blockIdx.z) * zchunk;
        z1 = min(z0 + zchunk, g.d);
        in = x < g.w && y < g.h;
        plane = size_t(g.w) * size_t(g.h);
        col = size_t(y) * size_t(g.w) + size_t(x);
    }
};

// Reference "backward difference" (tgv_cuda/tgv_common.cu:133-197): -a[i] + a[i-1], -a[0] at 0.
__device__ __forceinline__ float kbwd(const float* a, float center, size_t i, size_t st, bool has) {
    return has ? -center + a[i - st] : -center;
}
// Reference forward difference (tgv_common.cu:66-130): a[i+1] - a[i], 0 at the last index.
__device__ __forceinline__ float kfwd(const float* a, float center, size_t i, size_t st, bool has) {
    return has ? a[i + st] - center : 0.0f;
}

__global__ __launch_bounds__(KTX * KTY) void tgvk_dual_kernel(
    float* const* __restrict__ F, float* const* __restrict__ LV,
    const float* __restrict__ alpha, Geometry g, int levels, float sigma, int planar, int zchunk) {
    const KCoord c(g, zchunk);
    if (!c.in) return;
    const int W = g.w, H = g.h, D = g.d;
    const size_t st[3] = {1, size_t(W), c.plane};
    const float a_p = alpha[0], a_q = alpha[1];
    const float* ub = F[F_UBAR];
    const float* vb[3] = {F[F_VBX], F[F_VBY], F[F_VBZ]};
    float* pp[3] = {F[F_PX], F[F_PY], F[F_PZ]};
    float* qq[6] = {F[F_QXX], F[F_QYY], F[F_QZZ], F[F_QXY], F[F_QXZ], F[F_QYZ]};

    for (int z = c.z0; z < c.z1; ++z) {
        const size_t i = size_t(z) * c.plane + c.col;
        const bool nxt[3] = {c.x < W - 1, c.y < H - 1, z < D - 1};
        const bool prv[3] = {c.x > 0, c.y > 0, z > 0};

        // p update + projection onto |p| <= alpha[0] (tgv2_common.cu:96-131)
        const float uc = ub[i];
        float vbc[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) vbc[k] = vb[k][i];
        float a3[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const float d = kfwd(ub, uc, i, st[k], nxt[k]);
            a3[k] = fmaf(sigma, d - vbc[k], pp[k][i]);
        }
        float n2 = a3[0] * a3[0];
        n2 = fmaf(a3[1], a3[1], n2);
        n2 = fmaf(a3[2], a3[2], n2);
        const float tp = sqrtf(n2) / a_p;
        const bool sp = tp > 1.0f;
#pragma unroll
        for (int k = 0; k < 3; ++k) pp[k][i] = sp ? a3[k] / tp : a3[k];

        // q update with eps_B(v_bar) - w_bar_0, projection with alpha[1]
        // (tgv2_common.cu:174-219, tgv3/tgv3_common.cu:28-86)
        float bxv[3], byv[3], bzv[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            bxv[k] = kbwd(vb[k], vbc[k], i, st[0], prv[0]);
            byv[k] = kbwd(vb[k], vbc[k], i, st[1], prv[1]);
            bzv[k] = planar ? 0.0f : kbwd(vb[k], vbc[k], i, st[2], prv[2]);
        }
        float e[6];
        e[0] = bxv[0];
        e[1] = byv[1];
        e[2] = bzv[2];
        e[3] = (bxv[1] + byv[0]) * 0.5f;
        e[4] = (bxv[2] + bzv[0]) * 0.5f;
        e[5] = (byv[2] + bzv[1]) * 0.5f;
        float wc[6];
#pragma unroll
        for (int k = 0; k < 6; ++k) wc[k] = LV[6 + k][i];
        float b6[6];
#pragma unroll
        for (int k = 0; k < 6; ++k) b6[k] = fmaf(sigma, e[k] - wc[k], qq[k][i]);
        float m2 = b6[0] * b6[0];
        m2 = fmaf(b6[1], b6[1], m2);
        m2 = fmaf(2.0f * b6[3], b6[3], m2);
        m2 = fmaf(b6[2], b6[2], m2);
        m2 = fmaf(2.0f * b6[4], b6[4], m2);
        m2 = fmaf(2.0f * b6[5], b6[5], m2);
        const float tq = sqrtf(m2) / a_q;
        const bool sq = tq > 1.0f;
#pragma unroll
        for (int k = 0; k < 6; ++k) qq[k][i] = sq ? b6[k] / tq : b6[k];

        // levels: r_i update + projection with alpha[i + 2]
        // (G: tgv3_common.cu:214-262 forward, tgvk_common.cu:89-135 backward;
        //  update: tgvk_common.cu:28-86, last level tgv3_common.cu:311-360)
        for (int li = 0; li < levels; ++li) {
            float* const* lv = LV + size_t(li) * 18;
            const bool forward = (li & 1) != 0;
            const bool last = li == levels - 1;
            float g6[6];
#pragma unroll
            for (int k = 0; k < 6; ++k) {
                const float* wb = lv[6 + k];
                const int a = comp_axis_a(k), b = comp_axis_b(k);
                float d = forward ? kfwd(wb, wc[k], i, st[a], nxt[a])
                                  : kbwd(wb, wc[k], i, st[a], prv[a]);
                if (b >= 0) {
                    const float d2 = forward ? kfwd(wb, wc[k], i, st[b], nxt[b])
                                             : kbwd(wb, wc[k], i, st[b], prv[b]);
                    d = (d + d2) * 0.5f;
                }
                g6[k] = d;
            }
            float rc[6];
            if (!last) {
                float* const* nx = lv + 18;
#pragma unroll
                for (int k = 0; k < 6; ++k) {
                    const float wn = nx[6 + k][i];
                    rc[k] = fmaf(sigma, g6[k] - wn, lv[12 + k][i]);
                    wc[k] = wn;   // the next level's centre
                }
            } else {
#pragma unroll
                for (int k = 0; k < 6; ++k) rc[k] = fmaf(sigma, g6[k], lv[12 + k][i]);
            }
            float r2 = rc[0] * rc[0];
            r2 = fmaf(rc[1], rc[1], r2);
            r2 = fmaf(2.0f * rc[3], rc[3], r2);
            r2 = fmaf(rc[2], rc[2], r2);
            r2 = fmaf(2.0f * rc[4], rc[4], r2);
            r2 = fmaf(2.0f * rc[5], rc[5], r2);
            const float tr = sqrtf(r2) / alpha[li + 2];
            const bool sr = tr > 1.0f;
#pragma unroll
            for (int k = 0; k < 6; ++k) lv[12 + k][i] = sr ? rc[k] / tr : rc[k];
        }
    }
}

__global__ __launch_bounds__(KTX * KTY) void tgvk_primal_kernel(
    float* const* __restrict__ F, float* const* __restrict__ LV, const float* __restrict__ f,
    Geometry g, int levels, float tau, float lambda, float theta, int zchunk) {
    const KCoord c(g, zchunk);
    if (!c.in) return;
    const int W = g.w, H = g.h, D = g.d;
    const size_t st[3] = {1, size_t(W), c.plane};
    const float tl = tau * lambda;
    float* uu = F[F_U];
    float* ub = F[F_UBAR];
    float* vv[3] = {F[F_VX], F[F_VY], F[F_VZ]};
    float* vb[3] = {F[F_VBX], F[F_VBY], F[F_VBZ]};
    const float* pp[3] = {F[F_PX], F[F_PY], F[F_PZ]};
    const float* qq[6] = {F[F_QXX], F[F_QYY], F[F_QZZ], F[F_QXY], F[F_QXZ], F[F_QYZ]};

    for (int z = c.z0; z < c.z1; ++z) {
        const size_t i = size_t(z) * c.plane + c.col;
        const bool nxt[3] = {c.x < W - 1, c.y < H - 1, z < D - 1};
        const bool prv[3] = {c.x > 0, c.y > 0, z > 0};

        // u update, L1 prox, over-relaxation (tgv2_l1.cu:27-60)
        float pc[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) pc[k] = pp[k][i];
        const float bx = kbwd(pp[0], pc[0], i, st[0], prv[0]);
        const float by = kbwd(pp[1], pc[1], i, st[1], prv[1]);
        const float bz = kbwd(pp[2], pc[2], i, st[2], prv[2]);
        const float div = (bx + by) + bz;
        const float uo = uu[i], fc = f[i];
        float un = fmaf(-tau, div, uo);
        const float umf = un - fc;
        if (umf > tl) un = un - tl;
        else if (umf < -tl) un = un + tl;
        else un = fc;
        uu[i] = un;
        ub[i] = fmaf(theta, un - uo, un);

        // div2(q) with forward differences, v update (tgv2_common.cu:279-354)
        float qc[6];
#pragma unroll
        for (int k = 0; k < 6; ++k) qc[k] = qq[k][i];
        const float dxx = kfwd(qq[0], qc[0], i, st[0], nxt[0]);
        const float dyxy = kfwd(qq[3], qc[3], i, st[1], nxt[1]);
        const float dzxz = kfwd(qq[4], qc[4], i, st[2], nxt[2]);
        const float dyy = kfwd(qq[1], qc[1], i, st[1], nxt[1]);
        const float dxxy = kfwd(qq[3], qc[3], i, st[0], nxt[0]);
        const float dzyz = kfwd(qq[5], qc[5], i, st[2], nxt[2]);
        const float dzz = kfwd(qq[2], qc[2], i, st[2], nxt[2]);
        const float dxxz = kfwd(qq[4], qc[4], i, st[0], nxt[0]);
        const float dyyz = kfwd(qq[5], qc[5], i, st[1], nxt[1]);
        const float w3[3] = {(dxx + dyxy) + dzxz, (dyy + dxxy) + dzyz, (dzz + dxxz) + dyyz};
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const float vo = vv[k][i];
            const float vn = fmaf(-tau, w3[k] - pc[k], vo);
            vv[k][i] = vn;
            vb[k][i] = fmaf(theta, vn - vo, vn);
        }

        // levels: w_i update from H(r_i) - q~ (tgv3_common.cu:362-405;
        // H: tgvk_common.cu:138-182 forward for even i, tgv3_common.cu:265-309 backward for odd i)
        float qt[6];
#pragma unroll
        for (int k = 0; k < 6; ++k) qt[k] = qc[k];
        for (int li = 0; li < levels; ++li) {
            float* const* lv = LV + size_t(li) * 18;
            const bool forward = (li & 1) == 0;
            float rc[6];
#pragma unroll
            for (int k = 0; k < 6; ++k) rc[k] = lv[12 + k][i];
#pragma unroll
            for (int k = 0; k < 6; ++k) {
                const float* rr = lv[12 + k];
                const int a = comp_axis_a(k), b = comp_axis_b(k);
                float h = forward ? kfwd(rr, rc[k], i, st[a], nxt[a])
                                  : kbwd(rr, rc[k], i, st[a], prv[a]);
                if (b >= 0)
                    h = h + (forward ? kfwd(rr, rc[k], i, st[b], nxt[b])
                                     : kbwd(rr, rc[k], i, st[b], prv[b]));
                const float wo = lv[k][i];
                const float wn = fmaf(-tau, h - qt[k], wo);
                lv[k][i] = wn;
                lv[6 + k][i] = fmaf(theta, wn - wo, wn);
            }
#pragma unroll
            for (int k = 0; k < 6; ++k) qt[k] = rc[k];
        }
    }
}

dim3 tgvk_grid(const Geometry& g, int* zchunk) {
    const int nbx = (g.w + KTX - 1) / KTX, nby = (g.h + KTY - 1) / KTY;
    const long cols = long(nbx) * nby;
    long nz = 4096 / (cols > 0 ? cols : 1);
    if (nz < 1) nz = 1;
    if (nz > g.d) nz = g.d;
    const int zc = int((g.d + nz - 1) / nz);
    *zchunk = zc;
    return dim3(unsigned(nbx), unsigned(nby), unsigned((g.d + zc - 1) / zc));
}

}  // namespace

hipError_t launch_tgvk_dual(hipStream_t s, float* const* fields_dev, float* const* levels_dev,
                            const float* alpha_dev, Geometry g, int levels, float sigma,
                            int planar) {
    int zc = 1;
    const dim3 grid = tgvk_grid(g, &zc);
    hipLaunchKernelGGL(tgvk_dual_kernel, grid, dim3(KTX, KTY, 1), 0, s, fields_dev, levels_dev,
                       alpha_dev, g, levels, sigma, planar, zc);
    return hipGetLastError();
}

hipError_t launch_tgvk_primal(hipStream_t s, float* const* fields_dev, float* const* levels_dev,
                              const float* f, Geometry g, int levels, float tau, float lambda,
                              float theta) {
    int zc = 1;
    const dim3 grid = tgvk_grid(g, &zc);
    hipLaunchKernelGGL(tgvk_primal_kernel, grid, dim3(KTX, KTY, 1), 0, s, fields_dev, levels_dev,
                       f, g, levels, tau, lambda, theta, zc);
    return hipGetLastError();
}

}  // namespace amb
