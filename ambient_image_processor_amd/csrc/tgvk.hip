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
// Each thread owns one (x, y) column of a 64 x 4 tile and marches along z over
// `zchunk` planes; tiles are dealt so that each XCD holds a contiguous range
// (whole z-slabs), so halo rows are read through its own L2.  Level fields are
// reached through a device table of 18 pointers per level (w[6], w_bar[6],
// r[6], components x, y, z, xy, xz, yz).  For 1 to 3 levels the kernels are
// instantiated per level count: every load of a plane is issued before any of
// its stores, and the -z neighbours (previous plane's centres) stay in
// registers.  Higher orders run the generic kernels, which loop over the
// levels at run time.
//
// Per-voxel arithmetic keeps the reference's FP32 operation order (fmaf where
// the reference writes fmaf or nvcc contracts a multiply-add; correctly
// rounded division and sqrt; built with -ffp-contract=off).
#include "amb_internal.hpp"

#include <algorithm>
#include <cstdlib>

namespace amb {
namespace {

constexpr int KTX = 64, KTY = 4, KT = KTX * KTY;

// tensor component c -> its first and second axis (-1: diagonal component)
__device__ __forceinline__ int comp_axis_a(int c) { return c < 3 ? c : (c == 5 ? 1 : 0); }
__device__ __forceinline__ int comp_axis_b(int c) { return c < 3 ? -1 : (c == 3 ? 1 : 2); }

// Block id -> tile id such that each XCD (blocks are dealt round-robin to the
// 8 XCDs) receives a contiguous range of tiles.
__device__ __forceinline__ int kxcd_tile(int bid, int nblocks) {
    if ((nblocks & 7) != 0) return bid;
    return (bid & 7) * (nblocks >> 3) + (bid >> 3);
}

struct KCoord {
    int x, y, z0, z1;
    size_t plane, col;
    bool in;
    __device__ __forceinline__ KCoord(const Geometry& g, int zchunk) {
        const int nbx = (g.w + KTX - 1) / KTX, nby = (g.h + KTY - 1) / KTY;
        const int t = kxcd_tile(int(blockIdx.x), int(gridDim.x));
        const int bx = t % nbx, r = t / nbx, by = r % nby, bz = r / nby;
        x = bx * KTX + int(threadIdx.x) % KTX;
        y = by * KTY + int(threadIdx.x) / KTX;
        z0 = bz * zchunk;
        z1 = min(z0 + zchunk, g.d);
        in = x < g.w && y < g.h;
        plane = size_t(g.w) * size_t(g.h);
        col = size_t(y) * size_t(g.w) + size_t(x);
    }
};

// Loads and stores at a 32-bit byte offset from a uniform base (one offset
// register serves every field; used where 4 * voxels < 2^31).
__device__ __forceinline__ float LDO(const float* p, uint32_t off) {
    return *reinterpret_cast<const float*>(reinterpret_cast<const char*>(p) + off);
}
__device__ __forceinline__ void STO(float* p, uint32_t off, float v) {
    *reinterpret_cast<float*>(reinterpret_cast<char*>(p) + off) = v;
}

// slot of a z-differenced tensor component (2: z, 4: xz, 5: yz) in the carries
__device__ __forceinline__ int zslot(int c) { return c == 2 ? 0 : (c == 4 ? 1 : 2); }

// Reference "backward difference" (tgv_cuda/tgv_common.cu:133-197): -a[i] + a[i-1], -a[0] at 0.
__device__ __forceinline__ float kbwd(const float* a, float center, size_t i, size_t st, bool has) {
    return has ? -center + a[i - st] : -center;
}
// Reference forward difference (tgv_common.cu:66-130): a[i+1] - a[i], 0 at the last index.
__device__ __forceinline__ float kfwd(const float* a, float center, size_t i, size_t st, bool has) {
    return has ? a[i + st] - center : 0.0f;
}

__global__ __launch_bounds__(KT) void tgvk_dual_kernel(
    float* const* __restrict__ F, float* const* __restrict__ LV,
    const float* __restrict__ alpha, const uint8_t* __restrict__ act, Geometry g, int levels,
    float sigma, int planar, int zchunk) {
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
        if (act && !act[i]) continue;   // masked loop: out-of-mask voxels keep their values
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

__global__ __launch_bounds__(KT) void tgvk_primal_kernel(
    float* const* __restrict__ F, float* const* __restrict__ LV, const float* __restrict__ f,
    const uint8_t* __restrict__ act, Geometry g, int levels, float tau, float lambda, float theta,
    int zchunk) {
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
        if (act && !act[i]) continue;
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

// The same updates for a compile-time level count, loads first.
template <int L>
__global__ __launch_bounds__(KT) void tgvk_dual_fixed(
    float* const* __restrict__ F, float* const* __restrict__ LV,
    const float* __restrict__ alpha, const uint8_t* __restrict__ act, Geometry g, float sigma,
    int planar, int zchunk) {
    const KCoord c(g, zchunk);
    if (!c.in) return;
    const int W = g.w, H = g.h, D = g.d;
    const uint32_t st[3] = {4u, 4u * uint32_t(W), 4u * uint32_t(c.plane)};
    const float a_p = alpha[0], a_q = alpha[1];
    const float* ub = F[F_UBAR];
    const float* vb[3] = {F[F_VBX], F[F_VBY], F[F_VBZ]};
    float* pp[3] = {F[F_PX], F[F_PY], F[F_PZ]};
    float* qq[6] = {F[F_QXX], F[F_QYY], F[F_QZZ], F[F_QXY], F[F_QXZ], F[F_QYZ]};
    const float* wb[L][6];
    float* rr[L][6];
    float ar[L];
#pragma unroll
    for (int l = 0; l < L; ++l) {
        ar[l] = alpha[l + 2];
#pragma unroll
        for (int k = 0; k < 6; ++k) {
            wb[l][k] = LV[18 * l + 6 + k];
            rr[l][k] = LV[18 * l + 12 + k];
        }
    }
    const bool prv0[2] = {c.x > 0, c.y > 0};
    const bool nxt0[2] = {c.x < W - 1, c.y < H - 1};

    // previous plane's v_bar and (even levels, backward G) w_bar z-components
    float vb_prev[3] = {0.f, 0.f, 0.f};
    float wb_prev[L][3];
#pragma unroll
    for (int l = 0; l < L; ++l) wb_prev[l][0] = wb_prev[l][1] = wb_prev[l][2] = 0.f;
    if (c.z0 > 0) {
        const uint32_t j = 4u * (uint32_t(c.z0 - 1) * uint32_t(c.plane) + uint32_t(c.col));
#pragma unroll
        for (int k = 0; k < 3; ++k) vb_prev[k] = LDO(vb[k], j);
#pragma unroll
        for (int l = 0; l < L; l += 2) {
            wb_prev[l][0] = LDO(wb[l][2], j);
            wb_prev[l][1] = LDO(wb[l][4], j);
            wb_prev[l][2] = LDO(wb[l][5], j);
        }
    }

    for (int z = c.z0; z < c.z1; ++z) {
        const uint32_t i = 4u * (uint32_t(z) * uint32_t(c.plane) + uint32_t(c.col));
        const bool prv[3] = {prv0[0], prv0[1], z > 0};
        const bool nxt[3] = {nxt0[0], nxt0[1], z < D - 1};
        // ---- loads
        const float u0 = LDO(ub, i);
        float un[3];
#pragma unroll
        for (int a = 0; a < 3; ++a) un[a] = nxt[a] ? LDO(ub, i + st[a]) : 0.f;
        float v0[3], vxm[3], vym[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            v0[k] = LDO(vb[k], i);
            vxm[k] = prv[0] ? LDO(vb[k], i - st[0]) : 0.f;
            vym[k] = prv[1] ? LDO(vb[k], i - st[1]) : 0.f;
        }
        float p0[3], q0[6];
#pragma unroll
        for (int k = 0; k < 3; ++k) p0[k] = LDO(pp[k], i);
#pragma unroll
        for (int k = 0; k < 6; ++k) q0[k] = LDO(qq[k], i);
        float w0[L][6], wa[L][6], wbn[L][6], r0[L][6];
#pragma unroll
        for (int l = 0; l < L; ++l) {
            const bool fw = (l & 1) != 0;
#pragma unroll
            for (int k = 0; k < 6; ++k) {
                w0[l][k] = LDO(wb[l][k], i);
                r0[l][k] = LDO(rr[l][k], i);
                const int a = comp_axis_a(k), b = comp_axis_b(k);
                if (fw) wa[l][k] = nxt[a] ? LDO(wb[l][k], i + st[a]) : 0.f;
                else wa[l][k] = (a == 2) ? wb_prev[l][0] : (prv[a] ? LDO(wb[l][k], i - st[a]) : 0.f);
                wbn[l][k] = 0.f;
                if (b == 2) wbn[l][k] = fw ? (nxt[2] ? LDO(wb[l][k], i + st[2]) : 0.f) : wb_prev[l][zslot(k)];
                else if (b >= 0) wbn[l][k] = fw ? (nxt[b] ? LDO(wb[l][k], i + st[b]) : 0.f)
                                                : (prv[b] ? LDO(wb[l][k], i - st[b]) : 0.f);
            }
        }

        // ---- p update + projection (tgv2_common.cu:96-131)
        float a3[3];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const float d = nxt[a] ? un[a] - u0 : 0.0f;
            a3[a] = fmaf(sigma, d - v0[a], p0[a]);
        }
        float n2 = a3[0] * a3[0];
        n2 = fmaf(a3[1], a3[1], n2);
        n2 = fmaf(a3[2], a3[2], n2);
        const float tp = sqrtf(n2) / a_p;
        const bool sp = tp > 1.0f;

        // ---- q update with eps_B(v_bar) - w_bar_0 (tgv2_common.cu:174-219, tgv3_common.cu:28-86)
        float bxv[3], byv[3], bzv[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            bxv[k] = prv[0] ? -v0[k] + vxm[k] : -v0[k];
            byv[k] = prv[1] ? -v0[k] + vym[k] : -v0[k];
            bzv[k] = planar ? 0.0f : (prv[2] ? -v0[k] + vb_prev[k] : -v0[k]);
        }
        float e[6];
        e[0] = bxv[0];
        e[1] = byv[1];
        e[2] = bzv[2];
        e[3] = (bxv[1] + byv[0]) * 0.5f;
        e[4] = (bxv[2] + bzv[0]) * 0.5f;
        e[5] = (byv[2] + bzv[1]) * 0.5f;
        float b6[6];
#pragma unroll
        for (int k = 0; k < 6; ++k) b6[k] = fmaf(sigma, e[k] - w0[0][k], q0[k]);
        float m2 = b6[0] * b6[0];
        m2 = fmaf(b6[1], b6[1], m2);
        m2 = fmaf(2.0f * b6[3], b6[3], m2);
        m2 = fmaf(b6[2], b6[2], m2);
        m2 = fmaf(2.0f * b6[4], b6[4], m2);
        m2 = fmaf(2.0f * b6[5], b6[5], m2);
        const float tq = sqrtf(m2) / a_q;
        const bool sq = tq > 1.0f;

        // ---- levels: r_l update + projection with alpha[l + 2]
        float rn[L][6];
#pragma unroll
        for (int l = 0; l < L; ++l) {
            const bool fw = (l & 1) != 0;
            float rc[6];
#pragma unroll
            for (int k = 0; k < 6; ++k) {
                const int a = comp_axis_a(k), b = comp_axis_b(k);
                float d = fw ? (nxt[a] ? wa[l][k] - w0[l][k] : 0.0f)
                             : (prv[a] ? -w0[l][k] + wa[l][k] : -w0[l][k]);
                if (b >= 0) {
                    const float d2 = fw ? (nxt[b] ? wbn[l][k] - w0[l][k] : 0.0f)
                                        : (prv[b] ? -w0[l][k] + wbn[l][k] : -w0[l][k]);
                    d = (d + d2) * 0.5f;
                }
                rc[k] = (l + 1 < L) ? fmaf(sigma, d - w0[l + 1][k], r0[l][k])
                                    : fmaf(sigma, d, r0[l][k]);
            }
            float r2 = rc[0] * rc[0];
            r2 = fmaf(rc[1], rc[1], r2);
            r2 = fmaf(2.0f * rc[3], rc[3], r2);
            r2 = fmaf(rc[2], rc[2], r2);
            r2 = fmaf(2.0f * rc[4], rc[4], r2);
            r2 = fmaf(2.0f * rc[5], rc[5], r2);
            const float tr = sqrtf(r2) / ar[l];
            const bool sr = tr > 1.0f;
#pragma unroll
            for (int k = 0; k < 6; ++k) rn[l][k] = sr ? rc[k] / tr : rc[k];
        }

        // ---- stores (masked loop: only in-mask voxels)
        if (!act || act[i >> 2]) {
#pragma unroll
            for (int k = 0; k < 3; ++k) STO(pp[k], i, sp ? a3[k] / tp : a3[k]);
#pragma unroll
            for (int k = 0; k < 6; ++k) STO(qq[k], i, sq ? b6[k] / tq : b6[k]);
#pragma unroll
            for (int l = 0; l < L; ++l)
#pragma unroll
                for (int k = 0; k < 6; ++k) STO(rr[l][k], i, rn[l][k]);
        }

        // ---- carries
#pragma unroll
        for (int k = 0; k < 3; ++k) vb_prev[k] = v0[k];
#pragma unroll
        for (int l = 0; l < L; l += 2) {
            wb_prev[l][0] = w0[l][2];
            wb_prev[l][1] = w0[l][4];
            wb_prev[l][2] = w0[l][5];
        }
    }
}

template <int L>
__global__ __launch_bounds__(KT) void tgvk_primal_fixed(
    float* const* __restrict__ F, float* const* __restrict__ LV, const float* __restrict__ f,
    const uint8_t* __restrict__ act, Geometry g, float tau, float lambda, float theta,
    int zchunk) {
    const KCoord c(g, zchunk);
    if (!c.in) return;
    const int W = g.w, H = g.h, D = g.d;
    const uint32_t st[3] = {4u, 4u * uint32_t(W), 4u * uint32_t(c.plane)};
    const float tl = tau * lambda;
    float* uu = F[F_U];
    float* ub = F[F_UBAR];
    float* vv[3] = {F[F_VX], F[F_VY], F[F_VZ]};
    float* vb[3] = {F[F_VBX], F[F_VBY], F[F_VBZ]};
    const float* pp[3] = {F[F_PX], F[F_PY], F[F_PZ]};
    const float* qq[6] = {F[F_QXX], F[F_QYY], F[F_QZZ], F[F_QXY], F[F_QXZ], F[F_QYZ]};
    float* ww[L][6];
    float* wbb[L][6];
    const float* rr[L][6];
#pragma unroll
    for (int l = 0; l < L; ++l)
#pragma unroll
        for (int k = 0; k < 6; ++k) {
            ww[l][k] = LV[18 * l + k];
            wbb[l][k] = LV[18 * l + 6 + k];
            rr[l][k] = LV[18 * l + 12 + k];
        }
    const bool prv0[2] = {c.x > 0, c.y > 0};
    const bool nxt0[2] = {c.x < W - 1, c.y < H - 1};

    // previous plane's p_z and (odd levels, backward H) r z-components
    float pz_prev = 0.f;
    float r_prev[L][3];
#pragma unroll
    for (int l = 0; l < L; ++l) r_prev[l][0] = r_prev[l][1] = r_prev[l][2] = 0.f;
    if (c.z0 > 0) {
        const uint32_t j = 4u * (uint32_t(c.z0 - 1) * uint32_t(c.plane) + uint32_t(c.col));
        pz_prev = LDO(pp[2], j);
#pragma unroll
        for (int l = 1; l < L; l += 2) {
            r_prev[l][0] = LDO(rr[l][2], j);
            r_prev[l][1] = LDO(rr[l][4], j);
            r_prev[l][2] = LDO(rr[l][5], j);
        }
    }

    for (int z = c.z0; z < c.z1; ++z) {
        const uint32_t i = 4u * (uint32_t(z) * uint32_t(c.plane) + uint32_t(c.col));
        const bool prv[3] = {prv0[0], prv0[1], z > 0};
        const bool nxt[3] = {nxt0[0], nxt0[1], z < D - 1};
        // ---- loads
        float p0[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) p0[k] = LDO(pp[k], i);
        const float pxm = prv[0] ? LDO(pp[0], i - st[0]) : 0.f;
        const float pym = prv[1] ? LDO(pp[1], i - st[1]) : 0.f;
        const float uo = LDO(uu, i), fc = LDO(f, i);
        float q0[6], qa[6], qbn[6];
#pragma unroll
        for (int k = 0; k < 6; ++k) {
            q0[k] = LDO(qq[k], i);
            const int a = comp_axis_a(k), b = comp_axis_b(k);
            qa[k] = nxt[a] ? LDO(qq[k], i + st[a]) : 0.f;
            qbn[k] = (b >= 0 && nxt[b]) ? LDO(qq[k], i + st[b]) : 0.f;
        }
        float v0[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) v0[k] = LDO(vv[k], i);
        float r0[L][6], ra[L][6], rbn[L][6], w0[L][6];
#pragma unroll
        for (int l = 0; l < L; ++l) {
            const bool fw = (l & 1) == 0;
#pragma unroll
            for (int k = 0; k < 6; ++k) {
                r0[l][k] = LDO(rr[l][k], i);
                w0[l][k] = LDO(ww[l][k], i);
                const int a = comp_axis_a(k), b = comp_axis_b(k);
                if (fw) ra[l][k] = nxt[a] ? LDO(rr[l][k], i + st[a]) : 0.f;
                else ra[l][k] = (a == 2) ? r_prev[l][0] : (prv[a] ? LDO(rr[l][k], i - st[a]) : 0.f);
                rbn[l][k] = 0.f;
                if (b == 2) rbn[l][k] = fw ? (nxt[2] ? LDO(rr[l][k], i + st[2]) : 0.f) : r_prev[l][zslot(k)];
                else if (b >= 0) rbn[l][k] = fw ? (nxt[b] ? LDO(rr[l][k], i + st[b]) : 0.f)
                                                : (prv[b] ? LDO(rr[l][k], i - st[b]) : 0.f);
            }
        }

        // ---- u update, L1 prox, over-relaxation (tgv2_l1.cu:27-60)
        const float bx = prv[0] ? -p0[0] + pxm : -p0[0];
        const float by = prv[1] ? -p0[1] + pym : -p0[1];
        const float bz = prv[2] ? -p0[2] + pz_prev : -p0[2];
        const float div = (bx + by) + bz;
        float un = fmaf(-tau, div, uo);
        const float umf = un - fc;
        if (umf > tl) un = un - tl;
        else if (umf < -tl) un = un + tl;
        else un = fc;

        // ---- div2(q), v update (tgv2_common.cu:279-354)
        const float dxx = nxt[0] ? qa[0] - q0[0] : 0.0f;
        const float dyxy = nxt[1] ? qbn[3] - q0[3] : 0.0f;
        const float dzxz = nxt[2] ? qbn[4] - q0[4] : 0.0f;
        const float dyy = nxt[1] ? qa[1] - q0[1] : 0.0f;
        const float dxxy = nxt[0] ? qa[3] - q0[3] : 0.0f;
        const float dzyz = nxt[2] ? qbn[5] - q0[5] : 0.0f;
        const float dzz = nxt[2] ? qa[2] - q0[2] : 0.0f;
        const float dxxz = nxt[0] ? qa[4] - q0[4] : 0.0f;
        const float dyyz = nxt[1] ? qa[5] - q0[5] : 0.0f;
        const float w3[3] = {(dxx + dyxy) + dzxz, (dyy + dxxy) + dzyz, (dzz + dxxz) + dyyz};
        float vn[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) vn[k] = fmaf(-tau, w3[k] - p0[k], v0[k]);

        // ---- levels: w_l from H(r_l) - q~ (tgv3_common.cu:362-405)
        float wn[L][6];
#pragma unroll
        for (int l = 0; l < L; ++l) {
            const bool fw = (l & 1) == 0;
#pragma unroll
            for (int k = 0; k < 6; ++k) {
                const int a = comp_axis_a(k), b = comp_axis_b(k);
                float h = fw ? (nxt[a] ? ra[l][k] - r0[l][k] : 0.0f)
                             : (prv[a] ? -r0[l][k] + ra[l][k] : -r0[l][k]);
                if (b >= 0)
                    h = h + (fw ? (nxt[b] ? rbn[l][k] - r0[l][k] : 0.0f)
                                : (prv[b] ? -r0[l][k] + rbn[l][k] : -r0[l][k]));
                const float qt = l == 0 ? q0[k] : r0[l > 0 ? l - 1 : 0][k];
                wn[l][k] = fmaf(-tau, h - qt, w0[l][k]);
            }
        }

        // ---- stores (masked loop: only in-mask voxels)
        if (!act || act[i >> 2]) {
            STO(uu, i, un);
            STO(ub, i, fmaf(theta, un - uo, un));
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                STO(vv[k], i, vn[k]);
                STO(vb[k], i, fmaf(theta, vn[k] - v0[k], vn[k]));
            }
#pragma unroll
            for (int l = 0; l < L; ++l)
#pragma unroll
                for (int k = 0; k < 6; ++k) {
                    STO(ww[l][k], i, wn[l][k]);
                    STO(wbb[l][k], i, fmaf(theta, wn[l][k] - w0[l][k], wn[l][k]));
                }
        }

        // ---- carries
        pz_prev = p0[2];
#pragma unroll
        for (int l = 1; l < L; l += 2) {
            r_prev[l][0] = r0[l][2];
            r_prev[l][1] = r0[l][4];
            r_prev[l][2] = r0[l][5];
        }
    }
}

// ---------------------------------------------------------------------------
// Fused sweep for one level (k = 3): the primal half of iteration k (u, v, w)
// and the dual half of iteration k + 1 (p, q, r) in one z-march, with the
// per-voxel arithmetic of tgvk_primal_fixed<1> and tgvk_dual_fixed<1> in the
// same order.
//
// The dual half at a voxel reads u_bar at its +x, +y, +z neighbours and
// v_bar, w_bar at its -x, -y, -z neighbours, i.e. primal results of
// neighbouring voxels.  A block owns a tile of `tx` x columns (one voxel per
// thread; tx = W: whole rows) and `ty` y rows; it computes the primal half
// also on one column left and right of the tile (tx < W), one row above and
// below it and one plane before and after its z chunk, and runs the dual half
// one plane behind the primal half, so u_bar, v_bar and w_bar never leave the
// CU.  Neighbouring blocks recompute the halo rows from the same
// inputs, so the iteration state is read from one copy (src) and written to
// another (dst).  Per voxel-iteration: reads u, f, v x3, p x3, q x6, w x6,
// r x6 = 26, writes u, v x3, p x3, q x6, w x6, r x6 = 25: 51 element accesses
// (204 B) instead of the 86 (344 B) of the two launches.
//
// Thread (c, r) holds y = y0 - 1 + r and x = x0 - 1 + c (tiled) or x = c.
// Rows 1..ty and columns 1..tx are owned; row and column 0 feed only v_bar and
// w_bar (the dual half's backward differences), row ty + 1 and column tx + 1
// only u_bar (its forward differences).  Step s computes the primal half at
// plane s and the dual half at plane s - 1.
//
// LDS arrays (one float per thread each): P region, plane s, the dual fields whose x
// or y neighbours the primal half reads; D region, plane s - 1, u_bar, v_bar
// and w_bar.  z neighbours stay in registers: p_z and the dual fields of
// plane s - 1, q and r z-components of plane s + 1 (loaded one plane ahead),
// v_bar and w_bar z-components of plane s - 2.
// ---------------------------------------------------------------------------
enum KFusedField { KF_U = 0, KF_VX, KF_VY, KF_VZ, KF_PX, KF_PY, KF_PZ, KF_QXX, KF_QYY, KF_QZZ,
                   KF_QXY, KF_QXZ, KF_QYZ, KF_WX, KF_WY, KF_WZ, KF_WXY, KF_WXZ, KF_WYZ, KF_RX,
                   KF_RY, KF_RZ, KF_RXY, KF_RXZ, KF_RYZ, N_KF };
static_assert(N_KF == N_TGVK_FUSED_FIELDS, "fused TGV-k field count");

// dst field k = src field k + dt bytes (k < 13) or + dl bytes (k >= 13): the
// host lays the second copies out so, which keeps the 25 destination pointers
// out of the scalar registers.

struct KFusedFields {
    const float* src[N_KF];
    long long dt, dl;
};

enum KFusedRow { KR_PX = 0, KR_PY, KR_QXX, KR_QXY, KR_QYY, KR_QXZ, KR_QYZ, KR_RX, KR_RY, KR_RXY,
                 KR_RXZ, KR_RYZ,                                            // P region
                 KR_UB, KR_VBX, KR_VBY, KR_VBZ, KR_WBX, KR_WBY, KR_WBZ, KR_WBXY, KR_WBXZ,
                 KR_WBYZ,                                                   // D region
                 N_KROWS };
constexpr int KFUSED_MAX_THREADS = 1024;

template <class I>
__device__ __forceinline__ float kf_ld(const float* p, I i) {
    if constexpr (sizeof(I) == 4) return LDO(p, uint32_t(i) * 4u);
    else return p[i];
}
template <class I>
__device__ __forceinline__ void kf_st(float* p, I i, float v) {
    if constexpr (sizeof(I) == 4)
        __builtin_nontemporal_store(v, reinterpret_cast<float*>(reinterpret_cast<char*>(p) +
                                                                uint32_t(i) * 4u));
    else
        __builtin_nontemporal_store(v, p + i);
}

template <class I, int MK>
__global__ __launch_bounds__(KFUSED_MAX_THREADS) void tgvk_fused_kernel(
    KFusedFields ff, const float* __restrict__ f_, const uint8_t* __restrict__ act, Geometry g,
    float tau, float sigma, float lambda, float theta, float a_p, float a_q, float a_r, int tx,
    int ty, int zchunk, int planar) {
    extern __shared__ __align__(16) float lds[];
    const int W = g.w, H = g.h, D = g.d;
    // tx < W: tiles of tx columns with one halo column on each side
    const bool tiled = tx < W;
    const int TXH = tiled ? tx + 2 : W, RT = ty + 2, RW = RT * TXH;
    const int c = int(threadIdx.x) % TXH, r = int(threadIdx.x) / TXH;
    const int nbx = (W + tx - 1) / tx, nb = (H + ty - 1) / ty;
    const int t = kxcd_tile(int(blockIdx.x), int(gridDim.x));
    const int tb = t / nbx;
    const int x = tiled ? (t % nbx) * tx - 1 + c : c;
    const int y0 = (tb % nb) * ty;
    const int z0 = (tb / nb) * zchunk, z1 = min(z0 + zchunk, D);
    const int y = y0 - 1 + r;
    const bool in = y >= 0 && y < H && x >= 0 && x < W;
    // column 0 feeds only v_bar and w_bar, column tx + 1 only u_bar (tiled)
    const bool cl = !tiled || c > 0, cr = !tiled || c < TXH - 1;
    const bool own = in && r >= 1 && r <= ty && cl && cr;
    const bool xp = x > 0, xn = x < W - 1, yp = y > 0, yn = y < H - 1;
    const I plane = I(W) * I(H);
    const I col = in ? I(y) * I(W) + I(x) : I(0);
    const float tl = tau * lambda;
    const int me = r * TXH + c;
    auto R = [&](int k) { return lds + k * RW; };
    const float* const* src = ff.src;
    auto dst = [&](int k) {
        return reinterpret_cast<float*>(
            const_cast<char*>(reinterpret_cast<const char*>(src[k])) +
            (k < KF_WX ? ff.dt : ff.dl));
    };

    // D region starts at zero (read before its first write only into unused carries)
    for (int k = int(threadIdx.x); k < (N_KROWS - KR_UB) * RW; k += int(blockDim.x))
        lds[KR_UB * RW + k] = 0.f;

    float p_c[3], q_c[6], r_c[6];     // dual fields at plane s - 1
    float qn[3], rn[3];               // q_zz, q_xz, q_yz and r_z, r_xz, r_yz at plane s
    float vb_m2[3], wb_m2[3];         // v_bar, and w_bar z, xz, yz at plane s - 2
#pragma unroll
    for (int k = 0; k < 3; ++k) p_c[k] = qn[k] = rn[k] = vb_m2[k] = wb_m2[k] = 0.f;
#pragma unroll
    for (int k = 0; k < 6; ++k) q_c[k] = r_c[k] = 0.f;
    bool act_c = true;

    const int zs = z0 > 0 ? z0 - 1 : 0;
    if (in && zs < D) {
        const I i = I(zs) * plane + col;
        qn[0] = kf_ld(src[KF_QZZ], i); qn[1] = kf_ld(src[KF_QXZ], i); qn[2] = kf_ld(src[KF_QYZ], i);
        rn[0] = kf_ld(src[KF_RZ], i); rn[1] = kf_ld(src[KF_RXZ], i); rn[2] = kf_ld(src[KF_RYZ], i);
    }

    for (int s = zs; s <= z1; ++s) {
        const bool pin = s < D;
        const bool zn = s < D - 1;
        const I i = I(s) * plane + col;
        // ---- loads at plane s (q and r z-components at s + 1)
        float px = 0.f, py = 0.f, pz = 0.f, qxx = 0.f, qxy = 0.f, qyy = 0.f, u = 0.f, fv = 0.f;
        float v[3] = {0.f, 0.f, 0.f}, qnn[3] = {0.f, 0.f, 0.f}, rnn[3] = {0.f, 0.f, 0.f};
        float rx = 0.f, ry = 0.f, rxy = 0.f;
        float w[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
        bool actv = true;
        if (in && pin) {
            px = kf_ld(src[KF_PX], i); py = kf_ld(src[KF_PY], i); pz = kf_ld(src[KF_PZ], i);
            qxx = kf_ld(src[KF_QXX], i); qxy = kf_ld(src[KF_QXY], i); qyy = kf_ld(src[KF_QYY], i);
            u = kf_ld(src[KF_U], i); fv = kf_ld(f_, i);
#pragma unroll
            for (int k = 0; k < 3; ++k) v[k] = kf_ld(src[KF_VX + k], i);
            rx = kf_ld(src[KF_RX], i); ry = kf_ld(src[KF_RY], i); rxy = kf_ld(src[KF_RXY], i);
#pragma unroll
            for (int k = 0; k < 6; ++k) w[k] = kf_ld(src[KF_WX + k], i);
            if (zn) {
                const I j = i + plane;
                qnn[0] = kf_ld(src[KF_QZZ], j); qnn[1] = kf_ld(src[KF_QXZ], j);
                qnn[2] = kf_ld(src[KF_QYZ], j);
                rnn[0] = kf_ld(src[KF_RZ], j); rnn[1] = kf_ld(src[KF_RXZ], j);
                rnn[2] = kf_ld(src[KF_RYZ], j);
            }
            if (MK) actv = act[i] != 0;
        }

        // ---- P region: plane s
        R(KR_PX)[me] = px; R(KR_PY)[me] = py;
        R(KR_QXX)[me] = qxx; R(KR_QXY)[me] = qxy; R(KR_QYY)[me] = qyy;
        R(KR_QXZ)[me] = qn[1]; R(KR_QYZ)[me] = qn[2];
        R(KR_RX)[me] = rx; R(KR_RY)[me] = ry; R(KR_RXY)[me] = rxy;
        R(KR_RXZ)[me] = rn[1]; R(KR_RYZ)[me] = rn[2];
        __syncthreads();

        // ---- primal half at plane s
        const bool need_u = in && pin && r > 0 && cl && s >= z0;
        const bool need_v = in && pin && r < RT - 1 && cr && s < z1;
        float ubn = 0.f, vbn[3] = {0.f, 0.f, 0.f}, wbn[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
        float uo = 0.f, vo[3], wo[6];
        if (need_u) {
            // u update, L1 prox, over-relaxation (tgv2_l1.cu:27-60)
            const float pxm = xp ? R(KR_PX)[me - 1] : 0.f;
            const float pym = yp ? R(KR_PY)[me - TXH] : 0.f;
            const float bx = xp ? -px + pxm : -px;
            const float by = yp ? -py + pym : -py;
            const float bz = s > 0 ? -pz + p_c[2] : -pz;
            const float div = (bx + by) + bz;
            float un = fmaf(-tau, div, u);
            const float umf = un - fv;
            if (umf > tl) un = un - tl;
            else if (umf < -tl) un = un + tl;
            else un = fv;
            if (MK && !actv) un = u;
            uo = un;
            ubn = fmaf(theta, un - u, un);
        }
        if (need_v) {
            // div2(q) with forward differences, v update (tgv2_common.cu:279-354)
            const float qxx_x = xn ? R(KR_QXX)[me + 1] : 0.f;
            const float qxy_x = xn ? R(KR_QXY)[me + 1] : 0.f;
            const float qxz_x = xn ? R(KR_QXZ)[me + 1] : 0.f;
            const float qxy_y = yn ? R(KR_QXY)[me + TXH] : 0.f;
            const float qyy_y = yn ? R(KR_QYY)[me + TXH] : 0.f;
            const float qyz_y = yn ? R(KR_QYZ)[me + TXH] : 0.f;
            const float dxx = xn ? qxx_x - qxx : 0.0f;
            const float dyxy = yn ? qxy_y - qxy : 0.0f;
            const float dzxz = zn ? qnn[1] - qn[1] : 0.0f;
            const float dyy = yn ? qyy_y - qyy : 0.0f;
            const float dxxy = xn ? qxy_x - qxy : 0.0f;
            const float dzyz = zn ? qnn[2] - qn[2] : 0.0f;
            const float dzz = zn ? qnn[0] - qn[0] : 0.0f;
            const float dxxz = xn ? qxz_x - qn[1] : 0.0f;
            const float dyyz = yn ? qyz_y - qn[2] : 0.0f;
            const float w3[3] = {(dxx + dyxy) + dzxz, (dyy + dxxy) + dzyz, (dzz + dxxz) + dyyz};
            const float p0[3] = {px, py, pz};
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                float vn = fmaf(-tau, w3[k] - p0[k], v[k]);
                if (MK && !actv) vn = v[k];
                vo[k] = vn;
                vbn[k] = fmaf(theta, vn - v[k], vn);
            }
            // w from H(r) (forward differences) - q (tgv3_common.cu:362-405,
            // tgvk_common.cu:138-182)
            const float rx_x = xn ? R(KR_RX)[me + 1] : 0.f;
            const float rxy_x = xn ? R(KR_RXY)[me + 1] : 0.f;
            const float rxz_x = xn ? R(KR_RXZ)[me + 1] : 0.f;
            const float ry_y = yn ? R(KR_RY)[me + TXH] : 0.f;
            const float rxy_y = yn ? R(KR_RXY)[me + TXH] : 0.f;
            const float ryz_y = yn ? R(KR_RYZ)[me + TXH] : 0.f;
            float h[6];
            h[0] = xn ? rx_x - rx : 0.0f;
            h[1] = yn ? ry_y - ry : 0.0f;
            h[2] = zn ? rnn[0] - rn[0] : 0.0f;
            h[3] = (xn ? rxy_x - rxy : 0.0f) + (yn ? rxy_y - rxy : 0.0f);
            h[4] = (xn ? rxz_x - rn[1] : 0.0f) + (zn ? rnn[1] - rn[1] : 0.0f);
            h[5] = (yn ? ryz_y - rn[2] : 0.0f) + (zn ? rnn[2] - rn[2] : 0.0f);
            const float q0[6] = {qxx, qyy, qn[0], qxy, qn[1], qn[2]};
#pragma unroll
            for (int k = 0; k < 6; ++k) {
                float wn = fmaf(-tau, h[k] - q0[k], w[k]);
                if (MK && !actv) wn = w[k];
                wo[k] = wn;
                wbn[k] = fmaf(theta, wn - w[k], wn);
            }
        }
        if (own && pin && s >= z0 && s < z1) {
            kf_st(dst(KF_U), i, uo);
#pragma unroll
            for (int k = 0; k < 3; ++k) kf_st(dst(KF_VX + k), i, vo[k]);
#pragma unroll
            for (int k = 0; k < 6; ++k) kf_st(dst(KF_WX + k), i, wo[k]);
        }

        // ---- dual half at plane zd = s - 1 (D region holds its u_bar, v_bar, w_bar)
        const int zd = s - 1;
        float vbc[3], wbc[6];
#pragma unroll
        for (int k = 0; k < 3; ++k) vbc[k] = R(KR_VBX + k)[me];
#pragma unroll
        for (int k = 0; k < 6; ++k) wbc[k] = R(KR_WBX + k)[me];
        if (own && zd >= z0) {
            // p update + projection onto |p| <= alpha[0] (tgv2_common.cu:96-131)
            const float ubc = R(KR_UB)[me];
            const float uxn = xn ? R(KR_UB)[me + 1] : 0.f;
            const float uyn = yn ? R(KR_UB)[me + TXH] : 0.f;
            const float dx = xn ? uxn - ubc : 0.0f;
            const float dy = yn ? uyn - ubc : 0.0f;
            const float dz = zd < D - 1 ? ubn - ubc : 0.0f;
            float a3[3];
            a3[0] = fmaf(sigma, dx - vbc[0], p_c[0]);
            a3[1] = fmaf(sigma, dy - vbc[1], p_c[1]);
            a3[2] = fmaf(sigma, dz - vbc[2], p_c[2]);
            float n2 = a3[0] * a3[0];
            n2 = fmaf(a3[1], a3[1], n2);
            n2 = fmaf(a3[2], a3[2], n2);
            const float tp = sqrtf(n2) / a_p;
            if (tp > 1.0f) {   // wave-uniform branch: the divisions only where a lane projects
#pragma unroll
                for (int k = 0; k < 3; ++k) a3[k] = a3[k] / tp;
            }

            // q update with eps_B(v_bar) - w_bar (tgv2_common.cu:174-219, tgv3_common.cu:28-86)
            float bxv[3], byv[3], bzv[3];
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                bxv[k] = xp ? -vbc[k] + R(KR_VBX + k)[me - 1] : -vbc[k];
                byv[k] = yp ? -vbc[k] + R(KR_VBX + k)[me - TXH] : -vbc[k];
                bzv[k] = planar ? 0.0f : (zd > 0 ? -vbc[k] + vb_m2[k] : -vbc[k]);
            }
            float e[6];
            e[0] = bxv[0];
            e[1] = byv[1];
            e[2] = bzv[2];
            e[3] = (bxv[1] + byv[0]) * 0.5f;
            e[4] = (bxv[2] + bzv[0]) * 0.5f;
            e[5] = (byv[2] + bzv[1]) * 0.5f;
            float b6[6];
#pragma unroll
            for (int k = 0; k < 6; ++k) b6[k] = fmaf(sigma, e[k] - wbc[k], q_c[k]);
            float m2 = b6[0] * b6[0];
            m2 = fmaf(b6[1], b6[1], m2);
            m2 = fmaf(2.0f * b6[3], b6[3], m2);
            m2 = fmaf(b6[2], b6[2], m2);
            m2 = fmaf(2.0f * b6[4], b6[4], m2);
            m2 = fmaf(2.0f * b6[5], b6[5], m2);
            const float tq = sqrtf(m2) / a_q;
            if (tq > 1.0f) {
#pragma unroll
                for (int k = 0; k < 6; ++k) b6[k] = b6[k] / tq;
            }

            // r update from G(w_bar) (backward differences), projection with
            // alpha[2] (tgv3_common.cu:214-262, 311-360)
            float gg[6];
            gg[0] = xp ? -wbc[0] + R(KR_WBX)[me - 1] : -wbc[0];
            gg[1] = yp ? -wbc[1] + R(KR_WBY)[me - TXH] : -wbc[1];
            gg[2] = zd > 0 ? -wbc[2] + wb_m2[0] : -wbc[2];
            {
                const float d = xp ? -wbc[3] + R(KR_WBXY)[me - 1] : -wbc[3];
                const float d2 = yp ? -wbc[3] + R(KR_WBXY)[me - TXH] : -wbc[3];
                gg[3] = (d + d2) * 0.5f;
            }
            {
                const float d = xp ? -wbc[4] + R(KR_WBXZ)[me - 1] : -wbc[4];
                const float d2 = zd > 0 ? -wbc[4] + wb_m2[1] : -wbc[4];
                gg[4] = (d + d2) * 0.5f;
            }
            {
                const float d = yp ? -wbc[5] + R(KR_WBYZ)[me - TXH] : -wbc[5];
                const float d2 = zd > 0 ? -wbc[5] + wb_m2[2] : -wbc[5];
                gg[5] = (d + d2) * 0.5f;
            }
            float rc[6];
#pragma unroll
            for (int k = 0; k < 6; ++k) rc[k] = fmaf(sigma, gg[k], r_c[k]);
            float r2 = rc[0] * rc[0];
            r2 = fmaf(rc[1], rc[1], r2);
            r2 = fmaf(2.0f * rc[3], rc[3], r2);
            r2 = fmaf(rc[2], rc[2], r2);
            r2 = fmaf(2.0f * rc[4], rc[4], r2);
            r2 = fmaf(2.0f * rc[5], rc[5], r2);
            const float tr = sqrtf(r2) / a_r;
            if (tr > 1.0f) {
#pragma unroll
                for (int k = 0; k < 6; ++k) rc[k] = rc[k] / tr;
            }

            const I id = I(zd) * plane + col;
            const bool keep = MK && !act_c;   // masked loop: out-of-mask voxels keep their values
#pragma unroll
            for (int k = 0; k < 3; ++k) kf_st(dst(KF_PX + k), id, keep ? p_c[k] : a3[k]);
#pragma unroll
            for (int k = 0; k < 6; ++k) kf_st(dst(KF_QXX + k), id, keep ? q_c[k] : b6[k]);
#pragma unroll
            for (int k = 0; k < 6; ++k) kf_st(dst(KF_RX + k), id, keep ? r_c[k] : rc[k]);
        }
        __syncthreads();

        // ---- D region <- plane s; carry plane s to the next step
        R(KR_UB)[me] = ubn;
#pragma unroll
        for (int k = 0; k < 3; ++k) R(KR_VBX + k)[me] = vbn[k];
#pragma unroll
        for (int k = 0; k < 6; ++k) R(KR_WBX + k)[me] = wbn[k];
#pragma unroll
        for (int k = 0; k < 3; ++k) vb_m2[k] = vbc[k];
        wb_m2[0] = wbc[2]; wb_m2[1] = wbc[4]; wb_m2[2] = wbc[5];
        p_c[0] = px; p_c[1] = py; p_c[2] = pz;
        q_c[0] = qxx; q_c[1] = qyy; q_c[2] = qn[0]; q_c[3] = qxy; q_c[4] = qn[1]; q_c[5] = qn[2];
        r_c[0] = rx; r_c[1] = ry; r_c[2] = rn[0]; r_c[3] = rxy; r_c[4] = rn[1]; r_c[5] = rn[2];
#pragma unroll
        for (int k = 0; k < 3; ++k) { qn[k] = qnn[k]; rn[k] = rnn[k]; }
        act_c = actv;
    }
}

template <class I, int MK>
hipError_t tgvk_fused_launch(hipStream_t s, const KFusedFields& ff, const float* f,
                             const uint8_t* act, Geometry g, float tau, float sigma, float lambda,
                             float theta, const float* alpha, int planar, const TgvkFusedPlan& plan) {
    const int RT = plan.ty + 2, TXH = plan.tx < g.w ? plan.tx + 2 : g.w;
    const int nbx = (g.w + plan.tx - 1) / plan.tx;
    const int nb = (g.h + plan.ty - 1) / plan.ty, nc = (g.d + plan.zchunk - 1) / plan.zchunk;
    const size_t lds = size_t(N_KROWS) * RT * TXH * sizeof(float);
    auto kern = tgvk_fused_kernel<I, MK>;
    hipError_t e = ensure_dynamic_lds(reinterpret_cast<const void*>(kern), lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(kern, dim3(nbx * nb * nc), dim3(TXH * RT), lds, s, ff, f, act, g, tau,
                       sigma, lambda, theta, alpha[0], alpha[1], alpha[2], plan.tx, plan.ty,
                       plan.zchunk, planar);
    return hipGetLastError();
}

// Planes per block: enough z-chunks for about 4096 blocks (the tgvk_zchunk
// override sets it, for tests of the chunk edges).
int tgvk_blocks(const Geometry& g, const PlanOverrides& ov, int* zchunk) {
    const int nbx = (g.w + KTX - 1) / KTX, nby = (g.h + KTY - 1) / KTY;
    const long cols = long(nbx) * nby;
    long nz = 4096 / (cols > 0 ? cols : 1);
    if (nz < 1) nz = 1;
    if (nz > g.d) nz = g.d;
    int zc = int((g.d + nz - 1) / nz);
    if (ov.tgvk_zchunk > 0) zc = std::min(ov.tgvk_zchunk, g.d);
    *zchunk = zc;
    return int(cols * ((g.d + zc - 1) / zc));
}

}  // namespace

// Plan of the fused TGV-k sweep (one level): blocks of up to 1024 threads
// (one per CU: they hold all its registers), all 22 LDS arrays within the
// device's limit.  Whole rows (tx = W) whenever a band of at least one row
// fits (W <= 341), with the tallest band that fits: rows then start on cache
// lines and halo rows fall on whole waves, which skip the dual half.  Wider
// rows run in tiles of at most 128 columns plus a halo column on each side.
// Measured on MI355X at 256^3: whole rows, band 2, 803 us per launch; 64- and
// 128-column tiles with bands of 4 to 13 rows, 1027-1385 us (their rows are
// 264- or 520-B pieces off cache-line boundaries, and every wave then holds
// owned lanes and runs the dual half).  The z chunk minimises rounds of
// resident blocks x (planes + 1).  The overrides tgvk_fused = 0 disable the
// fused sweep, tgvk_fused_tx, _ty and _zchunk set the plan.
bool tgvk_fused_plan(const Geometry& g, int levels, int device, const PlanOverrides& ov,
                     TgvkFusedPlan* plan) {
    // 32-bit offsets only (volumes under 2^30 voxels): the 64-bit form spills
    if (levels != 1 || needs_index64(g, ov)) return false;
    if (ov.tgvk_fused == 0) return false;
    int max_lds = 0, cus = 0;
    if (hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, device) !=
            hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess)
        return false;
    const int want_tx = std::max(0, ov.tgvk_fused_tx), want_ty = std::max(0, ov.tgvk_fused_ty);
    const int want_zc = std::max(0, ov.tgvk_fused_zchunk);
    int tx = g.w;
    if (want_tx > 0) tx = std::min(want_tx, g.w);
    else if (3 * g.w > KFUSED_MAX_THREADS) tx = (g.w + (g.w + 127) / 128 - 1) / ((g.w + 127) / 128);
    const int txh = tx < g.w ? tx + 2 : g.w;
    int ty = std::min(KFUSED_MAX_THREADS / txh - 2, g.h);
    while (ty >= 1 && size_t(N_KROWS) * (ty + 2) * txh * sizeof(float) > size_t(max_lds)) --ty;
    if (want_ty > 0) {
        if (want_ty > ty) return false;   // does not fit
        ty = want_ty;
    }
    if (ty < 1) return false;
    const long nb = long((g.w + tx - 1) / tx) * ((g.h + ty - 1) / ty);
    int best_zc = g.d;
    long best = -1;
    for (int nc = 1; nc <= g.d; ++nc) {
        const int zc = (g.d + nc - 1) / nc;
        const long rounds = (nb * ((g.d + zc - 1) / zc) + cus - 1) / cus;
        const long cost = rounds * (zc + 1);
        if (best < 0 || cost < best) { best = cost; best_zc = zc; }
    }
    if (want_zc > 0) best_zc = std::min(want_zc, g.d);
    *plan = TgvkFusedPlan{tx, ty, best_zc};
    return true;
}

hipError_t launch_tgvk_fused(hipStream_t s, const float* const* src, float* const* dst,
                             const float* f, const uint8_t* active, Geometry g, float tau,
                             float sigma, float lambda, float theta, const float* alpha,
                             int planar, const TgvkFusedPlan& plan) {
    KFusedFields ff;
    for (int k = 0; k < N_KF; ++k) ff.src[k] = src[k];
    ff.dt = reinterpret_cast<const char*>(dst[0]) - reinterpret_cast<const char*>(src[0]);
    ff.dl = reinterpret_cast<const char*>(dst[KF_WX]) - reinterpret_cast<const char*>(src[KF_WX]);
    if (g.n() * 4 >= (size_t(1) << 32)) return hipErrorInvalidValue;   // see tgvk_fused_plan
    if (active)
        return tgvk_fused_launch<uint32_t, 1>(s, ff, f, active, g, tau, sigma, lambda, theta, alpha,
                                              planar, plan);
    return tgvk_fused_launch<uint32_t, 0>(s, ff, f, active, g, tau, sigma, lambda, theta, alpha,
                                          planar, plan);
}

hipError_t launch_tgvk_dual(hipStream_t s, float* const* fields_dev, float* const* levels_dev,
                            const float* alpha_dev, const uint8_t* active, Geometry g, int levels,
                            float sigma, int planar, const PlanOverrides& ov) {
    int zc = 1;
    const int nb = tgvk_blocks(g, ov, &zc);
    // fixed-level kernels address with 32-bit byte offsets
    switch (g.n() <= (size_t(1) << 29) && ov.index64 <= 0 ? levels : 0) {
    case 1: hipLaunchKernelGGL(tgvk_dual_fixed<1>, dim3(nb), dim3(KT), 0, s, fields_dev,
                               levels_dev, alpha_dev, active, g, sigma, planar, zc); break;
    case 2: hipLaunchKernelGGL(tgvk_dual_fixed<2>, dim3(nb), dim3(KT), 0, s, fields_dev,
                               levels_dev, alpha_dev, active, g, sigma, planar, zc); break;
    case 3: hipLaunchKernelGGL(tgvk_dual_fixed<3>, dim3(nb), dim3(KT), 0, s, fields_dev,
                               levels_dev, alpha_dev, active, g, sigma, planar, zc); break;
    default:
        hipLaunchKernelGGL(tgvk_dual_kernel, dim3(nb), dim3(KT), 0, s, fields_dev, levels_dev,
                           alpha_dev, active, g, levels, sigma, planar, zc);
    }
    return hipGetLastError();
}

hipError_t launch_tgvk_primal(hipStream_t s, float* const* fields_dev, float* const* levels_dev,
                              const float* f, const uint8_t* active, Geometry g, int levels,
                              float tau, float lambda, float theta, const PlanOverrides& ov) {
    int zc = 1;
    const int nb = tgvk_blocks(g, ov, &zc);
    // fixed-level kernels address with 32-bit byte offsets
    switch (g.n() <= (size_t(1) << 29) && ov.index64 <= 0 ? levels : 0) {
    case 1: hipLaunchKernelGGL(tgvk_primal_fixed<1>, dim3(nb), dim3(KT), 0, s, fields_dev,
                               levels_dev, f, active, g, tau, lambda, theta, zc); break;
    case 2: hipLaunchKernelGGL(tgvk_primal_fixed<2>, dim3(nb), dim3(KT), 0, s, fields_dev,
                               levels_dev, f, active, g, tau, lambda, theta, zc); break;
    case 3: hipLaunchKernelGGL(tgvk_primal_fixed<3>, dim3(nb), dim3(KT), 0, s, fields_dev,
                               levels_dev, f, active, g, tau, lambda, theta, zc); break;
    default:
        hipLaunchKernelGGL(tgvk_primal_kernel, dim3(nb), dim3(KT), 0, s, fields_dev, levels_dev, f,
                           active, g, levels, tau, lambda, theta, zc);
    }
    return hipGetLastError();
}

}  // namespace amb
