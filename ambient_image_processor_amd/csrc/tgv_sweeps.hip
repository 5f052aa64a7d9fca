// Fused TGV2 primal-dual sweeps for gfx950.
//
// One reference iteration (tgv_deshade/tgv2_l1_deshade.cu:169-264: 37 launches,
// 10 device syncs, ~139 FP32 array passes per voxel) becomes two kernels:
//
//   dual sweep   (p, q)   <- p, q, u_bar, v_bar
//       p = proj_a1(p + sigma (D+ u_bar - v_bar))        tgv_cuda/tgv2_common.cu:96-131
//       q = proj_a0(q + sigma eps_B(v_bar))              tgv_cuda/tgv2_common.cu:174-276
//   primal sweep (u, u_bar, v, v_bar) <- u, f, v, p, q
//       u = prox(u - tau (B p))                          tgv_cuda/tgv2_l1.cu:27-60, tgv2_l2.cu:27-56
//       u_bar = u + theta (u - u_old)
//       v = v - tau (div2(q) - p), v_bar = v + theta (v - v_old)   tgv_cuda/tgv2_common.cu:279-354
//
// The dual sweep reads only the old primal fields and the primal sweep only
// the new dual fields, so neither sweep writes anything it (or a neighbouring
// tile) reads: tiles and z-chunks are independent and each sweep is a single
// launch.  Per voxel-iteration: dual 13 reads + 9 writes, primal 14 reads + 8
// writes = 44 element accesses (176 B FP32, 90 B with BF16 fields).
//
// Tiling: a block of TX x TY threads owns a TX x TY column tile and marches
// along z over `zchunk` planes.  x/y neighbours come from an LDS copy of the
// current plane (with a one-voxel halo ring loaded by the edge threads); z
// neighbours live in registers (previous/next plane).  The next plane's loads
// are issued before the current plane's arithmetic.  Blocks are remapped so
// that consecutive tiles run on the same XCD and share its L2 for halo rows.
//
// Per-voxel arithmetic keeps the reference's FP32 operation order, with fmaf
// where the reference writes fmaf or nvcc contracts a multiply-add, and
// correctly rounded division and sqrt (this file is built with
// -ffp-contract=off).
#include "amb_internal.hpp"
#include "div_rn.hpp"

#include <map>
#include <mutex>
#include <type_traits>
#include <utility>

#include <algorithm>
#include <cstdlib>

namespace amb {
namespace {

// Block id -> tile id such that each XCD (blocks are dealt round-robin to the
// 8 XCDs) receives a contiguous range of tiles.
__device__ __forceinline__ int xcd_tile(int bid, int nblocks) {
    if ((nblocks & 7) != 0) return bid;
    return (bid & 7) * (nblocks >> 3) + (bid >> 3);
}

template <int TX, int TY>
struct TileCoord {
    int x0, y0, z0, z1, tx, ty, x, y;
    __device__ __forceinline__ TileCoord(const Geometry& g, int zchunk) {
        const int nbx = (g.w + TX - 1) / TX, nby = (g.h + TY - 1) / TY;
        const int t = xcd_tile(blockIdx.x, gridDim.x);
        const int bx = t % nbx, r = t / nbx, by = r % nby, bz = r / nby;
        x0 = bx * TX; y0 = by * TY;
        z0 = bz * zchunk; z1 = min(z0 + zchunk, g.d);
        tx = threadIdx.x % TX; ty = threadIdx.x / TX;
        x = x0 + tx; y = y0 + ty;
    }
};

// Reference "backward difference" (tgv_cuda/tgv_common.cu:133-197): -a[i] + a[i-1], -a[0] at 0.
__device__ __forceinline__ float bwd(float center, float prev, bool has_prev) {
    return has_prev ? -center + prev : -center;
}
// Reference forward difference (tgv_common.cu:66-130): a[i+1] - a[i], 0 at the last index.
__device__ __forceinline__ float fwd(float center, float next, bool has_next) {
    return has_next ? next - center : 0.0f;
}

// ---------------------------------------------------------------------------
// Dual sweep
// ---------------------------------------------------------------------------
template <class S, class I, int TX, int TY, int MK>
__global__ __launch_bounds__(TX * TY) void dual_sweep_kernel(
    const uint8_t* __restrict__ act, typename S::T* __restrict__ ubar, typename S::T* __restrict__ vbx,
    typename S::T* __restrict__ vby, typename S::T* __restrict__ vbz,
    typename S::T* __restrict__ px_, typename S::T* __restrict__ py_,
    typename S::T* __restrict__ pz_, typename S::T* __restrict__ qxx_,
    typename S::T* __restrict__ qyy_, typename S::T* __restrict__ qzz_,
    typename S::T* __restrict__ qxy_, typename S::T* __restrict__ qxz_,
    typename S::T* __restrict__ qyz_, Geometry g, float sigma, float alpha0, float alpha1,
    int zchunk, int planar) {
    // u_bar plane with +x/+y halo, v_bar planes with -x/-y halo; centre at [ty+1][tx+1].
    __shared__ float sU[TY + 2][TX + 2];
    __shared__ float sV[3][TY + 2][TX + 2];

    const TileCoord<TX, TY> c(g, zchunk);
    const int W = g.w, H = g.h, D = g.d;
    const I plane = I(W) * I(H);
    const bool in = (c.x < W) && (c.y < H);
    const I col = in ? I(c.y) * I(W) + I(c.x) : I(0);
    const bool hx = (c.tx == 0) && (c.y < H);    // loads the x halo of its row
    const bool hy = (c.ty == 0) && (c.x < W);    // loads the y halo of its column
    const bool has_xp = c.x0 + TX < W, has_xm = c.x0 > 0;
    const bool has_yp = c.y0 + TY < H, has_ym = c.y0 > 0;
    const I hx_p = has_xp && hx ? I(c.y) * I(W) + I(c.x0 + TX) : I(0);
    const I hx_m = has_xm && hx ? I(c.y) * I(W) + I(c.x0 - 1) : I(0);
    const I hy_p = has_yp && hy ? I(c.y0 + TY) * I(W) + I(c.x) : I(0);
    const I hy_m = has_ym && hy ? I(c.y0 - 1) * I(W) + I(c.x) : I(0);

    typename S::T* const vb[3] = {vbx, vby, vbz};
    typename S::T* const pf[3] = {px_, py_, pz_};
    typename S::T* const qf[6] = {qxx_, qyy_, qzz_, qxy_, qxz_, qyz_};

    // z-pipeline registers: u_bar (z, z+1), v_bar (z-1, z), p and q (z)
    float ub_c = 0.f, ub_n = 0.f, vb_p[3] = {0.f, 0.f, 0.f}, vb_c[3] = {0.f, 0.f, 0.f};
    float pc[3] = {0.f, 0.f, 0.f}, qc[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    int z = c.z0;
    if (in) {
        const I i0 = I(z) * plane + col;
        ub_c = S::ld(ubar, i0);
        if (z + 1 < D) ub_n = S::ld(ubar, I(i0 + plane));
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            vb_c[k] = S::ld(vb[k], i0);
            if (z > 0) vb_p[k] = S::ld(vb[k], I(i0 - plane));
            pc[k] = S::ld(pf[k], i0);
        }
#pragma unroll
        for (int k = 0; k < 6; ++k) qc[k] = S::ld(qf[k], i0);
    }

    for (; z < c.z1; ++z) {
        const I zo = I(z) * plane;
        const I i = zo + col;
        // stage this plane in LDS
        sU[c.ty + 1][c.tx + 1] = ub_c;
#pragma unroll
        for (int k = 0; k < 3; ++k) sV[k][c.ty + 1][c.tx + 1] = vb_c[k];
        if (hx) {
            sU[c.ty + 1][TX + 1] = has_xp ? S::ld(ubar, I(zo + hx_p)) : 0.f;
#pragma unroll
            for (int k = 0; k < 3; ++k) sV[k][c.ty + 1][0] = has_xm ? S::ld(vb[k], I(zo + hx_m)) : 0.f;
        }
        if (hy) {
            sU[TY + 1][c.tx + 1] = has_yp ? S::ld(ubar, I(zo + hy_p)) : 0.f;
#pragma unroll
            for (int k = 0; k < 3; ++k) sV[k][0][c.tx + 1] = has_ym ? S::ld(vb[k], I(zo + hy_m)) : 0.f;
        }
        __syncthreads();

        // prefetch the next plane
        float ub_nn = 0.f, vb_n[3] = {0.f, 0.f, 0.f}, pn[3] = {0.f, 0.f, 0.f};
        float qn[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
        if (in && z + 1 < c.z1) {
            const I j = i + plane;
            if (z + 2 < D) ub_nn = S::ld(ubar, I(j + plane));
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                vb_n[k] = S::ld(vb[k], j);
                pn[k] = S::ld(pf[k], j);
            }
#pragma unroll
            for (int k = 0; k < 6; ++k) qn[k] = S::ld(qf[k], j);
        }

        if (in && (!MK || *elem(act, i))) {
            // D+ u_bar
            const float dx = fwd(ub_c, sU[c.ty + 1][c.tx + 2], c.x < W - 1);
            const float dy = fwd(ub_c, sU[c.ty + 2][c.tx + 1], c.y < H - 1);
            const float dz = fwd(ub_c, ub_n, z < D - 1);
            // p update + projection onto |p| <= alpha1 (tgv2_common.cu:118-128)
            const float ax = fmaf(sigma, dx - vb_c[0], pc[0]);
            const float ay = fmaf(sigma, dy - vb_c[1], pc[1]);
            const float az = fmaf(sigma, dz - vb_c[2], pc[2]);
            float n2 = ax * ax;
            n2 = fmaf(ay, ay, n2);
            n2 = fmaf(az, az, n2);
            // p / max(1, |p|/alpha1): the division is skipped when the max is 1 (x / 1 == x)
            const float tp = sqrtf(n2) / alpha1;
            const bool sp = tp > 1.0f;
            S::st(px_, i, sp ? ax / tp : ax);
            S::st(py_, i, sp ? ay / tp : ay);
            S::st(pz_, i, sp ? az / tp : az);

            // eps_B(v_bar) (tgv2_common.cu:174-219): B_a applied to v_bar_c
            float bxv[3], byv[3], bzv[3];
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                bxv[k] = bwd(vb_c[k], sV[k][c.ty + 1][c.tx], c.x > 0);
                byv[k] = bwd(vb_c[k], sV[k][c.ty][c.tx + 1], c.y > 0);
                bzv[k] = planar ? 0.0f : bwd(vb_c[k], vb_p[k], z > 0);
            }
            const float exx = bxv[0], eyy = byv[1], ezz = bzv[2];
            const float exy = (bxv[1] + byv[0]) * 0.5f;
            const float exz = (bxv[2] + bzv[0]) * 0.5f;
            const float eyz = (byv[2] + bzv[1]) * 0.5f;
            // q update + projection onto |q|_w <= alpha0 (tgv2_common.cu:246-270)
            const float bxx = fmaf(sigma, exx, qc[0]);
            const float byy = fmaf(sigma, eyy, qc[1]);
            const float bxy = fmaf(sigma, exy, qc[3]);
            const float bzz = fmaf(sigma, ezz, qc[2]);
            const float bxz = fmaf(sigma, exz, qc[4]);
            const float byz = fmaf(sigma, eyz, qc[5]);
            float m2 = bxx * bxx;
            m2 = fmaf(byy, byy, m2);
            m2 = fmaf(2.0f * bxy, bxy, m2);
            m2 = fmaf(bzz, bzz, m2);
            m2 = fmaf(2.0f * bxz, bxz, m2);
            m2 = fmaf(2.0f * byz, byz, m2);
            const float tq = sqrtf(m2) / alpha0;
            const bool sq = tq > 1.0f;
            S::st(qxx_, i, sq ? bxx / tq : bxx);
            S::st(qyy_, i, sq ? byy / tq : byy);
            S::st(qxy_, i, sq ? bxy / tq : bxy);
            S::st(qzz_, i, sq ? bzz / tq : bzz);
            S::st(qxz_, i, sq ? bxz / tq : bxz);
            S::st(qyz_, i, sq ? byz / tq : byz);
        }
        __syncthreads();
        ub_c = ub_n; ub_n = ub_nn;
#pragma unroll
        for (int k = 0; k < 3; ++k) { vb_p[k] = vb_c[k]; vb_c[k] = vb_n[k]; pc[k] = pn[k]; }
#pragma unroll
        for (int k = 0; k < 6; ++k) qc[k] = qn[k];
    }
}

// ---------------------------------------------------------------------------
// Primal sweep
// ---------------------------------------------------------------------------
// LDS slots: fields needing an x or y neighbour.  -x/-y halo: p_x, p_y.
// +x: q_xx, q_xy, q_xz.  +y: q_xy, q_yy, q_yz.
enum { L_PX = 0, L_PY = 1, L_QXX = 2, L_QXY = 3, L_QXZ = 4, L_QYY = 5, L_QYZ = 6, L_N = 7 };

template <class S, class I, int TX, int TY, int L2, int MK>
__global__ __launch_bounds__(TX * TY) void primal_sweep_kernel(
    const uint8_t* __restrict__ act, typename S::T* __restrict__ u_, typename S::T* __restrict__ ubar_,
    typename S::T* __restrict__ vx_, typename S::T* __restrict__ vy_,
    typename S::T* __restrict__ vz_, typename S::T* __restrict__ vbx_,
    typename S::T* __restrict__ vby_, typename S::T* __restrict__ vbz_,
    const typename S::T* __restrict__ px_, const typename S::T* __restrict__ py_,
    const typename S::T* __restrict__ pz_, const typename S::T* __restrict__ qxx_,
    const typename S::T* __restrict__ qyy_, const typename S::T* __restrict__ qzz_,
    const typename S::T* __restrict__ qxy_, const typename S::T* __restrict__ qxz_,
    const typename S::T* __restrict__ qyz_, const float* __restrict__ f_, Geometry g,
    float tau, float lambda, float theta, int zchunk) {
    __shared__ float sL[L_N][TY + 2][TX + 2];

    const TileCoord<TX, TY> c(g, zchunk);
    const int W = g.w, H = g.h, D = g.d;
    const I plane = I(W) * I(H);
    const bool in = (c.x < W) && (c.y < H);
    const I col = in ? I(c.y) * I(W) + I(c.x) : I(0);
    const bool hx = (c.tx == 0) && (c.y < H);
    const bool hy = (c.ty == 0) && (c.x < W);
    const bool has_xp = c.x0 + TX < W, has_xm = c.x0 > 0;
    const bool has_yp = c.y0 + TY < H, has_ym = c.y0 > 0;
    const I hx_p = has_xp && hx ? I(c.y) * I(W) + I(c.x0 + TX) : I(0);
    const I hx_m = has_xm && hx ? I(c.y) * I(W) + I(c.x0 - 1) : I(0);
    const I hy_p = has_yp && hy ? I(c.y0 + TY) * I(W) + I(c.x) : I(0);
    const I hy_m = has_ym && hy ? I(c.y0 - 1) * I(W) + I(c.x) : I(0);
    // depth == 1 drops B_z p_z in the L2 update (tgv2_l2.cu:39-42); L1 always keeps it.
    const bool use_bz = !L2 || D > 1;
    const float tl = tau * lambda;

    // z-pipeline registers: p_z (z-1, z), q_zz/q_xz/q_yz (z, z+1), pointwise (z)
    float pz_p = 0.f, pz_c = 0.f;
    float qzz_c = 0.f, qxz_c = 0.f, qyz_c = 0.f, qzz_n = 0.f, qxz_n = 0.f, qyz_n = 0.f;
    float px_c = 0.f, py_c = 0.f, qxx_c = 0.f, qxy_c = 0.f, qyy_c = 0.f;
    float u_c = 0.f, f_c = 0.f, v_c[3] = {0.f, 0.f, 0.f};
    int z = c.z0;
    if (in) {
        const I i0 = I(z) * plane + col;
        if (z > 0) pz_p = S::ld(pz_, I(i0 - plane));
        pz_c = S::ld(pz_, i0);
        qzz_c = S::ld(qzz_, i0); qxz_c = S::ld(qxz_, i0); qyz_c = S::ld(qyz_, i0);
        if (z + 1 < D) {
            const I j = i0 + plane;
            qzz_n = S::ld(qzz_, j); qxz_n = S::ld(qxz_, j); qyz_n = S::ld(qyz_, j);
        }
        px_c = S::ld(px_, i0); py_c = S::ld(py_, i0);
        qxx_c = S::ld(qxx_, i0); qxy_c = S::ld(qxy_, i0); qyy_c = S::ld(qyy_, i0);
        u_c = S::ld(u_, i0); f_c = *elem(f_, i0);
        v_c[0] = S::ld(vx_, i0); v_c[1] = S::ld(vy_, i0); v_c[2] = S::ld(vz_, i0);
    }

    for (; z < c.z1; ++z) {
        const I zo = I(z) * plane;
        const I i = zo + col;
        const int ly = c.ty + 1, lx = c.tx + 1;
        sL[L_PX][ly][lx] = px_c;
        sL[L_PY][ly][lx] = py_c;
        sL[L_QXX][ly][lx] = qxx_c;
        sL[L_QXY][ly][lx] = qxy_c;
        sL[L_QXZ][ly][lx] = qxz_c;
        sL[L_QYY][ly][lx] = qyy_c;
        sL[L_QYZ][ly][lx] = qyz_c;
        if (hx) {
            sL[L_PX][ly][0] = has_xm ? S::ld(px_, I(zo + hx_m)) : 0.f;
            sL[L_QXX][ly][TX + 1] = has_xp ? S::ld(qxx_, I(zo + hx_p)) : 0.f;
            sL[L_QXY][ly][TX + 1] = has_xp ? S::ld(qxy_, I(zo + hx_p)) : 0.f;
            sL[L_QXZ][ly][TX + 1] = has_xp ? S::ld(qxz_, I(zo + hx_p)) : 0.f;
        }
        if (hy) {
            sL[L_PY][0][lx] = has_ym ? S::ld(py_, I(zo + hy_m)) : 0.f;
            sL[L_QXY][TY + 1][lx] = has_yp ? S::ld(qxy_, I(zo + hy_p)) : 0.f;
            sL[L_QYY][TY + 1][lx] = has_yp ? S::ld(qyy_, I(zo + hy_p)) : 0.f;
            sL[L_QYZ][TY + 1][lx] = has_yp ? S::ld(qyz_, I(zo + hy_p)) : 0.f;
        }
        __syncthreads();

        // prefetch the next plane
        float n_pz = 0.f, n_qzz = 0.f, n_qxz = 0.f, n_qyz = 0.f, n_px = 0.f, n_py = 0.f,
              n_qxx = 0.f, n_qxy = 0.f, n_qyy = 0.f, n_u = 0.f, n_f = 0.f;
        float n_v[3] = {0.f, 0.f, 0.f};
        if (in && z + 1 < c.z1) {
            const I j = i + plane;
            n_pz = S::ld(pz_, j);
            if (z + 2 < D) {
                const I j2 = j + plane;
                n_qzz = S::ld(qzz_, j2); n_qxz = S::ld(qxz_, j2); n_qyz = S::ld(qyz_, j2);
            }
            n_px = S::ld(px_, j); n_py = S::ld(py_, j);
            n_qxx = S::ld(qxx_, j); n_qxy = S::ld(qxy_, j); n_qyy = S::ld(qyy_, j);
            n_u = S::ld(u_, j); n_f = *elem(f_, j);
            n_v[0] = S::ld(vx_, j); n_v[1] = S::ld(vy_, j); n_v[2] = S::ld(vz_, j);
        }

        if (in && (!MK || *elem(act, i))) {
            // B p  (tgv2_l1_deshade.cu:186-193)
            const float bx = bwd(px_c, sL[L_PX][ly][lx - 1], c.x > 0);
            const float by = bwd(py_c, sL[L_PY][ly - 1][lx], c.y > 0);
            const float bz = bwd(pz_c, pz_p, z > 0);
            const float div = use_bz ? (bx + by) + bz : bx + by;
            float un = fmaf(-tau, div, u_c);
            if (L2) {
                un = fmaf(tl, f_c, un) / (1.0f + tl);           // tgv2_l2.cu:44
            } else {
                const float umf = un - f_c;                     // tgv2_l1.cu:41-48
                if (umf > tl) un = un - tl;
                else if (umf < -tl) un = un + tl;
                else un = f_c;
            }
            S::st(u_, i, un);
            S::st(ubar_, i, fmaf(theta, un - u_c, un));

            // div2(q) with forward differences (tgv2_common.cu:291-318)
            const bool xn = c.x < W - 1, yn = c.y < H - 1, zn = z < D - 1;
            const float dxx = fwd(qxx_c, sL[L_QXX][ly][lx + 1], xn);
            const float dyxy = fwd(qxy_c, sL[L_QXY][ly + 1][lx], yn);
            const float dzxz = fwd(qxz_c, qxz_n, zn);
            const float dyy = fwd(qyy_c, sL[L_QYY][ly + 1][lx], yn);
            const float dxxy = fwd(qxy_c, sL[L_QXY][ly][lx + 1], xn);
            const float dzyz = fwd(qyz_c, qyz_n, zn);
            const float dzz = fwd(qzz_c, qzz_n, zn);
            const float dxxz = fwd(qxz_c, sL[L_QXZ][ly][lx + 1], xn);
            const float dyyz = fwd(qyz_c, sL[L_QYZ][ly + 1][lx], yn);
            const float w0 = (dxx + dyxy) + dzxz;
            const float w1 = (dyy + dxxy) + dzyz;
            const float w2 = (dzz + dxxz) + dyyz;
            // v update + over-relaxation (tgv2_common.cu:338-347)
            const float vn0 = fmaf(-tau, w0 - px_c, v_c[0]);
            const float vn1 = fmaf(-tau, w1 - py_c, v_c[1]);
            const float vn2 = fmaf(-tau, w2 - pz_c, v_c[2]);
            S::st(vx_, i, vn0);
            S::st(vy_, i, vn1);
            S::st(vz_, i, vn2);
            S::st(vbx_, i, fmaf(theta, vn0 - v_c[0], vn0));
            S::st(vby_, i, fmaf(theta, vn1 - v_c[1], vn1));
            S::st(vbz_, i, fmaf(theta, vn2 - v_c[2], vn2));
        }
        __syncthreads();
        pz_p = pz_c; pz_c = n_pz;
        qzz_c = qzz_n; qxz_c = qxz_n; qyz_c = qyz_n;
        qzz_n = n_qzz; qxz_n = n_qxz; qyz_n = n_qyz;
        px_c = n_px; py_c = n_py; qxx_c = n_qxx; qxy_c = n_qxy; qyy_c = n_qyy;
        u_c = n_u; f_c = n_f;
#pragma unroll
        for (int k = 0; k < 3; ++k) v_c[k] = n_v[k];
    }
}

// ---------------------------------------------------------------------------
// Vector sweeps: each thread owns VX consecutive x voxels (W % VX == 0), so
// every field moves as one VX-wide load/store per thread (16 B for FP32, 8 B
// for BF16 at VX = 4).  y neighbours come from full LDS rows (+/- one halo row
// loaded by the ty == 0 threads); x neighbours inside the vector are in
// registers, across threads one edge element per field goes through LDS.
// PF = 1 issues the next plane's loads before the current plane's arithmetic.
// ---------------------------------------------------------------------------
template <class S, int VX>
struct VecIO;

template <>
struct VecIO<StoreF32, 4> {
    template <class I>
    static __device__ __forceinline__ void ld(const float* p, I i, float (&o)[4]) {
        const float4 t = *reinterpret_cast<const float4*>(elem(p, i));
        o[0] = t.x; o[1] = t.y; o[2] = t.z; o[3] = t.w;
    }
    template <class I>
    static __device__ __forceinline__ void st(float* p, I i, const float (&v)[4]) {
        // written once per launch and not read back in it: a streaming store
        typedef float f4 __attribute__((ext_vector_type(4)));
        const f4 t = {v[0], v[1], v[2], v[3]};
        __builtin_nontemporal_store(t, reinterpret_cast<f4*>(elem(p, i)));
    }
};

template <>
struct VecIO<StoreBF16, 4> {
    template <class I>
    static __device__ __forceinline__ void ld(const uint16_t* p, I i, float (&o)[4]) {
        const uint2 t = *reinterpret_cast<const uint2*>(elem(p, i));
        o[0] = __uint_as_float(t.x << 16); o[1] = __uint_as_float(t.x & 0xffff0000u);
        o[2] = __uint_as_float(t.y << 16); o[3] = __uint_as_float(t.y & 0xffff0000u);
    }
    template <class I>
    static __device__ __forceinline__ void st(uint16_t* p, I i, const float (&v)[4]) {
        uint2 t;
        t.x = StoreBF16::rne2(v[0], v[1]);
        t.y = StoreBF16::rne2(v[2], v[3]);
        *reinterpret_cast<uint2*>(elem(p, i)) = t;
    }
};

template <class S>
struct VecIO<S, 1> {   // one voxel per lane (odd widths)
    template <class I>
    static __device__ __forceinline__ void ld(const typename S::T* p, I i, float (&o)[1]) {
        o[0] = S::ld(p, i);
    }
    template <class I>
    static __device__ __forceinline__ void st(typename S::T* p, I i, const float (&v)[1]) {
        S::st(p, i, v[0]);
    }
};

template <>
struct VecIO<StoreF32, 2> {
    template <class I>
    static __device__ __forceinline__ void ld(const float* p, I i, float (&o)[2]) {
        const float2 t = *reinterpret_cast<const float2*>(elem(p, i));
        o[0] = t.x; o[1] = t.y;
    }
    template <class I>
    static __device__ __forceinline__ void st(float* p, I i, const float (&v)[2]) {
        typedef float f2 __attribute__((ext_vector_type(2)));
        const f2 t = {v[0], v[1]};
        __builtin_nontemporal_store(t, reinterpret_cast<f2*>(elem(p, i)));
    }
};

template <>
struct VecIO<StoreBF16, 2> {
    template <class I>
    static __device__ __forceinline__ void ld(const uint16_t* p, I i, float (&o)[2]) {
        const uint32_t t = *reinterpret_cast<const uint32_t*>(elem(p, i));
        o[0] = __uint_as_float(t << 16); o[1] = __uint_as_float(t & 0xffff0000u);
    }
    template <class I>
    static __device__ __forceinline__ void st(uint16_t* p, I i, const float (&v)[2]) {
        *reinterpret_cast<uint32_t*>(elem(p, i)) = StoreBF16::rne2(v[0], v[1]);
    }
};

// Active bytes of VX consecutive voxels (byte e of the result = voxel e).
template <int VX, class I>
__device__ __forceinline__ uint32_t load_active(const uint8_t* act, I i) {
    if constexpr (VX == 4) return *reinterpret_cast<const uint32_t*>(elem(act, i));
    else if constexpr (VX == 2) return uint32_t(*reinterpret_cast<const uint16_t*>(elem(act, i)));
    else return uint32_t(*elem(act, i));
}

template <int VX>
__device__ __forceinline__ void zero(float (&a)[VX]) {
#pragma unroll
    for (int e = 0; e < VX; ++e) a[e] = 0.f;
}
template <int VX>
__device__ __forceinline__ void copyv(float (&d)[VX], const float (&s)[VX]) {
#pragma unroll
    for (int e = 0; e < VX; ++e) d[e] = s[e];
}
template <int VX>
__device__ __forceinline__ void lds_put(float* row, const float (&v)[VX]) {
#pragma unroll
    for (int e = 0; e < VX; ++e) row[e] = v[e];
}
template <int VX>
__device__ __forceinline__ void lds_get(const float* row, float (&v)[VX]) {
#pragma unroll
    for (int e = 0; e < VX; ++e) v[e] = row[e];
}

// Vector store; with MK only the elements whose active byte is set are written.
template <class S, int VX, int MK, class I>
__device__ __forceinline__ void vstore(typename S::T* p, I i, const float (&v)[VX],
                                       const uint32_t actv) {
    if (!MK) {
        VecIO<S, VX>::st(p, i, v);
    } else {
#pragma unroll
        for (int e = 0; e < VX; ++e)
            if ((actv >> (8 * e)) & 0xffu) S::st(p, I(i + e), v[e]);
    }
}

template <int TX, int TY, int VX>
struct TileCoordV {
    int x0, y0, z0, z1, tx, ty, xb, y;
    __device__ __forceinline__ TileCoordV(const Geometry& g, int zchunk) {
        constexpr int RW = TX * VX;
        const int nbx = (g.w + RW - 1) / RW, nby = (g.h + TY - 1) / TY;
        const int t = xcd_tile(blockIdx.x, gridDim.x);
        const int bx = t % nbx, r = t / nbx, by = r % nby, bz = r / nby;
        x0 = bx * RW; y0 = by * TY;
        z0 = bz * zchunk; z1 = min(z0 + zchunk, g.d);
        tx = threadIdx.x % TX; ty = threadIdx.x / TX;
        xb = x0 + tx * VX; y = y0 + ty;
    }
};

template <int VX>
struct DualPlane {
    float ubn[VX];        // u_bar at z + 1
    float vb[3][VX];      // v_bar at z
    float p[3][VX];
    float q[6][VX];       // xx, yy, zz, xy, xz, yz
};

template <class S, class I, int TX, int TY, int VX, int PF, int MK>
__global__ __launch_bounds__(TX * TY) void dual_sweep_vec_kernel(
    const uint8_t* __restrict__ act, typename S::T* __restrict__ ubar, typename S::T* __restrict__ vbx,
    typename S::T* __restrict__ vby, typename S::T* __restrict__ vbz,
    typename S::T* __restrict__ px_, typename S::T* __restrict__ py_,
    typename S::T* __restrict__ pz_, typename S::T* __restrict__ qxx_,
    typename S::T* __restrict__ qyy_, typename S::T* __restrict__ qzz_,
    typename S::T* __restrict__ qxy_, typename S::T* __restrict__ qxz_,
    typename S::T* __restrict__ qyz_, Geometry g, float sigma, float alpha0, float alpha1,
    int zchunk, int planar) {
    typedef VecIO<S, VX> V;
    constexpr int RW = TX * VX;
    __shared__ __align__(16) float sY[4][TY + 2][RW];   // 0: u_bar (+y halo row TY+1); 1-3: v_bar (-y halo row 0)
    __shared__ float sXp[TY][TX + 1];                   // u_bar element 0 per thread; [ty][TX] = +x halo
    __shared__ float sXm[3][TY][TX + 1];                // v_bar element VX-1 at [ty][tx+1]; [ty][0] = -x halo

    const TileCoordV<TX, TY, VX> c(g, zchunk);
    const int W = g.w, H = g.h, D = g.d;
    const I plane = I(W) * I(H);
    const bool in = (c.xb < W) && (c.y < H);
    const I col = in ? I(c.y) * I(W) + I(c.xb) : I(0);
    const bool hx = (c.tx == 0) && (c.y < H);
    const bool hy = (c.ty == 0) && (c.xb < W);
    const bool has_xp = c.x0 + RW < W, has_xm = c.x0 > 0;
    const bool has_yp = c.y0 + TY < H, has_ym = c.y0 > 0;
    const I hx_p = has_xp && hx ? I(c.y) * I(W) + I(c.x0 + RW) : I(0);
    const I hx_m = has_xm && hx ? I(c.y) * I(W) + I(c.x0 - 1) : I(0);
    const I hy_p = has_yp && hy ? I(c.y0 + TY) * I(W) + I(c.xb) : I(0);
    const I hy_m = has_ym && hy ? I(c.y0 - 1) * I(W) + I(c.xb) : I(0);
    typename S::T* const vb[3] = {vbx, vby, vbz};
    typename S::T* const pf[3] = {px_, py_, pz_};
    typename S::T* const qf[6] = {qxx_, qyy_, qzz_, qxy_, qxz_, qyz_};
    const int lx = c.tx * VX;

    auto load_plane = [&](int zz, DualPlane<VX>& P) {
        const I i = I(zz) * plane + col;
        if (zz + 1 < D) V::ld(ubar, I(i + plane), P.ubn); else zero(P.ubn);
#pragma unroll
        for (int k = 0; k < 3; ++k) { V::ld(vb[k], i, P.vb[k]); V::ld(pf[k], i, P.p[k]); }
#pragma unroll
        for (int k = 0; k < 6; ++k) V::ld(qf[k], i, P.q[k]);
    };

    float ub_c[VX], vb_p[3][VX];
    DualPlane<VX> P;
    zero(ub_c);
#pragma unroll
    for (int k = 0; k < 3; ++k) zero(vb_p[k]);
    if (in) {
        const I i0 = I(c.z0) * plane + col;
        V::ld(ubar, i0, ub_c);
        if (c.z0 > 0)
#pragma unroll
            for (int k = 0; k < 3; ++k) V::ld(vb[k], I(i0 - plane), vb_p[k]);
        if (PF) load_plane(c.z0, P);
    }

    for (int z = c.z0; z < c.z1; ++z) {
        const I zo = I(z) * plane;
        const I i = zo + col;
        if (!PF && in) load_plane(z, P);
        // stage this plane
        lds_put<VX>(&sY[0][c.ty + 1][lx], ub_c);
        sXp[c.ty][c.tx] = ub_c[0];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            lds_put<VX>(&sY[1 + k][c.ty + 1][lx], P.vb[k]);
            sXm[k][c.ty][c.tx + 1] = P.vb[k][VX - 1];
        }
        if (hx) {
            sXp[c.ty][TX] = has_xp ? S::ld(ubar, I(zo + hx_p)) : 0.f;
#pragma unroll
            for (int k = 0; k < 3; ++k) sXm[k][c.ty][0] = has_xm ? S::ld(vb[k], I(zo + hx_m)) : 0.f;
        }
        if (hy) {
            float h[VX];
            if (has_yp) V::ld(ubar, I(zo + hy_p), h); else zero(h);
            lds_put<VX>(&sY[0][TY + 1][lx], h);
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                if (has_ym) V::ld(vb[k], I(zo + hy_m), h); else zero(h);
                lds_put<VX>(&sY[1 + k][0][lx], h);
            }
        }
        __syncthreads();

        DualPlane<VX> Pn;
        if (PF && in && z + 1 < c.z1) load_plane(z + 1, Pn);

        const uint32_t actv = (MK && in) ? *reinterpret_cast<const uint32_t*>(elem(act, i)) : 0xffffffffu;
        if (in && (!MK || actv)) {
            float uyn[VX], vym[3][VX], vxm[3];
            lds_get<VX>(&sY[0][c.ty + 2][lx], uyn);
            const float uxn = sXp[c.ty][c.tx + 1];
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                lds_get<VX>(&sY[1 + k][c.ty][lx], vym[k]);
                vxm[k] = sXm[k][c.ty][c.tx];
            }
            float po[3][VX], qo[6][VX];
#pragma unroll
            for (int e = 0; e < VX; ++e) {
                const int x = c.xb + e;
                const float ub = ub_c[e];
                const float dx = fwd(ub, e < VX - 1 ? ub_c[e + 1 < VX ? e + 1 : e] : uxn, x < W - 1);
                const float dy = fwd(ub, uyn[e], c.y < H - 1);
                const float dz = fwd(ub, P.ubn[e], z < D - 1);
                const float ax = fmaf(sigma, dx - P.vb[0][e], P.p[0][e]);
                const float ay = fmaf(sigma, dy - P.vb[1][e], P.p[1][e]);
                const float az = fmaf(sigma, dz - P.vb[2][e], P.p[2][e]);
                float n2 = ax * ax;
                n2 = fmaf(ay, ay, n2);
                n2 = fmaf(az, az, n2);
                // p / max(1, |p|/alpha1): no division when the max is 1 (x / 1 == x)
                const float tp = sqrtf(n2) / alpha1;
                if (tp > 1.0f) {
                    po[0][e] = ax / tp; po[1][e] = ay / tp; po[2][e] = az / tp;
                } else {
                    po[0][e] = ax; po[1][e] = ay; po[2][e] = az;
                }

                float bxv[3], byv[3], bzv[3];
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    const float vc = P.vb[k][e];
                    bxv[k] = bwd(vc, e > 0 ? P.vb[k][e > 0 ? e - 1 : 0] : vxm[k], x > 0);
                    byv[k] = bwd(vc, vym[k][e], c.y > 0);
                    bzv[k] = planar ? 0.0f : bwd(vc, vb_p[k][e], z > 0);
                }
                const float exy = (bxv[1] + byv[0]) * 0.5f;
                const float exz = (bxv[2] + bzv[0]) * 0.5f;
                const float eyz = (byv[2] + bzv[1]) * 0.5f;
                const float bxx = fmaf(sigma, bxv[0], P.q[0][e]);
                const float byy = fmaf(sigma, byv[1], P.q[1][e]);
                const float bxy = fmaf(sigma, exy, P.q[3][e]);
                const float bzz = fmaf(sigma, bzv[2], P.q[2][e]);
                const float bxz = fmaf(sigma, exz, P.q[4][e]);
                const float byz = fmaf(sigma, eyz, P.q[5][e]);
                float m2 = bxx * bxx;
                m2 = fmaf(byy, byy, m2);
                m2 = fmaf(2.0f * bxy, bxy, m2);
                m2 = fmaf(bzz, bzz, m2);
                m2 = fmaf(2.0f * bxz, bxz, m2);
                m2 = fmaf(2.0f * byz, byz, m2);
                const float tq = sqrtf(m2) / alpha0;
                if (tq > 1.0f) {
                    qo[0][e] = bxx / tq; qo[1][e] = byy / tq; qo[2][e] = bzz / tq;
                    qo[3][e] = bxy / tq; qo[4][e] = bxz / tq; qo[5][e] = byz / tq;
                } else {
                    qo[0][e] = bxx; qo[1][e] = byy; qo[2][e] = bzz;
                    qo[3][e] = bxy; qo[4][e] = bxz; qo[5][e] = byz;
                }
            }
#pragma unroll
            for (int k = 0; k < 3; ++k) vstore<S, VX, MK>(pf[k], i, po[k], actv);
#pragma unroll
            for (int k = 0; k < 6; ++k) vstore<S, VX, MK>(qf[k], i, qo[k], actv);
        }
        __syncthreads();
        copyv(ub_c, P.ubn);
#pragma unroll
        for (int k = 0; k < 3; ++k) copyv(vb_p[k], P.vb[k]);
        if (PF) P = Pn;
    }
}

template <int VX>
struct PrimalPlane {
    float px[VX], py[VX], pz[VX], qxx[VX], qxy[VX], qyy[VX], u[VX], f[VX], v[3][VX];
    float qzzn[VX], qxzn[VX], qyzn[VX];     // at z + 1
};

template <class S, class I, int TX, int TY, int VX, int PF, int L2, int MK>
__global__ __launch_bounds__(TX * TY) void primal_sweep_vec_kernel(
    const uint8_t* __restrict__ act, typename S::T* __restrict__ u_, typename S::T* __restrict__ ubar_,
    typename S::T* __restrict__ vx_, typename S::T* __restrict__ vy_,
    typename S::T* __restrict__ vz_, typename S::T* __restrict__ vbx_,
    typename S::T* __restrict__ vby_, typename S::T* __restrict__ vbz_,
    const typename S::T* __restrict__ px_, const typename S::T* __restrict__ py_,
    const typename S::T* __restrict__ pz_, const typename S::T* __restrict__ qxx_,
    const typename S::T* __restrict__ qyy_, const typename S::T* __restrict__ qzz_,
    const typename S::T* __restrict__ qxy_, const typename S::T* __restrict__ qxz_,
    const typename S::T* __restrict__ qyz_, const float* __restrict__ f_, Geometry g,
    float tau, float lambda, float theta, int zchunk) {
    typedef VecIO<S, VX> V;
    typedef VecIO<StoreF32, VX> VF;
    constexpr int RW = TX * VX;
    // y rows: 0 p_y (-y halo row 0), 1 q_xy, 2 q_yy, 3 q_yz (+y halo row TY+1)
    __shared__ __align__(16) float sY[4][TY + 2][RW];
    __shared__ float sXp[3][TY][TX + 1];    // q_xx, q_xy, q_xz element 0; [ty][TX] = +x halo
    __shared__ float sXm[TY][TX + 1];       // p_x element VX-1 at [ty][tx+1]; [ty][0] = -x halo

    const TileCoordV<TX, TY, VX> c(g, zchunk);
    const int W = g.w, H = g.h, D = g.d;
    const I plane = I(W) * I(H);
    const bool in = (c.xb < W) && (c.y < H);
    const I col = in ? I(c.y) * I(W) + I(c.xb) : I(0);
    const bool hx = (c.tx == 0) && (c.y < H);
    const bool hy = (c.ty == 0) && (c.xb < W);
    const bool has_xp = c.x0 + RW < W, has_xm = c.x0 > 0;
    const bool has_yp = c.y0 + TY < H, has_ym = c.y0 > 0;
    const I hx_p = has_xp && hx ? I(c.y) * I(W) + I(c.x0 + RW) : I(0);
    const I hx_m = has_xm && hx ? I(c.y) * I(W) + I(c.x0 - 1) : I(0);
    const I hy_p = has_yp && hy ? I(c.y0 + TY) * I(W) + I(c.xb) : I(0);
    const I hy_m = has_ym && hy ? I(c.y0 - 1) * I(W) + I(c.xb) : I(0);
    const bool use_bz = !L2 || D > 1;
    const float tl = tau * lambda;
    const int lx = c.tx * VX;

    auto load_plane = [&](int zz, PrimalPlane<VX>& P) {
        const I i = I(zz) * plane + col;
        V::ld(px_, i, P.px); V::ld(py_, i, P.py); V::ld(pz_, i, P.pz);
        V::ld(qxx_, i, P.qxx); V::ld(qxy_, i, P.qxy); V::ld(qyy_, i, P.qyy);
        V::ld(u_, i, P.u); VF::ld(f_, i, P.f);
        V::ld(vx_, i, P.v[0]); V::ld(vy_, i, P.v[1]); V::ld(vz_, i, P.v[2]);
        if (zz + 1 < D) {
            const I j = i + plane;
            V::ld(qzz_, j, P.qzzn); V::ld(qxz_, j, P.qxzn); V::ld(qyz_, j, P.qyzn);
        } else {
            zero(P.qzzn); zero(P.qxzn); zero(P.qyzn);
        }
    };

    float pz_p[VX], qzz_c[VX], qxz_c[VX], qyz_c[VX];
    PrimalPlane<VX> P;
    zero(pz_p); zero(qzz_c); zero(qxz_c); zero(qyz_c);
    if (in) {
        const I i0 = I(c.z0) * plane + col;
        if (c.z0 > 0) V::ld(pz_, I(i0 - plane), pz_p);
        V::ld(qzz_, i0, qzz_c); V::ld(qxz_, i0, qxz_c); V::ld(qyz_, i0, qyz_c);
        if (PF) load_plane(c.z0, P);
    }

    for (int z = c.z0; z < c.z1; ++z) {
        const I zo = I(z) * plane;
        const I i = zo + col;
        if (!PF && in) load_plane(z, P);
        lds_put<VX>(&sY[0][c.ty + 1][lx], P.py);
        lds_put<VX>(&sY[1][c.ty + 1][lx], P.qxy);
        lds_put<VX>(&sY[2][c.ty + 1][lx], P.qyy);
        lds_put<VX>(&sY[3][c.ty + 1][lx], qyz_c);
        sXp[0][c.ty][c.tx] = P.qxx[0];
        sXp[1][c.ty][c.tx] = P.qxy[0];
        sXp[2][c.ty][c.tx] = qxz_c[0];
        sXm[c.ty][c.tx + 1] = P.px[VX - 1];
        if (hx) {
            sXm[c.ty][0] = has_xm ? S::ld(px_, I(zo + hx_m)) : 0.f;
            sXp[0][c.ty][TX] = has_xp ? S::ld(qxx_, I(zo + hx_p)) : 0.f;
            sXp[1][c.ty][TX] = has_xp ? S::ld(qxy_, I(zo + hx_p)) : 0.f;
            sXp[2][c.ty][TX] = has_xp ? S::ld(qxz_, I(zo + hx_p)) : 0.f;
        }
        if (hy) {
            float h[VX];
            if (has_ym) V::ld(py_, I(zo + hy_m), h); else zero(h);
            lds_put<VX>(&sY[0][0][lx], h);
            if (has_yp) V::ld(qxy_, I(zo + hy_p), h); else zero(h);
            lds_put<VX>(&sY[1][TY + 1][lx], h);
            if (has_yp) V::ld(qyy_, I(zo + hy_p), h); else zero(h);
            lds_put<VX>(&sY[2][TY + 1][lx], h);
            if (has_yp) V::ld(qyz_, I(zo + hy_p), h); else zero(h);
            lds_put<VX>(&sY[3][TY + 1][lx], h);
        }
        __syncthreads();

        PrimalPlane<VX> Pn;
        if (PF && in && z + 1 < c.z1) load_plane(z + 1, Pn);

        const uint32_t actv = (MK && in) ? *reinterpret_cast<const uint32_t*>(elem(act, i)) : 0xffffffffu;
        if (in && (!MK || actv)) {
            float pym[VX], qxyn_y[VX], qyyn_y[VX], qyzn_y[VX];
            lds_get<VX>(&sY[0][c.ty][lx], pym);
            lds_get<VX>(&sY[1][c.ty + 2][lx], qxyn_y);
            lds_get<VX>(&sY[2][c.ty + 2][lx], qyyn_y);
            lds_get<VX>(&sY[3][c.ty + 2][lx], qyzn_y);
            const float pxm = sXm[c.ty][c.tx];
            const float qxxn_x = sXp[0][c.ty][c.tx + 1];
            const float qxyn_x = sXp[1][c.ty][c.tx + 1];
            const float qxzn_x = sXp[2][c.ty][c.tx + 1];
            float uo[VX], ubo[VX], vo[3][VX], vbo[3][VX];
#pragma unroll
            for (int e = 0; e < VX; ++e) {
                const int x = c.xb + e;
                const int en = e + 1 < VX ? e + 1 : e, ep = e > 0 ? e - 1 : 0;
                const bool xl = e < VX - 1;
                const float bx = bwd(P.px[e], e > 0 ? P.px[ep] : pxm, x > 0);
                const float by = bwd(P.py[e], pym[e], c.y > 0);
                const float bz = bwd(P.pz[e], pz_p[e], z > 0);
                const float div = use_bz ? (bx + by) + bz : bx + by;
                const float uc = P.u[e], fc = P.f[e];
                float un = fmaf(-tau, div, uc);
                if (L2) {
                    un = fmaf(tl, fc, un) / (1.0f + tl);
                } else {
                    const float umf = un - fc;
                    if (umf > tl) un = un - tl;
                    else if (umf < -tl) un = un + tl;
                    else un = fc;
                }
                uo[e] = un;
                ubo[e] = fmaf(theta, un - uc, un);
                const bool xn = x < W - 1, yn = c.y < H - 1, zn = z < D - 1;
                const float dxx = fwd(P.qxx[e], xl ? P.qxx[en] : qxxn_x, xn);
                const float dyxy = fwd(P.qxy[e], qxyn_y[e], yn);
                const float dzxz = fwd(qxz_c[e], P.qxzn[e], zn);
                const float dyy = fwd(P.qyy[e], qyyn_y[e], yn);
                const float dxxy = fwd(P.qxy[e], xl ? P.qxy[en] : qxyn_x, xn);
                const float dzyz = fwd(qyz_c[e], P.qyzn[e], zn);
                const float dzz = fwd(qzz_c[e], P.qzzn[e], zn);
                const float dxxz = fwd(qxz_c[e], xl ? qxz_c[en] : qxzn_x, xn);
                const float dyyz = fwd(qyz_c[e], qyzn_y[e], yn);
                const float w0 = (dxx + dyxy) + dzxz;
                const float w1 = (dyy + dxxy) + dzyz;
                const float w2 = (dzz + dxxz) + dyyz;
                const float vn0 = fmaf(-tau, w0 - P.px[e], P.v[0][e]);
                const float vn1 = fmaf(-tau, w1 - P.py[e], P.v[1][e]);
                const float vn2 = fmaf(-tau, w2 - P.pz[e], P.v[2][e]);
                vo[0][e] = vn0; vo[1][e] = vn1; vo[2][e] = vn2;
                vbo[0][e] = fmaf(theta, vn0 - P.v[0][e], vn0);
                vbo[1][e] = fmaf(theta, vn1 - P.v[1][e], vn1);
                vbo[2][e] = fmaf(theta, vn2 - P.v[2][e], vn2);
            }
            vstore<S, VX, MK>(u_, i, uo, actv);
            vstore<S, VX, MK>(ubar_, i, ubo, actv);
            vstore<S, VX, MK>(vx_, i, vo[0], actv);
            vstore<S, VX, MK>(vy_, i, vo[1], actv);
            vstore<S, VX, MK>(vz_, i, vo[2], actv);
            vstore<S, VX, MK>(vbx_, i, vbo[0], actv);
            vstore<S, VX, MK>(vby_, i, vbo[1], actv);
            vstore<S, VX, MK>(vbz_, i, vbo[2], actv);
        }
        __syncthreads();
        copyv(pz_p, P.pz);
        copyv(qzz_c, P.qzzn); copyv(qxz_c, P.qxzn); copyv(qyz_c, P.qyzn);
        if (PF) P = Pn;
    }
}

// ---------------------------------------------------------------------------
// Fused sweep: the primal half of iteration k and the dual half of iteration
// k + 1 in one z-march (same per-voxel arithmetic as the two sweeps above).
//
// The dual update at a voxel reads u_bar / v_bar of its x, y, z neighbours,
// i.e. primal results of neighbouring voxels.  A block therefore owns whole
// x rows and a band of `ty` y rows, computes the primal half on one extra row
// above and below the band (and one extra plane before and after its z
// chunk), and runs the dual half one plane behind the primal half.  u_bar and
// v_bar never leave the CU.  Neighbouring blocks recompute the halo rows from
// the same inputs, so u, v, p, q are read from one copy of the state (src)
// and written to another (dst).  Per voxel-iteration: reads u, f, v x3, p x3,
// q x6 = 14, writes u, v x3, p x3, q x6 = 13: 27 element accesses (108 B FP32).
//
// Threads: blockDim.x = L * (ty + 2) with L = W / VX lanes per row; thread row
// r holds y = y0 - 1 + r.  Rows 1..ty are owned; row 0 only feeds v_bar
// (its u half is not needed), row ty + 1 only u_bar.  Step s computes the
// primal half at plane s and the dual half at plane s - 1 for s = z0 - 1 ..
// z1 (plane z0 - 1 needs only v_bar, plane z1 only u_bar).
//
// LDS (dynamic, floats): P region for the primal half's x/y neighbours of
// p and q at plane s, D region for the dual half's x/y neighbours of u_bar and
// v_bar at plane s - 1; rows of W floats plus one edge float per lane.
// ---------------------------------------------------------------------------
enum FusedField { G_U = 0, G_VX, G_VY, G_VZ, G_PX, G_PY, G_PZ, G_QXX, G_QYY, G_QZZ, G_QXY,
                  G_QXZ, G_QYZ, N_GFIELDS };

struct FusedFields {
    const void* src[N_GFIELDS];
    void* dst[N_GFIELDS];
    void* bars[4];   // SB = 1: u_bar, v_bar x, y, z of the primal half (else unused)
};

constexpr int FUSED_MAX_THREADS = 768;

// LDS floats of a fused block: `row` floats per row (W, or tx + 2 VX for tiles);
// the energy form keeps u and v of the finished iteration in 4 more rows.
template <int VX>
inline size_t fused_lds_floats(int row, int ty, int en = 0) {
    const int L = row / VX, RT = ty + 2;
    return size_t(en ? 12 : 8) * (size_t(RT) * row + size_t(RT) * (L + 1));
}

// Sum over the first nact lanes of a wave (lanes beyond nact do not exist in a
// partial last wave; their shuffled values are ignored).
__device__ __forceinline__ double wave_sum_n(double v, int lane, int nact) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const double t = __shfl_down(v, o, 64);
        if (lane + o < nact) v += t;
    }
    return v;
}

template <int VX>
struct FusedPlane {
    float px[VX], py[VX], pz[VX], qxx[VX], qxy[VX], qyy[VX], u[VX], f[VX], v[3][VX];
    float qnn[3][VX];     // q_zz, q_xz, q_yz at the next plane
    uint32_t act;
};

// PF = 1 loads plane s + 1 right after the first barrier of step s, so the
// loads are in flight during both halves of step s.
// TL = 1: the block owns `tx` columns (tx % VX == 0) instead of whole rows, with
// one halo lane (VX voxels) on each side: lane 0 feeds only v_bar at x0 - 1,
// the last lane only u_bar at x0 + tx (rows wider than whole-row bands allow).
// P2 = 1 (the 2-D loop, depth 1): v_z, p_z, q_zz, q_xz and q_yz stay +0 through
// every update there, so they are taken as +0 instead of loaded and are not
// stored (both copies of them must hold +0); the arithmetic is unchanged.
// P2 = 2 (the 2-D loop marched along the image's rows): the host passes the
// 1 x H x W image as an H x 1 x W volume with the fields' y and z roles
// exchanged (v_y <-> v_z, p_y <-> p_z, q_yy <-> q_zz, q_xy <-> q_xz); the
// kernel's y components are then the image's zero z components (taken as +0,
// not stored), its y backward differences of eps_B(v_bar) are 0 (the 2-D
// algebra), bands have no halo rows (height 1), and every nonzero term of
// every update keeps its place in the reference's operation order.
// EN = 1 (3-D loop, P2 = 0, not planar): the launch also reduces, in FP64, the
// energy and primal-dual gap of the state after the primal half - u and v just
// computed, p and q read - with the per-voxel formulas of the standalone
// energy pass (energy.hip): the data term, the -(D+)^T p bracket, its L1
// infeasibility and the v residual in the primal half at plane s (the p and q
// neighbours are the ones the updates read), the alpha1 and alpha0 terms and
// brackets one plane behind, with the dual half (u at z + 1 is then in
// registers; u and v of the plane in 4 more LDS rows).  One row of ten sums
// per block goes to epart[blockIdx.x * 10 ..] (slots as the standalone pass).
// SB = 1 (3-D loop, P2 = 0): u_bar and v_bar of the primal half, which the
// sweep otherwise keeps in LDS only, are also stored (to ff.bars), so that dst
// then holds the whole state after the primal half with p and q one dual half
// ahead (the adjoint's recompute of a segment, one launch per iteration).
template <class S, class I, int VX, int L2, int MK, int MAXT, int PF, int TL, int P2, int EN = 0,
          int SB = 0>
__global__ __launch_bounds__(MAXT) void fused_sweep_kernel(
    FusedFields ff, const float* __restrict__ f_, const uint8_t* __restrict__ act, Geometry g,
    float tau, float sigma, float lambda, float alpha0, float alpha1, float theta, int ty,
    int zchunk, int planar, int tx, double* __restrict__ epart,
    const uint8_t* __restrict__ tile_active, DivConsts dc) {
    static_assert(!EN || P2 == 0, "the in-sweep energy covers the 3-D loop");
    static_assert(!SB || P2 == 0, "u_bar / v_bar stores cover the 3-D loop");
    typedef typename S::T T;
    typedef VecIO<S, VX> V;
    typedef VecIO<StoreF32, VX> VF;
    extern __shared__ __align__(16) float lds[];

    const int W = g.w, H = g.h, D = g.d;
    constexpr bool SKZ = P2 == 1, IMG = P2 == 2;
    const int L = TL ? tx / VX + 2 : W / VX, RT = IMG ? ty : ty + 2;
    const int RWD = L * VX;               // floats per LDS row
    const int lane = int(threadIdx.x) % L, r = int(threadIdx.x) / L;
    const int nbx = TL ? (W + tx - 1) / tx : 1, nb = (H + ty - 1) / ty;
    const int t = xcd_tile(blockIdx.x, gridDim.x);
    // masked loop: a tile without an active voxel among the voxels it owns would
    // store its inputs unchanged; once both copies of the state agree outside
    // the active set (the host passes tile_active only then), it can skip
    if (MK && !EN && P2 == 0 && tile_active && !tile_active[t]) return;
    const int tb = TL ? t / nbx : t;
    const int y0 = (tb % nb) * ty;
    const int z0 = (tb / nb) * zchunk, z1 = min(z0 + zchunk, D);
    const int y = IMG ? y0 + r : y0 - 1 + r;
    const int xb = TL ? (t % nbx) * tx - VX + lane * VX : lane * VX;   // global x of element 0
    // (W % VX == 0 and x0 % VX == 0: a lane lies wholly inside or outside the volume)
    const bool rowin = y >= 0 && y < H && (!TL || (xb >= 0 && xb < W));
    const bool cl = !TL || lane > 0, cr = !TL || lane < L - 1;
    const bool own = rowin && (IMG || (r >= 1 && r <= ty)) && cl && cr;
    const I plane = I(W) * I(H);
    const I col = rowin ? I(y) * I(W) + I(xb) : I(0);
    const bool use_bz = !L2 || D > 1;
    const float tl = tau * lambda;

    // LDS carve-up
    const int RW = RT * RWD, RE = RT * (L + 1);
    float* const sPY = lds;               // P region: p_y, q_xy, q_yy, q_yz rows
    float* const sQXY = sPY + RW;
    float* const sQYY = sQXY + RW;
    float* const sQYZ = sQYY + RW;
    float* const sPXm = sQYZ + RW;        // p_x element VX-1 of lane at [r][lane+1]
    float* const sQXXp = sPXm + RE;       // q_xx element 0 of lane at [r][lane]
    float* const sQXYp = sQXXp + RE;
    float* const sQXZp = sQXYp + RE;
    float* const sUB = sQXZp + RE;        // D region: u_bar, v_bar x3 rows
    float* const sVB = sUB + RW;          // [3][RT][RWD]
    float* const sUBp = sVB + 3 * RW;     // u_bar element 0 at [r][lane]
    float* const sVBm = sUBp + RE;        // [3][RT][L+1]: v_bar element VX-1 at [r][lane+1]
    // EN: u and v of the finished iteration at plane s - 1 (stored values)
    float* const sUK = sVBm + 3 * RE;     // u rows
    float* const sUKp = sUK + RW;         // u element 0 at [r][lane]
    float* const sVK = sUKp + RE;         // [3][RT][RWD] v rows
    float* const sVKm = sVK + 3 * RW;     // [3][RT][L+1]: v element VX-1 at [r][lane+1]

    // zero everything once: edge slots at x = -1 / x = W (and beyond the halo
    // lanes) and rows outside the volume are never written afterwards and read
    // only under false flags
    for (int k = int(threadIdx.x); k < (EN ? 12 : 8) * (RW + RE); k += int(blockDim.x))
        lds[k] = 0.f;
    __syncthreads();

    const T* const* src = reinterpret_cast<const T* const*>(ff.src);
    T* const* dst = reinterpret_cast<T* const*>(ff.dst);
    const int rw = r * RWD + lane * VX;   // this thread's offset in a row array
    const int re = r * (L + 1) + lane;    // this thread's edge slot

    float p_c[3][VX], q_c[6][VX];         // p, q at plane s - 1 (dual base; p_z for B_z)
    float qn[3][VX];                      // q_zz, q_xz, q_yz at plane s
    float vb_m2[3][VX];                   // v_bar at plane s - 2
    uint32_t act_c = 0xffffffffu;         // active bytes at plane s - 1
    // EN: FP64 sums (slots of the standalone energy pass; 0 and 4 formed at the end)
    double acc[10];
    float vk_m[3][VX];                    // v (stored value) at plane s - 2
    const double lamd = double(lambda), a0d = double(alpha0), a1d = double(alpha1);
#pragma unroll
    for (int k = 0; k < 10; ++k) acc[k] = 0.0;
#pragma unroll
    for (int k = 0; k < 3; ++k) { zero(p_c[k]); zero(qn[k]); zero(vb_m2[k]); zero(vk_m[k]); }
#pragma unroll
    for (int k = 0; k < 6; ++k) zero(q_c[k]);

    const int zs = z0 > 0 ? z0 - 1 : 0;
    if (!SKZ && rowin && zs < D) {
        const I i = I(zs) * plane + col;
        V::ld(src[G_QZZ], i, qn[0]); V::ld(src[G_QXZ], i, qn[1]);
        if (!IMG) V::ld(src[G_QYZ], i, qn[2]);
    }

    // BF16 fields (the sweep is bound by instruction issue there): lanes outside
    // the volume (halo rows and columns at its faces, the plane past the last)
    // load the volume's first voxel instead, and the next-plane fields of the
    // last plane load the plane itself - no branch and no zeroed registers
    // (215.9 vs 225.5 us per launch at 256^3).  Those values reach no result:
    // every use of a neighbour or of the next plane is selected by its face
    // flag (fwd, bwd, and the energy form's flags), and such lanes neither
    // update nor store.  FP32 fields (bound by HBM; the same change measured
    // 365.3 vs 363.6 us): every field is zeroed and the loads overwrite what
    // exists.  With a zeroing else-branch instead, the compiler merged the two
    // branches' stores (in different field orders) through pointer phis and
    // kept the plane in scratch memory in the 2-D forms.  Only the 2-D forms'
    // absent components are +0 constants in both.
    auto load_plane = [&](int zz, FusedPlane<VX>& P) {
        constexpr bool CLAMP = std::is_same<S, StoreBF16>::value;
        const bool ok = rowin && zz < D;
        if (!CLAMP) {
            zero(P.px); zero(P.py); zero(P.pz); zero(P.qxx); zero(P.qxy); zero(P.qyy);
            zero(P.u); zero(P.f);
#pragma unroll
            for (int k = 0; k < 3; ++k) { zero(P.v[k]); zero(P.qnn[k]); }
            P.act = 0xffffffffu;
        }
        if (CLAMP || ok) {
            const I i = !CLAMP || ok ? I(zz) * plane + col : I(0);
            V::ld(src[G_PX], i, P.px);
            if (IMG) zero(P.py); else V::ld(src[G_PY], i, P.py);
            if (SKZ) zero(P.pz); else V::ld(src[G_PZ], i, P.pz);
            V::ld(src[G_QXX], i, P.qxx);
            if (IMG) { zero(P.qxy); zero(P.qyy); }
            else { V::ld(src[G_QXY], i, P.qxy); V::ld(src[G_QYY], i, P.qyy); }
            V::ld(src[G_U], i, P.u); VF::ld(f_, i, P.f);
            V::ld(src[G_VX], i, P.v[0]);
            if (IMG) zero(P.v[1]); else V::ld(src[G_VY], i, P.v[1]);
            if (SKZ) zero(P.v[2]); else V::ld(src[G_VZ], i, P.v[2]);
            if (SKZ) {
                zero(P.qnn[0]); zero(P.qnn[1]); zero(P.qnn[2]);
            } else if (CLAMP) {
                const I j = ok && zz + 1 < D ? i + plane : i;
                V::ld(src[G_QZZ], j, P.qnn[0]); V::ld(src[G_QXZ], j, P.qnn[1]);
                if (IMG) zero(P.qnn[2]); else V::ld(src[G_QYZ], j, P.qnn[2]);
            } else if (zz + 1 < D) {
                const I j = i + plane;
                V::ld(src[G_QZZ], j, P.qnn[0]); V::ld(src[G_QXZ], j, P.qnn[1]);
                if (!IMG) V::ld(src[G_QYZ], j, P.qnn[2]);
            }
            P.act = MK ? load_active<VX>(act, i) : 0xffffffffu;
        }
    };

    FusedPlane<VX> P;
    if (PF) load_plane(zs, P);

    for (int s = zs; s <= z1; ++s) {
        const bool pin = s < D;                       // plane s exists
        const I zo = I(s) * plane;
        const I i = zo + col;
        if (!PF) load_plane(s, P);
        float (&px)[VX] = P.px;
        float (&py)[VX] = P.py;
        float (&pz)[VX] = P.pz;
        float (&qxx)[VX] = P.qxx;
        float (&qxy)[VX] = P.qxy;
        float (&qyy)[VX] = P.qyy;
        float (&u)[VX] = P.u;
        float (&fv)[VX] = P.f;
        float (&v)[3][VX] = P.v;
        float (&qnn)[3][VX] = P.qnn;
        const uint32_t actv = P.act;

        // P region: plane s
        lds_put<VX>(&sPY[rw], py);
        lds_put<VX>(&sQXY[rw], qxy);
        lds_put<VX>(&sQYY[rw], qyy);
        lds_put<VX>(&sQYZ[rw], qn[2]);
        sPXm[re + 1] = px[VX - 1];
        sQXXp[re] = qxx[0];
        sQXYp[re] = qxy[0];
        sQXZp[re] = qn[1][0];
        __syncthreads();

        FusedPlane<VX> Pn;
        if (PF && s < z1) load_plane(s + 1, Pn);

        // ---- primal half at plane s
        const bool need_u = rowin && pin && (IMG || r > 0) && cl && s >= z0;
        const bool need_v = rowin && pin && (IMG || r < RT - 1) && cr && s < z1;
        float ubn[VX], vbn[3][VX];
        zero(ubn);
#pragma unroll
        for (int k = 0; k < 3; ++k) zero(vbn[k]);
        float uo[VX], vo[3][VX];
        if (EN) {
            zero(uo);
#pragma unroll
            for (int k = 0; k < 3; ++k) zero(vo[k]);
        }
        if (need_u) {
            float pym[VX];
            if (IMG) zero(pym); else lds_get<VX>(&sPY[rw - RWD], pym);
            const float pxm = sPXm[re];
#pragma unroll
            for (int e = 0; e < VX; ++e) {
                const int x = xb + e;
                const float bx = bwd(px[e], e > 0 ? px[e > 0 ? e - 1 : 0] : pxm, x > 0);
                const float by = bwd(py[e], pym[e], y > 0);
                const float bz = bwd(pz[e], p_c[2][e], s > 0);
                const float div = use_bz ? (bx + by) + bz : bx + by;
                const float uc = u[e], fc = fv[e];
                float un = fmaf(-tau, div, uc);
                if (L2) {
                    un = fmaf(tl, fc, un) / (1.0f + tl);
                } else {
                    const float umf = un - fc;
                    if (umf > tl) un = un - tl;
                    else if (umf < -tl) un = un + tl;
                    else un = fc;
                }
                if (MK && !((actv >> (8 * e)) & 0xffu)) un = uc;
                uo[e] = un;
                ubn[e] = fmaf(theta, un - uc, un);
            }
            S::template rt_all<VX>(ubn);   // u_bar as stored (rounded in pairs for BF16)
        }
        if (need_v) {
            float qxyn_y[VX], qyyn_y[VX], qyzn_y[VX];
            if (IMG) {   // no y neighbours (height 1)
                zero(qxyn_y); zero(qyyn_y); zero(qyzn_y);
            } else {
                lds_get<VX>(&sQXY[rw + RWD], qxyn_y);
                lds_get<VX>(&sQYY[rw + RWD], qyyn_y);
                lds_get<VX>(&sQYZ[rw + RWD], qyzn_y);
            }
            const float qxxn_x = sQXXp[re + 1];
            const float qxyn_x = sQXYp[re + 1];
            const float qxzn_x = sQXZp[re + 1];
#pragma unroll
            for (int e = 0; e < VX; ++e) {
                const int x = xb + e;
                const int en = e + 1 < VX ? e + 1 : e;
                const bool xl = e < VX - 1;
                const bool xn = x < W - 1, yn = y < H - 1, zn = s < D - 1;
                const float dxx = fwd(qxx[e], xl ? qxx[en] : qxxn_x, xn);
                const float dyxy = fwd(qxy[e], qxyn_y[e], yn);
                const float dzxz = fwd(qn[1][e], qnn[1][e], zn);
                const float dyy = fwd(qyy[e], qyyn_y[e], yn);
                const float dxxy = fwd(qxy[e], xl ? qxy[en] : qxyn_x, xn);
                const float dzyz = fwd(qn[2][e], qnn[2][e], zn);
                const float dzz = fwd(qn[0][e], qnn[0][e], zn);
                const float dxxz = fwd(qn[1][e], xl ? qn[1][en] : qxzn_x, xn);
                const float dyyz = fwd(qn[2][e], qyzn_y[e], yn);
                const float w0 = (dxx + dyxy) + dzxz;
                const float w1 = (dyy + dxxy) + dzyz;
                const float w2 = (dzz + dxxz) + dyyz;
                float vn0 = fmaf(-tau, w0 - px[e], v[0][e]);
                float vn1 = fmaf(-tau, w1 - py[e], v[1][e]);
                float vn2 = fmaf(-tau, w2 - pz[e], v[2][e]);
                if (MK && !((actv >> (8 * e)) & 0xffu)) { vn0 = v[0][e]; vn1 = v[1][e]; vn2 = v[2][e]; }
                vo[0][e] = vn0; vo[1][e] = vn1; vo[2][e] = vn2;
                vbn[0][e] = fmaf(theta, vn0 - v[0][e], vn0);
                vbn[1][e] = fmaf(theta, vn1 - v[1][e], vn1);
                vbn[2][e] = fmaf(theta, vn2 - v[2][e], vn2);
            }
#pragma unroll
            for (int k = 0; k < 3; ++k) S::template rt_all<VX>(vbn[k]);
        }
        if constexpr (EN) {
            // primal-half terms at plane s: data term with s = -(D+)^T p (true
            // transpose: p at the -x/-y/-z neighbours), L1 infeasibility, and
            // the v residual (eps_B^T q - p with B^T = D+ - e_last e_last^T:
            // q at the +x/+y/+z neighbours)
            if (own && pin && s >= z0 && s < z1) {
                float pym[VX], qxyn_y[VX], qyyn_y[VX], qyzn_y[VX];
                lds_get<VX>(&sPY[rw - RWD], pym);
                lds_get<VX>(&sQXY[rw + RWD], qxyn_y);
                lds_get<VX>(&sQYY[rw + RWD], qyyn_y);
                lds_get<VX>(&sQYZ[rw + RWD], qyzn_y);
                const float pxm = sPXm[re];
                const float qxxn_x = sQXXp[re + 1], qxyn_x = sQXYp[re + 1], qxzn_x = sQXZp[re + 1];
#pragma unroll
                for (int e = 0; e < VX; ++e) {
                    const int x = xb + e;
                    const int en = e + 1 < VX ? e + 1 : e;
                    const bool xl = e < VX - 1;
                    const bool hx = x > 0, hy = y > 0, hz = s > 0;
                    const bool nx = x < W - 1, ny = y < H - 1, nz = s < D - 1;
                    double sdiv = 0.0;
                    {
                        double t = 0.0;
                        if (hx) t += double(e > 0 ? px[e > 0 ? e - 1 : 0] : pxm);
                        if (nx) t -= double(px[e]);
                        sdiv -= t;
                    }
                    {
                        double t = 0.0;
                        if (hy) t += double(pym[e]);
                        if (ny) t -= double(py[e]);
                        sdiv -= t;
                    }
                    {
                        double t = 0.0;
                        if (hz) t += double(p_c[2][e]);
                        if (nz) t -= double(pz[e]);
                        sdiv -= t;
                    }
                    const double umf = double(S::rt(uo[e])) - double(fv[e]);
                    if (!L2) {
                        acc[1] += lamd * fabs(umf);
                        acc[5] += lamd * fabs(umf) - umf * sdiv;
                        acc[8] += fmax(0.0, fabs(sdiv) - lamd);
                    } else {
                        acc[1] += 0.5 * lamd * umf * umf;
                        const double rr = lamd * umf - sdiv;
                        acc[5] += rr * rr / (2.0 * lamd);
                    }
                    // (eps_B^T q)_c: sum over axes a of -q_ac + q_ac(+a) (last index: -q_ac)
                    const float qxz_c = qn[1][e], qyz_c = qn[2][e], qzz_c = qn[0][e];
                    const float qxx_x = xl ? qxx[en] : qxxn_x, qxy_x = xl ? qxy[en] : qxyn_x;
                    const float qxz_x = xl ? qn[1][en] : qxzn_x;
                    const float cq[3][3] = {{qxx[e], qxy[e], qxz_c},     // c = x: (xx, xy, xz)
                                            {qxy[e], qyy[e], qyz_c},     // c = y: (xy, yy, yz)
                                            {qxz_c, qyz_c, qzz_c}};      // c = z: (xz, yz, zz)
                    const float nq[3][3] = {{qxx_x, qxyn_y[e], qnn[1][e]},
                                            {qxy_x, qyyn_y[e], qnn[2][e]},
                                            {qxz_x, qyzn_y[e], qnn[0][e]}};
                    const bool na[3] = {nx, ny, nz};
                    const float pc[3] = {px[e], py[e], pz[e]};
#pragma unroll
                    for (int c = 0; c < 3; ++c) {
                        double t = 0.0;
#pragma unroll
                        for (int a = 0; a < 3; ++a) {
                            double bt = -double(cq[c][a]);
                            if (na[a]) bt += double(nq[c][a]);
                            t += bt;
                        }
                        const double rr = t - double(pc[c]);
                        acc[9] += rr * rr;
                    }
                }
            }
        }
        if (own && pin && s >= z0 && s < z1) {
            V::st(dst[G_U], i, uo);
            V::st(dst[G_VX], i, vo[0]);
            if (!IMG) V::st(dst[G_VY], i, vo[1]);
            if (!SKZ) V::st(dst[G_VZ], i, vo[2]);
            if constexpr (SB) {
                T* const* bars = reinterpret_cast<T* const*>(ff.bars);
                V::st(bars[0], i, ubn);
#pragma unroll
                for (int k = 0; k < 3; ++k) V::st(bars[1 + k], i, vbn[k]);
            }
        }

        // ---- dual half at plane zd = s - 1 (D region holds u_bar, v_bar at zd)
        const int zd = s - 1;
        float vbc[3][VX];
#pragma unroll
        for (int k = 0; k < 3; ++k) lds_get<VX>(&sVB[k * RW + rw], vbc[k]);
        float vkc[3][VX];                 // EN: v at plane zd
        if (EN) {
#pragma unroll
            for (int k = 0; k < 3; ++k) lds_get<VX>(&sVK[k * RW + rw], vkc[k]);
        }
        if (own && zd >= z0) {
            float ubc[VX], uyn[VX], vym[3][VX], vxm[3];
            lds_get<VX>(&sUB[rw], ubc);
            if (IMG) zero(uyn); else lds_get<VX>(&sUB[rw + RWD], uyn);
            const float uxn = sUBp[re + 1];
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                if (IMG) zero(vym[k]); else lds_get<VX>(&sVB[k * RW + rw - RWD], vym[k]);
                vxm[k] = sVBm[k * RE + re];
            }
            const I id = I(zd) * plane + col;
            float po[3][VX], qo[6][VX];
#pragma unroll
            for (int e = 0; e < VX; ++e) {
                const int x = xb + e;
                const float ub = ubc[e];
                const float dx = fwd(ub, e < VX - 1 ? ubc[e + 1 < VX ? e + 1 : e] : uxn, x < W - 1);
                const float dy = fwd(ub, uyn[e], y < H - 1);
                const float dz = fwd(ub, ubn[e], zd < D - 1);
                const float ax = fmaf(sigma, dx - vbc[0][e], p_c[0][e]);
                const float ay = fmaf(sigma, dy - vbc[1][e], p_c[1][e]);
                const float az = fmaf(sigma, dz - vbc[2][e], p_c[2][e]);
                float n2 = ax * ax;
                n2 = fmaf(ay, ay, n2);
                n2 = fmaf(az, az, n2);
                // |p| / alpha1 by the host's RN(1/alpha1) and one correction
                // (div_rn.hpp): the quotient rounded as the IEEE division rounds it
                const float tp = div_rn(sqrtf(n2), alpha1, dc.inv_alpha1);
                if (tp > 1.0f) {
                    po[0][e] = ax / tp; po[1][e] = ay / tp; po[2][e] = az / tp;
                } else {
                    po[0][e] = ax; po[1][e] = ay; po[2][e] = az;
                }
                float bxv[3], byv[3], bzv[3];
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    const float vc = vbc[k][e];
                    bxv[k] = bwd(vc, e > 0 ? vbc[k][e > 0 ? e - 1 : 0] : vxm[k], x > 0);
                    byv[k] = IMG ? 0.0f : bwd(vc, vym[k][e], y > 0);
                    bzv[k] = planar ? 0.0f : bwd(vc, vb_m2[k][e], zd > 0);
                }
                const float exy = (bxv[1] + byv[0]) * 0.5f;
                const float exz = (bxv[2] + bzv[0]) * 0.5f;
                const float eyz = (byv[2] + bzv[1]) * 0.5f;
                const float bxx = fmaf(sigma, bxv[0], q_c[0][e]);
                const float byy = fmaf(sigma, byv[1], q_c[1][e]);
                const float bxy = fmaf(sigma, exy, q_c[3][e]);
                const float bzz = fmaf(sigma, bzv[2], q_c[2][e]);
                const float bxz = fmaf(sigma, exz, q_c[4][e]);
                const float byz = fmaf(sigma, eyz, q_c[5][e]);
                float m2 = bxx * bxx;
                m2 = fmaf(byy, byy, m2);
                m2 = fmaf(2.0f * bxy, bxy, m2);
                m2 = fmaf(bzz, bzz, m2);
                m2 = fmaf(2.0f * bxz, bxz, m2);
                m2 = fmaf(2.0f * byz, byz, m2);
                const float tq = div_rn(sqrtf(m2), alpha0, dc.inv_alpha0);
                if (tq > 1.0f) {
                    qo[0][e] = bxx / tq; qo[1][e] = byy / tq; qo[2][e] = bzz / tq;
                    qo[3][e] = bxy / tq; qo[4][e] = bxz / tq; qo[5][e] = byz / tq;
                } else {
                    qo[0][e] = bxx; qo[1][e] = byy; qo[2][e] = bzz;
                    qo[3][e] = bxy; qo[4][e] = bxz; qo[5][e] = byz;
                }
                if (MK && !((act_c >> (8 * e)) & 0xffu)) {
#pragma unroll
                    for (int k = 0; k < 3; ++k) po[k][e] = p_c[k][e];
#pragma unroll
                    for (int k = 0; k < 6; ++k) qo[k][e] = q_c[k][e];
                }
            }
            if constexpr (EN) {
                // terms of the finished iteration at plane zd: alpha1 |D+u - v|
                // and its bracket with p, alpha0 |eps_B v|_w and its bracket with q
                float ukc[VX], uky[VX], vky[3][VX], vkxm[3];
                lds_get<VX>(&sUK[rw], ukc);
                lds_get<VX>(&sUK[rw + RWD], uky);
                const float ukx = sUKp[re + 1];
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    lds_get<VX>(&sVK[k * RW + rw - RWD], vky[k]);
                    vkxm[k] = sVKm[k * RE + re];
                }
#pragma unroll
                for (int e = 0; e < VX; ++e) {
                    const int x = xb + e;
                    const int en = e + 1 < VX ? e + 1 : e;
                    const bool nx = x < W - 1, ny = y < H - 1, nz = zd < D - 1;
                    const bool hx = x > 0, hy = y > 0, hz = zd > 0;
                    const double uc = ukc[e];
                    const double du[3] = {
                        nx ? double(e < VX - 1 ? ukc[en] : ukx) - uc : 0.0,
                        ny ? double(uky[e]) - uc : 0.0,
                        nz ? double(S::rt(uo[e])) - uc : 0.0};   // u at zd + 1 = s
                    double bv[3][3];
#pragma unroll
                    for (int a = 0; a < 3; ++a)
#pragma unroll
                        for (int c = 0; c < 3; ++c) {
                            const double vc = vkc[c][e];
                            const float nb = a == 0 ? (e > 0 ? vkc[c][e > 0 ? e - 1 : 0] : vkxm[c])
                                           : a == 1 ? vky[c][e] : vk_m[c][e];
                            const bool has = a == 0 ? hx : a == 1 ? hy : hz;
                            bv[a][c] = has ? double(nb) - vc : -vc;
                        }
                    double gn = 0.0, gp = 0.0;
#pragma unroll
                    for (int a = 0; a < 3; ++a) {
                        const double ga = du[a] - double(vkc[a][e]);
                        gn += ga * ga;
                        gp += ga * double(p_c[a][e]);
                    }
                    gn = sqrt(gn);
                    acc[2] += a1d * gn;
                    acc[6] += a1d * gn - gp;
                    const double ev[6] = {bv[0][0], bv[1][1], bv[2][2], 0.5 * (bv[0][1] + bv[1][0]),
                                          0.5 * (bv[0][2] + bv[2][0]), 0.5 * (bv[1][2] + bv[2][1])};
                    double enr = 0.0, eq = 0.0;
#pragma unroll
                    for (int k = 0; k < 6; ++k) {
                        const double wgt = k < 3 ? 1.0 : 2.0;
                        enr += wgt * ev[k] * ev[k];
                        eq += wgt * ev[k] * double(q_c[k][e]);
                    }
                    enr = sqrt(enr);
                    acc[3] += a0d * enr;
                    acc[7] += a0d * enr - eq;
                }
            }
            V::st(dst[G_PX], id, po[0]); V::st(dst[G_QXX], id, qo[0]);
            if (!IMG) {
                V::st(dst[G_PY], id, po[1]); V::st(dst[G_QYY], id, qo[1]);
                V::st(dst[G_QXY], id, qo[3]);
            }
            if (!SKZ) {
                V::st(dst[G_PZ], id, po[2]);
                V::st(dst[G_QZZ], id, qo[2]); V::st(dst[G_QXZ], id, qo[4]);
            }
            if (!SKZ && !IMG) V::st(dst[G_QYZ], id, qo[5]);
        }
        __syncthreads();

        // D region <- u_bar, v_bar at plane s; carry plane s to the next step
        lds_put<VX>(&sUB[rw], ubn);
        sUBp[re] = ubn[0];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            lds_put<VX>(&sVB[k * RW + rw], vbn[k]);
            sVBm[k * RE + re + 1] = vbn[k][VX - 1];
            copyv(vb_m2[k], vbc[k]);
        }
        if constexpr (EN) {   // u, v of plane s as stored
            float t[VX];
#pragma unroll
            for (int e = 0; e < VX; ++e) t[e] = S::rt(uo[e]);
            lds_put<VX>(&sUK[rw], t);
            sUKp[re] = t[0];
#pragma unroll
            for (int k = 0; k < 3; ++k) {
#pragma unroll
                for (int e = 0; e < VX; ++e) t[e] = S::rt(vo[k][e]);
                lds_put<VX>(&sVK[k * RW + rw], t);
                sVKm[k * RE + re + 1] = t[VX - 1];
                copyv(vk_m[k], vkc[k]);
            }
        }
        copyv(p_c[0], px); copyv(p_c[1], py); copyv(p_c[2], pz);
        copyv(q_c[0], qxx); copyv(q_c[1], qyy); copyv(q_c[2], qn[0]);
        copyv(q_c[3], qxy); copyv(q_c[4], qn[1]); copyv(q_c[5], qn[2]);
#pragma unroll
        for (int k = 0; k < 3; ++k) copyv(qn[k], qnn[k]);
        act_c = actv;
        if (PF && s < z1) P = Pn;
    }
    if constexpr (EN) {
        acc[0] = acc[1] + acc[2] + acc[3];
        acc[4] = acc[5] + acc[6] + acc[7];
        __syncthreads();   // the LDS rows are free again: [waves][10] doubles
        double* red = reinterpret_cast<double*>(lds);
        const int wl = int(threadIdx.x) & 63, wv = int(threadIdx.x) >> 6;
        const int nact = min(64, int(blockDim.x) - wv * 64);
#pragma unroll
        for (int k = 0; k < 10; ++k) {
            const double t = wave_sum_n(acc[k], wl, nact);
            if (wl == 0) red[wv * 10 + k] = t;
        }
        __syncthreads();
        if (threadIdx.x < 10) {
            const int nw = (int(blockDim.x) + 63) / 64;
            double t = 0.0;
            for (int w = 0; w < nw; ++w) t += red[w * 10 + threadIdx.x];
            epart[size_t(blockIdx.x) * 10 + threadIdx.x] = t;
        }
    }
}

template <class S>
__global__ void init_kernel(typename S::T* u, typename S::T* ubar, typename S::T* zero_base,
                            size_t zero_count, const float* f, size_t n) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const size_t stride = size_t(gridDim.x) * blockDim.x;
    for (size_t k = i; k < n; k += stride) {
        const float fv = f[k];
        S::st(u, k, fv);
        S::st(ubar, k, fv);
    }
    for (size_t k = i; k < zero_count; k += stride) S::st(zero_base, k, 0.0f);
}

template <class SI, class SO>
__global__ void convert_kernel(const typename SI::T* src, typename SO::T* dst, size_t n) {
    const size_t stride = size_t(gridDim.x) * blockDim.x;
    for (size_t k = size_t(blockIdx.x) * blockDim.x + threadIdx.x; k < n; k += stride)
        SO::st(dst, k, SI::ld(src, k));
}

template <int TX, int TY, int VX = 1>
int grid_blocks(const Geometry& g, int zchunk) {
    const int nbx = (g.w + TX * VX - 1) / (TX * VX), nby = (g.h + TY - 1) / TY;
    const int nbz = (g.d + zchunk - 1) / zchunk;
    return nbx * nby * nbz;
}

// Next-plane prefetch of the vector sweeps: on for FP32 fields, off for BF16
// (measured on MI355X at 256^3); SweepParams::sweep_pf = 0/1 sets it (tuning).
// Both variants compute identical results.
template <class S>
inline int sweep_prefetch(const SweepParams& p) {
    if (p.sweep_pf == 0 || p.sweep_pf == 1) return p.sweep_pf;
    return sizeof(typename S::T) == 4 ? 1 : 0;
}

template <class S, class I, int TX, int TY, int PF, int MK = 0>
hipError_t dual_vec_launch(hipStream_t s, void* const* fv, Geometry g, const SweepParams& p) {
    typedef typename S::T T;
    T* const* F = reinterpret_cast<T* const*>(fv);
    hipLaunchKernelGGL((dual_sweep_vec_kernel<S, I, TX, TY, 4, PF, MK>),
                       dim3(grid_blocks<TX, TY, 4>(g, p.zchunk)), dim3(TX * TY), 0, s, p.active, F[F_UBAR],
                       F[F_VBX], F[F_VBY], F[F_VBZ], F[F_PX], F[F_PY], F[F_PZ], F[F_QXX],
                       F[F_QYY], F[F_QZZ], F[F_QXY], F[F_QXZ], F[F_QYZ], g, p.sigma, p.alpha0,
                       p.alpha1, p.zchunk, p.planar);
    return hipGetLastError();
}

template <class S, class I, int TX, int TY, int PF, int MK = 0>
hipError_t primal_vec_launch(hipStream_t s, void* const* fv, const float* f, Geometry g,
                             const SweepParams& p) {
    typedef typename S::T T;
    T* const* F = reinterpret_cast<T* const*>(fv);
    const dim3 grid(grid_blocks<TX, TY, 4>(g, p.zchunk)), block(TX * TY);
    if (p.data_term == 1)
        hipLaunchKernelGGL((primal_sweep_vec_kernel<S, I, TX, TY, 4, PF, 1, MK>), grid, block, 0, s,
                           p.active, F[F_U], F[F_UBAR], F[F_VX], F[F_VY], F[F_VZ], F[F_VBX], F[F_VBY],
                           F[F_VBZ], F[F_PX], F[F_PY], F[F_PZ], F[F_QXX], F[F_QYY], F[F_QZZ],
                           F[F_QXY], F[F_QXZ], F[F_QYZ], f, g, p.tau, p.lambda, p.theta,
                           p.zchunk);
    else
        hipLaunchKernelGGL((primal_sweep_vec_kernel<S, I, TX, TY, 4, PF, 0, MK>), grid, block, 0, s,
                           p.active, F[F_U], F[F_UBAR], F[F_VX], F[F_VY], F[F_VZ], F[F_VBX], F[F_VBY],
                           F[F_VBZ], F[F_PX], F[F_PY], F[F_PZ], F[F_QXX], F[F_QYY], F[F_QZZ],
                           F[F_QXY], F[F_QXZ], F[F_QYZ], f, g, p.tau, p.lambda, p.theta,
                           p.zchunk);
    return hipGetLastError();
}

template <class S, class I, int TX, int TY>
hipError_t dual_vec_pf(hipStream_t s, void* const* fv, Geometry g, const SweepParams& p) {
    return sweep_prefetch<S>(p) ? dual_vec_launch<S, I, TX, TY, 1>(s, fv, g, p)
                            : dual_vec_launch<S, I, TX, TY, 0>(s, fv, g, p);
}
template <class S, class I, int TX, int TY>
hipError_t primal_vec_pf(hipStream_t s, void* const* fv, const float* f, Geometry g,
                         const SweepParams& p) {
    return sweep_prefetch<S>(p) ? primal_vec_launch<S, I, TX, TY, 1>(s, fv, f, g, p)
                            : primal_vec_launch<S, I, TX, TY, 0>(s, fv, f, g, p);
}


template <class S, class I, int TX, int TY, int MK = 0>
hipError_t dual_launch(hipStream_t s, void* const* fv, Geometry g, const SweepParams& p) {
    typedef typename S::T T;
    T* const* F = reinterpret_cast<T* const*>(fv);
    hipLaunchKernelGGL((dual_sweep_kernel<S, I, TX, TY, MK>), dim3(grid_blocks<TX, TY>(g, p.zchunk)),
                       dim3(TX * TY), 0, s, p.active, F[F_UBAR], F[F_VBX], F[F_VBY], F[F_VBZ], F[F_PX],
                       F[F_PY], F[F_PZ], F[F_QXX], F[F_QYY], F[F_QZZ], F[F_QXY], F[F_QXZ],
                       F[F_QYZ], g, p.sigma, p.alpha0, p.alpha1, p.zchunk, p.planar);
    return hipGetLastError();
}

template <class S, class I, int TX, int TY, int MK = 0>
hipError_t primal_launch(hipStream_t s, void* const* fv, const float* f, Geometry g,
                         const SweepParams& p) {
    typedef typename S::T T;
    T* const* F = reinterpret_cast<T* const*>(fv);
    const dim3 grid(grid_blocks<TX, TY>(g, p.zchunk)), block(TX * TY);
    if (p.data_term == 1)
        hipLaunchKernelGGL((primal_sweep_kernel<S, I, TX, TY, 1, MK>), grid, block, 0, s, p.active, F[F_U],
                           F[F_UBAR], F[F_VX], F[F_VY], F[F_VZ], F[F_VBX], F[F_VBY], F[F_VBZ],
                           F[F_PX], F[F_PY], F[F_PZ], F[F_QXX], F[F_QYY], F[F_QZZ], F[F_QXY],
                           F[F_QXZ], F[F_QYZ], f, g, p.tau, p.lambda, p.theta, p.zchunk);
    else
        hipLaunchKernelGGL((primal_sweep_kernel<S, I, TX, TY, 0, MK>), grid, block, 0, s, p.active, F[F_U],
                           F[F_UBAR], F[F_VX], F[F_VY], F[F_VZ], F[F_VBX], F[F_VBY], F[F_VBZ],
                           F[F_PX], F[F_PY], F[F_PZ], F[F_QXX], F[F_QYY], F[F_QZZ], F[F_QXY],
                           F[F_QXZ], F[F_QYZ], f, g, p.tau, p.lambda, p.theta, p.zchunk);
    return hipGetLastError();
}

// Masked sweeps use the default shapes only (64x4 threads; vector tiles when
// W % 4 == 0, with the storage's default prefetch).
template <class S, class I>
hipError_t dual_tiles(hipStream_t s, void* const* fv, Geometry g, const SweepParams& p) {
    if (p.active) {
        if (p.tile >= 4)
            return sizeof(typename S::T) == 4 ? dual_vec_launch<S, I, 64, 4, 1, 1>(s, fv, g, p)
                                              : dual_vec_launch<S, I, 64, 4, 0, 1>(s, fv, g, p);
        return dual_launch<S, I, 64, 4, 1>(s, fv, g, p);
    }
    switch (p.tile) {
        case 4: return dual_vec_pf<S, I, 64, 4>(s, fv, g, p);
        case 5: return dual_vec_pf<S, I, 64, 2>(s, fv, g, p);
        case 6: return dual_vec_pf<S, I, 64, 8>(s, fv, g, p);
        case 7: return dual_vec_pf<S, I, 32, 8>(s, fv, g, p);
        case 1: return dual_launch<S, I, 64, 8>(s, fv, g, p);
        case 2: return dual_launch<S, I, 32, 8>(s, fv, g, p);
        case 3: return dual_launch<S, I, 64, 2>(s, fv, g, p);
        default: return dual_launch<S, I, 64, 4>(s, fv, g, p);
    }
}

template <class S, class I>
hipError_t primal_tiles(hipStream_t s, void* const* fv, const float* f, Geometry g,
                        const SweepParams& p) {
    if (p.active) {
        if (p.tile >= 4)
            return sizeof(typename S::T) == 4 ? primal_vec_launch<S, I, 64, 4, 1, 1>(s, fv, f, g, p)
                                              : primal_vec_launch<S, I, 64, 4, 0, 1>(s, fv, f, g, p);
        return primal_launch<S, I, 64, 4, 1>(s, fv, f, g, p);
    }
    switch (p.tile) {
        case 4: return primal_vec_pf<S, I, 64, 4>(s, fv, f, g, p);
        case 5: return primal_vec_pf<S, I, 64, 2>(s, fv, f, g, p);
        case 6: return primal_vec_pf<S, I, 64, 8>(s, fv, f, g, p);
        case 7: return primal_vec_pf<S, I, 32, 8>(s, fv, f, g, p);
        case 1: return primal_launch<S, I, 64, 8>(s, fv, f, g, p);
        case 2: return primal_launch<S, I, 32, 8>(s, fv, f, g, p);
        case 3: return primal_launch<S, I, 64, 2>(s, fv, f, g, p);
        default: return primal_launch<S, I, 64, 4>(s, fv, f, g, p);
    }
}

template <class S, class I, int VX, int L2, int MK, int MAXT, int PF, int TL = 0, int P2 = 0,
          int EN = 0, int SB = 0>
hipError_t fused_launch(hipStream_t s, const FusedFields& ff, const float* f, Geometry g,
                        const SweepParams& p, const FusedPlan& plan) {
    const int row = TL ? plan.tx + 2 * VX : g.w;
    const int L = row / VX, RT = P2 == 2 ? plan.ty : plan.ty + 2;
    if (L * RT > MAXT) return hipErrorInvalidConfiguration;
    if (!fused_divisors_ok(p.alpha0, p.alpha1)) return hipErrorInvalidValue;
    const int nbx = TL ? (g.w + plan.tx - 1) / plan.tx : 1;
    const int nb = (g.h + plan.ty - 1) / plan.ty, nc = (g.d + plan.zchunk - 1) / plan.zchunk;
    size_t lds = (size_t(EN ? 12 : 8) * RT * row + size_t(EN ? 12 : 8) * RT * (L + 1)) * sizeof(float);
    if (EN) lds = std::max(lds, size_t(16 * 10 * sizeof(double)));   // the block reduction
    auto kern = fused_sweep_kernel<S, I, VX, L2, MK, MAXT, PF, TL, P2, EN, SB>;
    hipError_t e = ensure_dynamic_lds(reinterpret_cast<const void*>(kern), lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(kern, dim3(nbx * nb * nc), dim3(L * RT), lds, s, ff, f, p.active, g, p.tau,
                       p.sigma, p.lambda, p.alpha0, p.alpha1, p.theta, plan.ty, plan.zchunk,
                       p.planar, TL ? plan.tx : 0, EN ? p.energy_partials : nullptr,
                       MK && !EN && P2 == 0 ? p.tile_active : nullptr,
                       div_consts(p.alpha0, p.alpha1));
    return hipGetLastError();
}

// The energy form (EN = 1): 32-bit offsets, the 3-D loop, two voxels per
// thread, no next-plane prefetch, blocks of at most EN_THREADS threads (two
// waves per SIMD: the FP64 sums and the carried u, v need the registers); the
// plan comes from fused_energy_plan.
constexpr int EN_THREADS = 512;

template <class S, int MK>
hipError_t fused_energy_mk(hipStream_t s, const FusedFields& ff, const float* f, Geometry g,
                           const SweepParams& p, const FusedPlan& plan) {
    typedef uint32_t I;
    constexpr int T = EN_THREADS;
    if (plan.vx == 1) {   // odd widths
        if (plan.tx > 0)
            return p.data_term == 1 ? fused_launch<S, I, 1, 1, MK, T, 0, 1, 0, 1>(s, ff, f, g, p, plan)
                                    : fused_launch<S, I, 1, 0, MK, T, 0, 1, 0, 1>(s, ff, f, g, p, plan);
        return p.data_term == 1 ? fused_launch<S, I, 1, 1, MK, T, 0, 0, 0, 1>(s, ff, f, g, p, plan)
                                : fused_launch<S, I, 1, 0, MK, T, 0, 0, 0, 1>(s, ff, f, g, p, plan);
    }
    if (plan.vx != 2) return hipErrorInvalidValue;
    if (plan.tx > 0)
        return p.data_term == 1 ? fused_launch<S, I, 2, 1, MK, T, 0, 1, 0, 1>(s, ff, f, g, p, plan)
                                : fused_launch<S, I, 2, 0, MK, T, 0, 1, 0, 1>(s, ff, f, g, p, plan);
    return p.data_term == 1 ? fused_launch<S, I, 2, 1, MK, T, 0, 0, 0, 1>(s, ff, f, g, p, plan)
                            : fused_launch<S, I, 2, 0, MK, T, 0, 0, 0, 1>(s, ff, f, g, p, plan);
}

template <class S>
hipError_t fused_energy_dispatch(hipStream_t s, const FusedFields& ff, const float* f, Geometry g,
                                 const SweepParams& p, const FusedPlan& plan) {
    return p.active ? fused_energy_mk<S, 1>(s, ff, f, g, p, plan)
                    : fused_energy_mk<S, 0>(s, ff, f, g, p, plan);
}

// VX = 4: blocks of up to 512 threads (two waves per SIMD, ~200-250 VGPRs).
// VX = 2: blocks of up to 1024 threads (four waves per SIMD, <= 128 VGPRs),
// no prefetch (it would spill).  64-bit offsets (volumes over 2^30 voxels):
// VX = 4 without prefetch.
template <class S, class I, int MK, int PF>
hipError_t fused_shape(hipStream_t s, const FusedFields& ff, const float* f, Geometry g,
                       const SweepParams& p, const FusedPlan& plan) {
    if constexpr (sizeof(I) == 4 && PF == 0) {
        if (plan.vx == 2)
            return p.data_term == 1 ? fused_launch<S, I, 2, 1, MK, 1024, 0>(s, ff, f, g, p, plan)
                                    : fused_launch<S, I, 2, 0, MK, 1024, 0>(s, ff, f, g, p, plan);
    }
    return p.data_term == 1 ? fused_launch<S, I, 4, 1, MK, 512, PF>(s, ff, f, g, p, plan)
                             : fused_launch<S, I, 4, 0, MK, 512, PF>(s, ff, f, g, p, plan);
}

template <class S, class I, int MK>
hipError_t fused_pf(hipStream_t s, const FusedFields& ff, const float* f, Geometry g,
                    const SweepParams& p, const FusedPlan& plan) {
    if constexpr (sizeof(I) == 8) {
        if (plan.tx > 0)   // x tiles, VX = 4 (64-bit offsets)
            return p.data_term == 1 ? fused_launch<S, I, 4, 1, MK, 512, 0, 1, 0>(s, ff, f, g, p, plan)
                                    : fused_launch<S, I, 4, 0, MK, 512, 0, 1, 0>(s, ff, f, g, p, plan);
        return fused_shape<S, I, MK, 0>(s, ff, f, g, p, plan);
    } else {
        if (plan.vx == 1) {   // odd widths: one voxel per thread, 1024-thread blocks
            if (plan.tx > 0)
                return p.data_term == 1 ? fused_launch<S, I, 1, 1, MK, 1024, 0, 1>(s, ff, f, g, p, plan)
                                        : fused_launch<S, I, 1, 0, MK, 1024, 0, 1>(s, ff, f, g, p, plan);
            return p.data_term == 1 ? fused_launch<S, I, 1, 1, MK, 1024, 0>(s, ff, f, g, p, plan)
                                    : fused_launch<S, I, 1, 0, MK, 1024, 0>(s, ff, f, g, p, plan);
        }
        if constexpr (std::is_same<S, StoreF32>::value) {
            if (p.planar_zero_z && plan.vx == 2) {   // the 2-D loop: z fields taken as +0
                const bool pf = plan.pf && (plan.tx > 0 ? (plan.tx / 2 + 2) : g.w / 2) *
                                               (plan.ty + 2) <= 768;
                if (plan.tx > 0) {
                    if (pf)
                        return p.data_term == 1
                                   ? fused_launch<S, I, 2, 1, MK, 768, 1, 1, 1>(s, ff, f, g, p, plan)
                                   : fused_launch<S, I, 2, 0, MK, 768, 1, 1, 1>(s, ff, f, g, p, plan);
                    return p.data_term == 1
                               ? fused_launch<S, I, 2, 1, MK, 1024, 0, 1, 1>(s, ff, f, g, p, plan)
                               : fused_launch<S, I, 2, 0, MK, 1024, 0, 1, 1>(s, ff, f, g, p, plan);
                }
                if (pf)
                    return p.data_term == 1
                               ? fused_launch<S, I, 2, 1, MK, 768, 1, 0, 1>(s, ff, f, g, p, plan)
                               : fused_launch<S, I, 2, 0, MK, 768, 1, 0, 1>(s, ff, f, g, p, plan);
                return p.data_term == 1
                           ? fused_launch<S, I, 2, 1, MK, 1024, 0, 0, 1>(s, ff, f, g, p, plan)
                           : fused_launch<S, I, 2, 0, MK, 1024, 0, 0, 1>(s, ff, f, g, p, plan);
            }
        }
        if (plan.tx > 0) {   // x tiles, VX = 2
            if (plan.pf && (plan.tx / 2 + 2) * (plan.ty + 2) <= 768)
                return p.data_term == 1
                           ? fused_launch<S, I, 2, 1, MK, 768, 1, 1>(s, ff, f, g, p, plan)
                           : fused_launch<S, I, 2, 0, MK, 768, 1, 1>(s, ff, f, g, p, plan);
            return p.data_term == 1 ? fused_launch<S, I, 2, 1, MK, 1024, 0, 1>(s, ff, f, g, p, plan)
                                    : fused_launch<S, I, 2, 0, MK, 1024, 0, 1>(s, ff, f, g, p, plan);
        }
        // VX = 2 with the prefetch: blocks of up to 768 threads (three waves per SIMD)
        if (plan.pf && plan.vx == 2 && (g.w / 2) * (plan.ty + 2) <= 768)
            return p.data_term == 1 ? fused_launch<S, I, 2, 1, MK, 768, 1>(s, ff, f, g, p, plan)
                                    : fused_launch<S, I, 2, 0, MK, 768, 1>(s, ff, f, g, p, plan);
        return plan.pf && plan.vx == 4 ? fused_shape<S, I, MK, 1>(s, ff, f, g, p, plan)
                                       : fused_shape<S, I, MK, 0>(s, ff, f, g, p, plan);
    }
}

template <class S, class I>
hipError_t fused_dispatch(hipStream_t s, const FusedFields& ff, const float* f, Geometry g,
                          const SweepParams& p, const FusedPlan& plan) {
    return p.active ? fused_pf<S, I, 1>(s, ff, f, g, p, plan)
                    : fused_pf<S, I, 0>(s, ff, f, g, p, plan);
}

// The SB form (u_bar / v_bar stored too) for the adjoint's recompute: FP32
// fields, 32-bit offsets, the 3-D loop, no mask, the L2 data term, two voxels
// per thread (one on odd widths); whole rows or x tiles and the prefetch as the
// plan says.
hipError_t fused_bars_dispatch(hipStream_t s, const FusedFields& ff, const float* f, Geometry g,
                               const SweepParams& p, const FusedPlan& plan) {
    typedef StoreF32 S;
    typedef uint32_t I;
    if (p.data_term != 1) return hipErrorInvalidValue;
    if (plan.vx == 1)
        return plan.tx > 0 ? fused_launch<S, I, 1, 1, 0, 1024, 0, 1, 0, 0, 1>(s, ff, f, g, p, plan)
                           : fused_launch<S, I, 1, 1, 0, 1024, 0, 0, 0, 0, 1>(s, ff, f, g, p, plan);
    if (plan.vx != 2) return hipErrorInvalidValue;
    const int lanes = plan.tx > 0 ? plan.tx / 2 + 2 : g.w / 2;
    const bool pf = plan.pf && lanes * (plan.ty + 2) <= 768;
    if (plan.tx > 0)
        return pf ? fused_launch<S, I, 2, 1, 0, 768, 1, 1, 0, 0, 1>(s, ff, f, g, p, plan)
                  : fused_launch<S, I, 2, 1, 0, 1024, 0, 1, 0, 0, 1>(s, ff, f, g, p, plan);
    return pf ? fused_launch<S, I, 2, 1, 0, 768, 1, 0, 0, 0, 1>(s, ff, f, g, p, plan)
              : fused_launch<S, I, 2, 1, 0, 1024, 0, 0, 0, 0, 1>(s, ff, f, g, p, plan);
}

}  // namespace

// Plan of the fused sweep for a geometry: voxels per thread (vx), band height
// ty and z chunk.  Cost model: rounds of resident blocks x blocks per CU x the
// rows and planes each block marches (the halo row / plane on each side does
// half the work of an owned one).  Blocks resident per CU follow from LDS
// (160 KB) and registers: vx = 4 kernels use ~200-250 VGPRs (two waves per
// SIMD), vx = 2 kernels at most 128 (four waves per SIMD).  With BF16 fields
// the sweep is latency-bound rather than bandwidth-bound (measured on MI355X
// at 256^3), so the cost is scaled by 16 / (waves per CU, at most 16) and the
// halo rows count half; with FP32 fields the halo traffic dominates.  Ties go
// to vx = 2 and to the lower band.
// The overrides: fused = 0 disables the fused sweep; fused_vx / _ty / _zchunk /
// _pf set the plan, and fused_tx the tile width of rows too wide for whole rows.
bool fused_plan(const Geometry& g, int storage, int device, const PlanOverrides& ov,
                FusedPlan* plan) {
    if (ov.fused == 0) return false;
    int max_lds = 0, cus = 0;
    if (hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, device) !=
            hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess)
        return false;
    const int want_vx = std::max(0, ov.fused_vx), want_ty = std::max(0, ov.fused_ty);
    const int want_zc = std::max(0, ov.fused_zchunk), want_tx = std::max(0, ov.fused_tx);
    const bool big_index = needs_index64(g, ov);
    const int want_pf = ov.fused_pf;   // 0/1, else the plan's own choice
    double best = 1e300;
    FusedPlan bp{0, 0, 1, 0, 0};
    // depth 1 (images) with rows too wide for whole-row bands: nothing along z
    // amortises the halo rows of a band, so such images run in tiles of 124
    // columns (64 two-voxel lanes) and bands of 6 rows, not the tile the cost
    // model below would pick.  Measured per iteration: 4096^2 456 us (the cost
    // model's tile 681 us, bands of 4 / 10 rows 543 / 501 us, 252-column tiles
    // 536 us, the two sweeps 564 us); 1024^2 38 us (the two sweeps 52 us).
    // Images of at most 682 columns keep whole rows (512^2: 11.5 us, tiles 14.2).
    const bool image = g.d == 1 && !big_index && g.w % 2 == 0 && !want_vx && want_tx == 0;
    if (image && 3 * (g.w / 2) > 1024) {
        const int ty = std::min(want_ty ? want_ty : 6, g.h);
        if (size_t(fused_lds_floats<2>(128, ty)) * sizeof(float) <= size_t(max_lds)) {
            FusedPlan ip{ty, 1, storage == 0 ? 1 : 0, 2, 124};
            if (want_pf == 0 || want_pf == 1) ip.pf = want_pf;
            *plan = ip;
            return true;
        }
    }
    // odd widths: one voxel per thread (the two- and four-voxel forms need
    // lanes that lie wholly inside the row), whole rows only while bands of two
    // rows fit (four rows of lanes): bands of one row read each row three times
    // (257^3 per launch: 755 us, tiles of 126 columns 549 us; 255^3 with bands
    // of two rows 431 us, of one row 662 us, tiles 542 us)
    for (int vx : {1, 2, 4}) {
        if (want_vx && vx != want_vx) continue;
        if (want_tx > 0) break;   // x tiles forced
        if (g.w % vx != 0 || (vx <= 2 && big_index) || (vx == 1 && g.w % 2 == 0)) continue;
        if (vx == 1 && 4 * g.w > 1024 && !want_ty) continue;
        const int L = g.w / vx;
        const int max_threads = vx == 4 ? 512 : 1024;
        const int waves_simd_cap = vx == 4 ? 2 : 4;
        for (int rt = 3; rt * L <= max_threads && rt - 2 <= g.h + 1; ++rt) {
            const int ty = rt - 2;
            if (want_ty && ty != want_ty) continue;
            const size_t lds = (vx == 4   ? fused_lds_floats<4>(g.w, ty)
                                : vx == 2 ? fused_lds_floats<2>(g.w, ty)
                                          : fused_lds_floats<1>(g.w, ty)) *
                               sizeof(float);
            if (lds > size_t(max_lds)) break;
            const int waves_simd = ((rt * L + 63) / 64 + 3) / 4;
            const int per_cu = std::max(
                1, std::min(int((160 * 1024) / lds), waves_simd_cap / waves_simd));
            const long conc = long(cus) * per_cu;
            const long nb = (g.h + ty - 1) / ty;
            for (int nc = 1; nc <= g.d; ++nc) {
                const int zc = want_zc ? want_zc : (g.d + nc - 1) / nc;
                const long ncz = (g.d + zc - 1) / zc;
                const long blocks = nb * ncz;
                const long rounds = (blocks + conc - 1) / conc;
                const int waves_cu = std::min(16, per_cu * ((rt * L + 63) / 64));
                const double occ = storage == 1 ? 16.0 / waves_cu : 1.0;
                const double halo = storage == 1 ? 0.5 : 1.0;
                const double cost = double(rounds) * per_cu * (ty + halo) * (zc + 1) * occ;
                if (cost < best) { best = cost; bp = FusedPlan{ty, zc, vx == 4 ? 1 : 0, vx}; }
            }
        }
    }
    if (bp.ty == 0 && big_index && g.w % 4 == 0 && (!want_vx || want_vx == 4)) {
        // 64-bit offsets and rows too wide for whole-row bands: tiles of 248
        // columns (64 four-voxel lanes with a halo lane on each side) and bands
        // of up to 6 rows in 512-thread blocks; the cost model picks the z chunk
        const int tx = want_tx > 0 ? want_tx : 248;
        if (tx % 4 == 0 && tx < g.w) {
            const int lanes = tx / 4 + 2, ty = std::min(want_ty ? want_ty : 6, g.h);
            const size_t lds = fused_lds_floats<4>(tx + 8, ty) * sizeof(float);
            if (lanes * (ty + 2) <= 512 && lds <= size_t(max_lds)) {
                const long nb = long((g.w + tx - 1) / tx) * ((g.h + ty - 1) / ty);
                // one block per CU: the four-voxel form's VGPRs allow two waves
                // per SIMD, which one 512-thread block fills
                for (int nc = 1; nc <= g.d; ++nc) {
                    const int zc = want_zc ? want_zc : (g.d + nc - 1) / nc;
                    const long blocks = nb * ((g.d + zc - 1) / zc);
                    const long rounds = (blocks + cus - 1) / cus;
                    const double cost = double(rounds) * (ty + 1.0) * (zc + 1);
                    if (cost < best) { best = cost; bp = FusedPlan{ty, zc, 0, 4, tx}; }
                    if (want_zc) break;
                }
            }
        }
        if (bp.ty == 0) return false;
        *plan = bp;
        return true;
    }
    if (bp.ty == 0 && !big_index && g.w % 2 == 0 && (!want_vx || want_vx == 2)) {
        // rows too wide for whole-row bands: tiles of 124 columns (64 two-voxel
        // lanes, one wave, with a halo lane on each side) and bands of up to 8
        // rows, FP32 with the prefetch (<= 768 threads), BF16 without (<= 1024);
        // the cost model picks the z chunk.  Measured at 768^3 per launch: about
        // 12-13 ms for 124 x 8, 124 x 10, 188 x 6 and 252 x 4 (10% run to run),
        // 13.4 ms for 380 x 2, 15.9 ms for 508 x 1; the two sweeps take 16.3 ms.
        const int maxt = storage == 0 ? 768 : 1024, cap = storage == 0 ? 3 : 4;
        for (int txc : {124}) {
            const int tx = want_tx ? want_tx : txc;
            if (tx % 2 != 0 || tx >= g.w) { if (want_tx) break; continue; }
            const int lanes = tx / 2 + 2;
            const long nbx = (g.w + tx - 1) / tx;
            for (int rt = 3; rt * lanes <= maxt && rt - 2 <= g.h + 1; ++rt) {
                const int ty = rt - 2;
                if (want_ty ? ty != want_ty : ty != std::min(8, g.h)) continue;
                const size_t lds = fused_lds_floats<2>(tx + 4, ty) * sizeof(float);
                if (lds > size_t(max_lds)) break;
                const int waves_simd = ((rt * lanes + 63) / 64 + 3) / 4;
                const int per_cu = std::max(1, std::min(int((160 * 1024) / lds), cap / waves_simd));
                const long conc = long(cus) * per_cu;
                const long nb = nbx * ((g.h + ty - 1) / ty);
                for (int nc = 1; nc <= g.d; ++nc) {
                    const int zc = want_zc ? want_zc : (g.d + nc - 1) / nc;
                    const long blocks = nb * ((g.d + zc - 1) / zc);
                    const long rounds = (blocks + conc - 1) / conc;
                    const double cost = double(rounds) * per_cu * (ty + 1.0) * (zc + 1) *
                                        (2.0 * lanes / tx);
                    if (cost < best) {
                        best = cost;
                        bp = FusedPlan{ty, zc, storage == 0 ? 1 : 0, 2, tx};
                    }
                    if (want_zc) break;
                }
            }
            if (want_tx) break;
        }
        if (bp.ty == 0) return false;
        if (want_pf == 0 || want_pf == 1) bp.pf = want_pf;
        *plan = bp;
        return true;
    }
    if (bp.ty == 0 && !big_index && g.w % 2 != 0 && g.d > 1 && (!want_vx || want_vx == 1)) {
        // odd rows too wide for whole-row bands: x tiles of 126 columns (128
        // one-voxel lanes with a halo lane on each side) and bands of up to 4
        // rows, 768-thread blocks (the kernel's ~70 VGPRs fit two per CU), no
        // prefetch; the cost model picks the z chunk.  Measured per launch
        // (FP32 / BF16 us; the best of 62, 94, 126 columns x 2..8 rows; the two
        // sweeps): 257^3 553 / 374 (522 / 328; 853 / 676); 128x256x511 529 /
        // 311 (489 / 311; 766 / 495); 200x200x383 473 / 298 (best; 733 / 512);
        // 100x300x765 685 / 449 (best; 1061 / 755); 48x128x2047 369 / 221 (340 /
        // 221; 554 / 379); 64x300x301 209 / 129 (182 / 118; 258 / 121).
        // Choosing the width by the fewest lanes over the row did worse (up to
        // 1.26x).  Not for images (depth 1): nothing along z amortises the halo
        // rows there (2047^2 per iteration: FP32 158 us, the two sweeps 161 us;
        // BF16 127 us, the two sweeps 102 us).
        const int tx = want_tx > 0 ? want_tx : 126;
        const int lanes = tx + 2, ty = std::min(want_ty ? want_ty : 4, g.h);
        if (tx >= g.w || lanes * (ty + 2) > 1024) return false;
        const size_t lds = fused_lds_floats<1>(tx + 2, ty) * sizeof(float);
        if (lds > size_t(max_lds)) return false;
        const int waves_simd = ((lanes * (ty + 2) + 63) / 64 + 3) / 4;
        const int per_cu = std::max(1, std::min(int((160 * 1024) / lds), 7 / waves_simd));
        const long conc = long(cus) * per_cu;
        const long nb = long((g.w + tx - 1) / tx) * ((g.h + ty - 1) / ty);
        for (int nc = 1; nc <= g.d; ++nc) {
            const int zc = want_zc ? want_zc : (g.d + nc - 1) / nc;
            const long blocks = nb * ((g.d + zc - 1) / zc);
            const long rounds = (blocks + conc - 1) / conc;
            const double cost = double(rounds) * per_cu * (ty + 1.0) * (zc + 1);
            if (cost < best) { best = cost; bp = FusedPlan{ty, zc, 0, 1, tx}; }
            if (want_zc) break;
        }
        if (bp.ty == 0) return false;
        *plan = bp;
        return true;
    }
    if (bp.ty == 0) return false;
    // FP32 fields, two voxels per thread, at most 768 threads: the next plane's
    // loads issued one step ahead fit the 168 VGPRs of three waves per SIMD
    // (measured at 256^3, band 4: 366.5 vs 373.7 us; with BF16 fields it is
    // slower, 528 vs 316 us)
    if (storage == 0 && bp.vx == 2 && !big_index && (g.w / 2) * (bp.ty + 2) <= 768) bp.pf = 1;
    if (want_pf == 0 || want_pf == 1) bp.pf = want_pf;
    if (bp.vx == 1) bp.pf = 0;   // the one-voxel form has no prefetch variant
    *plan = bp;
    return true;
}

// The energy form's plan (results do not depend on the plan): two voxels per
// thread (one on odd widths), no prefetch, the iteration plan's z chunk; whole
// rows in bands of up to the iteration plan's height while three rows fit
// EN_THREADS threads, else x tiles of 124 columns (64 lanes) in bands of up to
// 6 rows (odd widths: 126 columns, 128 lanes, bands of 2 rows).  None for
// 64-bit offsets.
bool fused_energy_plan(const Geometry& g, int device, const PlanOverrides& ov,
                       const FusedPlan& base, FusedPlan* plan) {
    if (needs_index64(g, ov) || base.ty < 1) return false;
    FusedPlan e = base;
    e.pf = 0;
    int max_lds = 0;
    if (hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, device) !=
        hipSuccess)
        return false;
    if (g.w % 2 != 0) {
        e.vx = 1;
        if (base.tx == 0 && 3 * g.w <= EN_THREADS) {
            e.tx = 0;
            e.ty = std::min(base.ty, EN_THREADS / g.w - 2);
        } else {
            e.tx = 126;
            e.ty = EN_THREADS / (e.tx + 2) - 2;
            if (e.tx >= g.w) return false;
        }
        e.ty = std::min(e.ty, g.h);
        if (e.ty < 1) return false;
        const int row = e.tx > 0 ? e.tx + 2 : g.w;
        if (fused_lds_floats<1>(row, e.ty, 1) * sizeof(float) > size_t(max_lds)) return false;
        *plan = e;
        return true;
    }
    e.vx = 2;
    if (base.tx == 0 && 3 * (g.w / 2) <= EN_THREADS) {
        e.tx = 0;
        e.ty = std::min(base.ty, EN_THREADS / (g.w / 2) - 2);
    } else {
        e.tx = base.tx > 0 && base.vx == 2 ? base.tx : 124;
        if (e.tx >= g.w) e.tx = 124;
        e.ty = std::min(std::min(base.tx > 0 ? base.ty : 6, 6), EN_THREADS / (e.tx / 2 + 2) - 2);
    }
    e.ty = std::min(e.ty, g.h);
    if (e.ty < 1) return false;
    const int row = e.tx > 0 ? e.tx + 4 : g.w;
    if (fused_lds_floats<2>(row, e.ty, 1) * sizeof(float) > size_t(max_lds)) return false;
    *plan = e;
    return true;
}

size_t fused_grid_blocks(const Geometry& g, const FusedPlan& plan) {
    const size_t nbx = plan.tx > 0 ? size_t((g.w + plan.tx - 1) / plan.tx) : 1;
    return nbx * size_t((g.h + plan.ty - 1) / plan.ty) * size_t((g.d + plan.zchunk - 1) / plan.zchunk);
}

// Per tile of a fused plan (the kernel's tile index t): 1 if any voxel the
// tile owns is active.  One block per tile, an OR over its rows and planes.
__global__ void tile_activity_kernel(const uint8_t* __restrict__ act, Geometry g, int ty,
                                     int zchunk, int tx, uint8_t* __restrict__ out) {
    const int t = int(blockIdx.x);
    const int nbx = tx > 0 ? (g.w + tx - 1) / tx : 1, nb = (g.h + ty - 1) / ty;
    const int tb = tx > 0 ? t / nbx : t;
    const int x0 = tx > 0 ? (t % nbx) * tx : 0, x1 = tx > 0 ? min(x0 + tx, g.w) : g.w;
    const int y0 = (tb % nb) * ty, y1 = min(y0 + ty, g.h);
    const int z0 = (tb / nb) * zchunk, z1 = min(z0 + zchunk, g.d);
    const int rw = x1 - x0;
    const long cnt = long(y1 - y0) * (z1 - z0) * rw;
    int any = 0;
    for (long k = threadIdx.x; k < cnt && !any; k += blockDim.x) {
        const int x = x0 + int(k % rw);
        const long r = k / rw;
        const int y = y0 + int(r % (y1 - y0)), z = z0 + int(r / (y1 - y0));
        any = act[(size_t(z) * g.h + y) * g.w + x] != 0;
    }
    any = __syncthreads_or(any);
    if (threadIdx.x == 0) out[t] = uint8_t(any ? 1 : 0);
}

hipError_t launch_tile_activity(hipStream_t s, const uint8_t* act, Geometry g,
                                const FusedPlan& plan, uint8_t* out) {
    const size_t nt = fused_grid_blocks(g, plan);
    hipLaunchKernelGGL(tile_activity_kernel, dim3(unsigned(nt)), dim3(256), 0, s, act, g, plan.ty,
                       plan.zchunk, plan.tx, out);
    return hipGetLastError();
}

size_t fused_lds_bytes(const Geometry& g, const FusedPlan& plan) {
    const int row = plan.tx > 0 ? plan.tx + 2 * plan.vx : g.w;
    return (plan.vx == 1   ? fused_lds_floats<1>(row, plan.ty)
            : plan.vx == 2 ? fused_lds_floats<2>(row, plan.ty)
                           : fused_lds_floats<4>(row, plan.ty)) *
           sizeof(float);
}

const int FUSED_FIELD_MAP[N_FUSED_FIELDS] = {F_U, F_VX, F_VY, F_VZ, F_PX, F_PY, F_PZ,
                                             F_QXX, F_QYY, F_QZZ, F_QXY, F_QXZ, F_QYZ};

hipError_t launch_fused_sweep(hipStream_t s, int storage, void* const* src, void* const* dst13,
                              const float* f, Geometry g, const SweepParams& p,
                              const FusedPlan& plan) {
    static_assert(int(N_GFIELDS) == N_FUSED_FIELDS, "fused field count");
    FusedFields ff{};
    for (int k = 0; k < N_GFIELDS; ++k) {
        ff.src[k] = src[FUSED_FIELD_MAP[k]];
        ff.dst[k] = dst13[k];
    }
    const bool small = !p.index64;
    if (p.energy_partials) {
        if (!small || p.planar || p.planar_zero_z || plan.pf) return hipErrorInvalidValue;
        return storage == 1 ? fused_energy_dispatch<StoreBF16>(s, ff, f, g, p, plan)
                            : fused_energy_dispatch<StoreF32>(s, ff, f, g, p, plan);
    }
    if (storage == 1)
        return small ? fused_dispatch<StoreBF16, uint32_t>(s, ff, f, g, p, plan)
                     : fused_dispatch<StoreBF16, size_t>(s, ff, f, g, p, plan);
    return small ? fused_dispatch<StoreF32, uint32_t>(s, ff, f, g, p, plan)
                 : fused_dispatch<StoreF32, size_t>(s, ff, f, g, p, plan);
}

bool fused_divisors_ok(float alpha0, float alpha1) {
    return div_rn_divisor_ok(alpha0) && div_rn_divisor_ok(alpha1);
}

bool fused_bars_supported(int storage, Geometry g, const SweepParams& p, const FusedPlan& plan) {
    return storage == 0 && !p.index64 && !p.active && !p.planar && !p.planar_zero_z &&
           p.data_term == 1 && (plan.vx == 2 || plan.vx == 1) && g.d > 1;
}

hipError_t launch_fused_sweep_bars(hipStream_t s, void* const* src, void* const* dst,
                                   const float* f, Geometry g, const SweepParams& p,
                                   const FusedPlan& plan) {
    if (!fused_bars_supported(0, g, p, plan) || p.energy_partials) return hipErrorInvalidValue;
    FusedFields ff{};
    for (int k = 0; k < N_GFIELDS; ++k) {
        ff.src[k] = src[FUSED_FIELD_MAP[k]];
        ff.dst[k] = dst[FUSED_FIELD_MAP[k]];
    }
    ff.bars[0] = dst[F_UBAR];
    for (int k = 0; k < 3; ++k) ff.bars[1 + k] = dst[F_VBX + k];
    return fused_bars_dispatch(s, ff, f, g, p, plan);
}

// The 2-D loop marched along the image's rows (fused_sweep_kernel with P2 = 2):
// the 1 x H x W image as an H x 1 x W volume, the fields' y and z roles
// exchanged.  Used for images wider than 682 pixels (narrower ones run
// whole-row bands, which measured faster: 512^2 11.7 vs 15.0 us per launch).
// Plan: blocks of one row of two-voxel lanes, a whole number of waves: tiles of
// 124 columns (64 lanes with the halo lane on each side; images up to 2048
// wide) or 508 columns (256 lanes), and chunks of image rows for about 24
// blocks per CU, 4 to 6 rows.  Measured per launch: 1024^2 26-28 us (tiled
// depth-1 bands 38, two sweeps 43); 2048^2 73 us (best of 252..1020 columns x
// 4..64 rows: 71); 4096^2 261 us with 508 columns x 6 rows (the best; 380:
// 267, 252: 271, 1020: 268; 12 rows 277; 512, i.e. 258 lanes: 390; tiled
// depth-1 bands 454, two sweeps 545); 8192^2 1000 us with 6 rows (the best;
// 22 rows, 24 blocks per CU: 1123).
bool fused_image_plan(const Geometry& g, int device, const PlanOverrides& ov, FusedPlan* plan) {
    if (g.d != 1 || g.w % 2 != 0 || needs_index64(g, ov)) return false;
    if (ov.fused_image == 0) return false;
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess)
        return false;
    const bool want_tx = ov.fused_image_tx >= 0;   // 0: whole rows
    if (!want_tx && 3 * (g.w / 2) <= 1024) return false;   // whole-row bands fit
    int tx = g.w <= 2048 ? 124 : 508;
    if (want_tx) tx = ov.fused_image_tx >= g.w ? 0 : ov.fused_image_tx;
    if (tx < 0 || tx % 2 != 0 || (tx == 0 && g.w / 2 > 768)) return false;
    const long nbx = tx ? (g.w + tx - 1) / tx : 1;
    int zc = int(std::min(6L, std::max(4L, long(g.h) * nbx / (24L * cus))));
    zc = std::min(zc, g.h);
    if (ov.fused_image_zchunk > 0) zc = std::min(ov.fused_image_zchunk, g.h);
    *plan = FusedPlan{1, zc, 1, 2, tx};
    return true;
}

hipError_t launch_fused_image(hipStream_t s, void* const* src, void* const* dst13, const float* f,
                              Geometry g, const SweepParams& p, const FusedPlan& plan) {
    // fused order u, vx, vy, vz, px, py, pz, qxx, qyy, qzz, qxy, qxz, qyz with
    // y and z exchanged: the kernel's y slots get the image's (zero) z fields
    static const int swap[N_GFIELDS] = {0, 1, 3, 2, 4, 6, 5, 7, 9, 8, 11, 10, 12};
    FusedFields ff{};
    for (int k = 0; k < N_GFIELDS; ++k) {
        ff.src[k] = src[FUSED_FIELD_MAP[swap[k]]];
        ff.dst[k] = dst13[swap[k]];
    }
    const Geometry gi{g.w, 1, g.h};
    SweepParams pi = p;
    pi.planar = 0;
    typedef StoreF32 S;
    typedef uint32_t I;
    if (plan.tx > 0) {
        if (p.active)
            return p.data_term == 1 ? fused_launch<S, I, 2, 1, 1, 768, 1, 1, 2>(s, ff, f, gi, pi, plan)
                                    : fused_launch<S, I, 2, 0, 1, 768, 1, 1, 2>(s, ff, f, gi, pi, plan);
        return p.data_term == 1 ? fused_launch<S, I, 2, 1, 0, 768, 1, 1, 2>(s, ff, f, gi, pi, plan)
                                : fused_launch<S, I, 2, 0, 0, 768, 1, 1, 2>(s, ff, f, gi, pi, plan);
    }
    if (p.active)
        return p.data_term == 1 ? fused_launch<S, I, 2, 1, 1, 768, 1, 0, 2>(s, ff, f, gi, pi, plan)
                                : fused_launch<S, I, 2, 0, 1, 768, 1, 0, 2>(s, ff, f, gi, pi, plan);
    return p.data_term == 1 ? fused_launch<S, I, 2, 1, 0, 768, 1, 0, 2>(s, ff, f, gi, pi, plan)
                            : fused_launch<S, I, 2, 0, 0, 768, 1, 0, 2>(s, ff, f, gi, pi, plan);
}

hipError_t launch_dual_sweep(hipStream_t s, int storage, void* const* fields, Geometry g,
                             const SweepParams& p) {
    const bool small = !p.index64;
    if (storage == 1)
        return small ? dual_tiles<StoreBF16, uint32_t>(s, fields, g, p)
                     : dual_tiles<StoreBF16, size_t>(s, fields, g, p);
    return small ? dual_tiles<StoreF32, uint32_t>(s, fields, g, p)
                 : dual_tiles<StoreF32, size_t>(s, fields, g, p);
}

hipError_t launch_primal_sweep(hipStream_t s, int storage, void* const* fields, const float* f,
                               Geometry g, const SweepParams& p) {
    const bool small = !p.index64;
    if (storage == 1)
        return small ? primal_tiles<StoreBF16, uint32_t>(s, fields, f, g, p)
                     : primal_tiles<StoreBF16, size_t>(s, fields, f, g, p);
    return small ? primal_tiles<StoreF32, uint32_t>(s, fields, f, g, p)
                 : primal_tiles<StoreF32, size_t>(s, fields, f, g, p);
}

// Fields are allocated as one block: field k starts at fields[0] + k * stride.
// init writes u = u_bar = f and zeroes the 15 fields from v_x on (A3:
// reference tgv_common.cu:36-63, tgv2_common.cu:357-371).
hipError_t launch_init(hipStream_t s, int storage, void* const* fields, const float* f, Geometry g) {
    const size_t n = g.n();
    const char* base = static_cast<const char*>(fields[F_VX]);
    const char* last = static_cast<const char*>(fields[F_QYZ]);
    const size_t elem_bytes = storage == 1 ? 2 : 4;
    const size_t zero_count = size_t(last - base) / elem_bytes + n;
    const int blocks = 2048;
    if (storage == 1)
        hipLaunchKernelGGL(init_kernel<StoreBF16>, dim3(blocks), dim3(256), 0, s,
                           static_cast<uint16_t*>(fields[F_U]), static_cast<uint16_t*>(fields[F_UBAR]),
                           static_cast<uint16_t*>(fields[F_VX]), zero_count, f, n);
    else
        hipLaunchKernelGGL(init_kernel<StoreF32>, dim3(blocks), dim3(256), 0, s,
                           static_cast<float*>(fields[F_U]), static_cast<float*>(fields[F_UBAR]),
                           static_cast<float*>(fields[F_VX]), zero_count, f, n);
    return hipGetLastError();
}

// Zero fill by a kernel, not hipMemsetAsync, for the clears inside the
// iteration calls (the 2-D loop's second z copies, the TGV-k block, the
// adjoint's accumulators): their work on the context's stream is then kernel
// launches and copies only, with no memset nodes should it ever be captured
// into a HIP graph (the calls synchronise at their end, so none is today).
template <bool VEC>
__global__ void zero_words_kernel(uint32_t* p, size_t nwords) {
    const size_t stride = size_t(gridDim.x) * blockDim.x;
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    size_t head = 0;
    if (VEC) {
        const size_t n4 = nwords / 4;
        uint4* p4 = reinterpret_cast<uint4*>(p);
        for (size_t k = i; k < n4; k += stride) p4[k] = make_uint4(0u, 0u, 0u, 0u);
        head = n4 * 4;
    }
    for (size_t k = head + i; k < nwords; k += stride) p[k] = 0u;
}

hipError_t launch_zero_fill(hipStream_t s, void* p, size_t bytes) {
    if (bytes == 0) return hipSuccess;
    if (bytes % 4 != 0 || reinterpret_cast<uintptr_t>(p) % 4 != 0) return hipErrorInvalidValue;
    const size_t nwords = bytes / 4;
    const bool vec = reinterpret_cast<uintptr_t>(p) % 16 == 0;
    const size_t units = vec ? (nwords + 3) / 4 : nwords;
    const int blocks = int(std::min<size_t>((units + 255) / 256, 4096));
    if (vec)
        hipLaunchKernelGGL(zero_words_kernel<true>, dim3(blocks), dim3(256), 0, s,
                           static_cast<uint32_t*>(p), nwords);
    else
        hipLaunchKernelGGL(zero_words_kernel<false>, dim3(blocks), dim3(256), 0, s,
                           static_cast<uint32_t*>(p), nwords);
    return hipGetLastError();
}

// to_storage = 1: float -> storage ; 0: storage -> float
hipError_t launch_convert(hipStream_t s, int storage, const void* src, void* dst, size_t n,
                          int to_storage) {
    const int blocks = 1024;
    if (storage == 1) {
        if (to_storage)
            hipLaunchKernelGGL((convert_kernel<StoreF32, StoreBF16>), dim3(blocks), dim3(256), 0, s,
                               static_cast<const float*>(src), static_cast<uint16_t*>(dst), n);
        else
            hipLaunchKernelGGL((convert_kernel<StoreBF16, StoreF32>), dim3(blocks), dim3(256), 0, s,
                               static_cast<const uint16_t*>(src), static_cast<float*>(dst), n);
    } else {
        hipLaunchKernelGGL((convert_kernel<StoreF32, StoreF32>), dim3(blocks), dim3(256), 0, s,
                           static_cast<const float*>(src), static_cast<float*>(dst), n);
    }
    return hipGetLastError();
}

hipError_t ensure_dynamic_lds(const void* kernel, size_t bytes) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    static std::mutex mu;
    static std::map<std::pair<int, const void*>, size_t> raised;
    std::lock_guard<std::mutex> lock(mu);
    size_t& have = raised[{dev, kernel}];
    if (bytes <= have) return hipSuccess;
    e = hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize, int(bytes));
    if (e == hipSuccess) have = bytes;
    return e;
}

}  // namespace amb
