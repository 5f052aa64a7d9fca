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

namespace amb {
namespace {

constexpr int TX = 64;
constexpr int TY = 4;
constexpr int NT = TX * TY;

// Block id -> tile id such that each XCD (blocks are dealt round-robin to the
// 8 XCDs) receives a contiguous range of tiles.
__device__ __forceinline__ int xcd_tile(int bid, int nblocks) {
    if ((nblocks & 7) != 0) return bid;
    return (bid & 7) * (nblocks >> 3) + (bid >> 3);
}

struct TileCoord {
    int x0, y0, z0, z1, tx, ty, x, y;
};

__device__ __forceinline__ TileCoord tile_coord(const Geometry& g, int zchunk) {
    const int nbx = (g.w + TX - 1) / TX, nby = (g.h + TY - 1) / TY;
    const int t = xcd_tile(blockIdx.x, gridDim.x);
    TileCoord c;
    const int bx = t % nbx, r = t / nbx, by = r % nby, bz = r / nby;
    c.x0 = bx * TX; c.y0 = by * TY;
    c.z0 = bz * zchunk; c.z1 = min(c.z0 + zchunk, g.d);
    c.tx = threadIdx.x % TX; c.ty = threadIdx.x / TX;
    c.x = c.x0 + c.tx; c.y = c.y0 + c.ty;
    return c;
}

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
template <class S>
__global__ __launch_bounds__(NT) void dual_sweep_kernel(
    typename S::T* __restrict__ ubar, typename S::T* __restrict__ vbx,
    typename S::T* __restrict__ vby, typename S::T* __restrict__ vbz,
    typename S::T* __restrict__ px_, typename S::T* __restrict__ py_,
    typename S::T* __restrict__ pz_, typename S::T* __restrict__ qxx_,
    typename S::T* __restrict__ qyy_, typename S::T* __restrict__ qzz_,
    typename S::T* __restrict__ qxy_, typename S::T* __restrict__ qxz_,
    typename S::T* __restrict__ qyz_, Geometry g, float sigma, float alpha0, float alpha1,
    int zchunk) {
    // ubar plane with +x/+y halo, v_bar planes with -x/-y halo; centre at [ty+1][tx+1].
    __shared__ float sU[TY + 2][TX + 2];
    __shared__ float sV[3][TY + 2][TX + 2];

    const TileCoord c = tile_coord(g, zchunk);
    const int W = g.w, H = g.h, D = g.d;
    const size_t plane = size_t(W) * size_t(H);
    const bool in = (c.x < W) && (c.y < H);
    const size_t col = in ? size_t(c.y) * W + c.x : 0;
    const bool hx = (c.tx == 0) && (c.y < H);    // loads the x halo of its row
    const bool hy = (c.ty == 0) && (c.x < W);    // loads the y halo of its column
    const bool has_xp = c.x0 + TX < W, has_xm = c.x0 > 0;
    const bool has_yp = c.y0 + TY < H, has_ym = c.y0 > 0;
    const size_t hx_p = size_t(c.y) * W + (c.x0 + TX), hx_m = size_t(c.y) * W + (c.x0 - 1);
    const size_t hy_p = size_t(c.y0 + TY) * W + c.x, hy_m = size_t(c.y0 - 1) * W + c.x;

    typename S::T* const vb[3] = {vbx, vby, vbz};

    // z-pipeline registers: u_bar (z, z+1), v_bar (z-1, z), p and q (z)
    float ub_c = 0.f, ub_n = 0.f, vb_p[3] = {0.f, 0.f, 0.f}, vb_c[3] = {0.f, 0.f, 0.f};
    float pc[3] = {0.f, 0.f, 0.f}, qc[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    typename S::T* const pf[3] = {px_, py_, pz_};
    typename S::T* const qf[6] = {qxx_, qyy_, qzz_, qxy_, qxz_, qyz_};
    int z = c.z0;
    if (in) {
        const size_t i0 = size_t(z) * plane + col;
        ub_c = S::ld(ubar, i0);
        if (z + 1 < D) ub_n = S::ld(ubar, i0 + plane);
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            vb_c[k] = S::ld(vb[k], i0);
            if (z > 0) vb_p[k] = S::ld(vb[k], i0 - plane);
            pc[k] = S::ld(pf[k], i0);
        }
#pragma unroll
        for (int k = 0; k < 6; ++k) qc[k] = S::ld(qf[k], i0);
    }

    for (; z < c.z1; ++z) {
        const size_t zo = size_t(z) * plane;
        const size_t i = zo + col;
        // stage this plane in LDS
        sU[c.ty + 1][c.tx + 1] = ub_c;
#pragma unroll
        for (int k = 0; k < 3; ++k) sV[k][c.ty + 1][c.tx + 1] = vb_c[k];
        if (hx) {
            sU[c.ty + 1][TX + 1] = has_xp ? S::ld(ubar, zo + hx_p) : 0.f;
#pragma unroll
            for (int k = 0; k < 3; ++k) sV[k][c.ty + 1][0] = has_xm ? S::ld(vb[k], zo + hx_m) : 0.f;
        }
        if (hy) {
            sU[TY + 1][c.tx + 1] = has_yp ? S::ld(ubar, zo + hy_p) : 0.f;
#pragma unroll
            for (int k = 0; k < 3; ++k) sV[k][0][c.tx + 1] = has_ym ? S::ld(vb[k], zo + hy_m) : 0.f;
        }
        __syncthreads();

        // prefetch the next plane
        float ub_nn = 0.f, vb_n[3] = {0.f, 0.f, 0.f}, pn[3] = {0.f, 0.f, 0.f};
        float qn[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
        if (in && z + 1 < c.z1) {
            if (z + 2 < D) ub_nn = S::ld(ubar, i + 2 * plane);
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                vb_n[k] = S::ld(vb[k], i + plane);
                pn[k] = S::ld(pf[k], i + plane);
            }
#pragma unroll
            for (int k = 0; k < 6; ++k) qn[k] = S::ld(qf[k], i + plane);
        }
        const float p0 = pc[0], p1 = pc[1], p2 = pc[2];
        const float q0 = qc[0], q1 = qc[1], q2 = qc[2], q3 = qc[3], q4 = qc[4], q5 = qc[5];

        if (in) {
            // D+ u_bar
            const float dx = fwd(ub_c, sU[c.ty + 1][c.tx + 2], c.x < W - 1);
            const float dy = fwd(ub_c, sU[c.ty + 2][c.tx + 1], c.y < H - 1);
            const float dz = fwd(ub_c, ub_n, z < D - 1);
            // p update + projection onto |p| <= alpha1 (tgv2_common.cu:118-128)
            float ax = fmaf(sigma, dx - vb_c[0], p0);
            float ay = fmaf(sigma, dy - vb_c[1], p1);
            float az = fmaf(sigma, dz - vb_c[2], p2);
            float n2 = ax * ax;
            n2 = fmaf(ay, ay, n2);
            n2 = fmaf(az, az, n2);
            const float np = fmaxf(1.0f, sqrtf(n2) / alpha1);
            S::st(px_, i, ax / np);
            S::st(py_, i, ay / np);
            S::st(pz_, i, az / np);

            // eps_B(v_bar) (tgv2_common.cu:174-219): B_a applied to v_bar_c
            float bxv[3], byv[3], bzv[3];
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                bxv[k] = bwd(vb_c[k], sV[k][c.ty + 1][c.tx], c.x > 0);
                byv[k] = bwd(vb_c[k], sV[k][c.ty][c.tx + 1], c.y > 0);
                bzv[k] = bwd(vb_c[k], vb_p[k], z > 0);
            }
            const float exx = bxv[0], eyy = byv[1], ezz = bzv[2];
            const float exy = (bxv[1] + byv[0]) * 0.5f;
            const float exz = (bxv[2] + bzv[0]) * 0.5f;
            const float eyz = (byv[2] + bzv[1]) * 0.5f;
            // q update + projection onto |q|_w <= alpha0 (tgv2_common.cu:246-270)
            const float bxx = fmaf(sigma, exx, q0);
            const float byy = fmaf(sigma, eyy, q1);
            const float bxy = fmaf(sigma, exy, q3);
            const float bzz = fmaf(sigma, ezz, q2);
            const float bxz = fmaf(sigma, exz, q4);
            const float byz = fmaf(sigma, eyz, q5);
            float m2 = bxx * bxx;
            m2 = fmaf(byy, byy, m2);
            m2 = fmaf(2.0f * bxy, bxy, m2);
            m2 = fmaf(bzz, bzz, m2);
            m2 = fmaf(2.0f * bxz, bxz, m2);
            m2 = fmaf(2.0f * byz, byz, m2);
            const float nq = fmaxf(1.0f, sqrtf(m2) / alpha0);
            S::st(qxx_, i, bxx / nq);
            S::st(qyy_, i, byy / nq);
            S::st(qxy_, i, bxy / nq);
            S::st(qzz_, i, bzz / nq);
            S::st(qxz_, i, bxz / nq);
            S::st(qyz_, i, byz / nq);
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

template <class S, int L2>
__global__ __launch_bounds__(NT) void primal_sweep_kernel(
    typename S::T* __restrict__ u_, typename S::T* __restrict__ ubar_,
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

    const TileCoord c = tile_coord(g, zchunk);
    const int W = g.w, H = g.h, D = g.d;
    const size_t plane = size_t(W) * size_t(H);
    const bool in = (c.x < W) && (c.y < H);
    const size_t col = in ? size_t(c.y) * W + c.x : 0;
    const bool hx = (c.tx == 0) && (c.y < H);
    const bool hy = (c.ty == 0) && (c.x < W);
    const bool has_xp = c.x0 + TX < W, has_xm = c.x0 > 0;
    const bool has_yp = c.y0 + TY < H, has_ym = c.y0 > 0;
    const size_t hx_p = size_t(c.y) * W + (c.x0 + TX), hx_m = size_t(c.y) * W + (c.x0 - 1);
    const size_t hy_p = size_t(c.y0 + TY) * W + c.x, hy_m = size_t(c.y0 - 1) * W + c.x;
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
        const size_t i0 = size_t(z) * plane + col;
        if (z > 0) pz_p = S::ld(pz_, i0 - plane);
        pz_c = S::ld(pz_, i0);
        qzz_c = S::ld(qzz_, i0); qxz_c = S::ld(qxz_, i0); qyz_c = S::ld(qyz_, i0);
        if (z + 1 < D) {
            qzz_n = S::ld(qzz_, i0 + plane); qxz_n = S::ld(qxz_, i0 + plane);
            qyz_n = S::ld(qyz_, i0 + plane);
        }
        px_c = S::ld(px_, i0); py_c = S::ld(py_, i0);
        qxx_c = S::ld(qxx_, i0); qxy_c = S::ld(qxy_, i0); qyy_c = S::ld(qyy_, i0);
        u_c = S::ld(u_, i0); f_c = f_[i0];
        v_c[0] = S::ld(vx_, i0); v_c[1] = S::ld(vy_, i0); v_c[2] = S::ld(vz_, i0);
    }

    for (; z < c.z1; ++z) {
        const size_t zo = size_t(z) * plane;
        const size_t i = zo + col;
        const int ly = c.ty + 1, lx = c.tx + 1;
        sL[L_PX][ly][lx] = px_c;
        sL[L_PY][ly][lx] = py_c;
        sL[L_QXX][ly][lx] = qxx_c;
        sL[L_QXY][ly][lx] = qxy_c;
        sL[L_QXZ][ly][lx] = qxz_c;
        sL[L_QYY][ly][lx] = qyy_c;
        sL[L_QYZ][ly][lx] = qyz_c;
        if (hx) {
            sL[L_PX][ly][0] = has_xm ? S::ld(px_, zo + hx_m) : 0.f;
            sL[L_QXX][ly][TX + 1] = has_xp ? S::ld(qxx_, zo + hx_p) : 0.f;
            sL[L_QXY][ly][TX + 1] = has_xp ? S::ld(qxy_, zo + hx_p) : 0.f;
            sL[L_QXZ][ly][TX + 1] = has_xp ? S::ld(qxz_, zo + hx_p) : 0.f;
        }
        if (hy) {
            sL[L_PY][0][lx] = has_ym ? S::ld(py_, zo + hy_m) : 0.f;
            sL[L_QXY][TY + 1][lx] = has_yp ? S::ld(qxy_, zo + hy_p) : 0.f;
            sL[L_QYY][TY + 1][lx] = has_yp ? S::ld(qyy_, zo + hy_p) : 0.f;
            sL[L_QYZ][TY + 1][lx] = has_yp ? S::ld(qyz_, zo + hy_p) : 0.f;
        }
        __syncthreads();

        // prefetch the next plane
        float n_pz = 0.f, n_qzz = 0.f, n_qxz = 0.f, n_qyz = 0.f, n_px = 0.f, n_py = 0.f,
              n_qxx = 0.f, n_qxy = 0.f, n_qyy = 0.f, n_u = 0.f, n_f = 0.f;
        float n_v[3] = {0.f, 0.f, 0.f};
        if (in && z + 1 < c.z1) {
            const size_t j = i + plane;
            n_pz = S::ld(pz_, j);
            if (z + 2 < D) {
                n_qzz = S::ld(qzz_, j + plane); n_qxz = S::ld(qxz_, j + plane);
                n_qyz = S::ld(qyz_, j + plane);
            }
            n_px = S::ld(px_, j); n_py = S::ld(py_, j);
            n_qxx = S::ld(qxx_, j); n_qxy = S::ld(qxy_, j); n_qyy = S::ld(qyy_, j);
            n_u = S::ld(u_, j); n_f = f_[j];
            n_v[0] = S::ld(vx_, j); n_v[1] = S::ld(vy_, j); n_v[2] = S::ld(vz_, j);
        }

        if (in) {
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
            const float pc[3] = {px_c, py_c, pz_c};
            const float wv[3] = {w0, w1, w2};
            typename S::T* const vo[3] = {vx_, vy_, vz_};
            typename S::T* const vbo[3] = {vbx_, vby_, vbz_};
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                const float vn = fmaf(-tau, wv[k] - pc[k], v_c[k]);
                S::st(vo[k], i, vn);
                S::st(vbo[k], i, fmaf(theta, vn - v_c[k], vn));
            }
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

template <class S>
__global__ void init_kernel(typename S::T* const* fields_unused, typename S::T* u,
                            typename S::T* ubar, typename S::T* zero_base, size_t zero_count,
                            const float* f, size_t n) {
    (void)fields_unused;
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

inline int grid_blocks(const Geometry& g, int zchunk) {
    const int nbx = (g.w + TX - 1) / TX, nby = (g.h + TY - 1) / TY;
    const int nbz = (g.d + zchunk - 1) / zchunk;
    return nbx * nby * nbz;
}

template <class S>
hipError_t dual_impl(hipStream_t s, void* const* fv, Geometry g, const SweepParams& p) {
    typedef typename S::T T;
    T* const* F = reinterpret_cast<T* const*>(fv);
    hipLaunchKernelGGL(dual_sweep_kernel<S>, dim3(grid_blocks(g, p.zchunk)), dim3(NT), 0, s,
                       F[F_UBAR], F[F_VBX], F[F_VBY], F[F_VBZ], F[F_PX], F[F_PY], F[F_PZ],
                       F[F_QXX], F[F_QYY], F[F_QZZ], F[F_QXY], F[F_QXZ], F[F_QYZ], g, p.sigma,
                       p.alpha0, p.alpha1, p.zchunk);
    return hipGetLastError();
}

template <class S>
hipError_t primal_impl(hipStream_t s, void* const* fv, const float* f, Geometry g,
                       const SweepParams& p) {
    typedef typename S::T T;
    T* const* F = reinterpret_cast<T* const*>(fv);
    const dim3 grid(grid_blocks(g, p.zchunk)), block(NT);
    if (p.data_term == 1)
        hipLaunchKernelGGL((primal_sweep_kernel<S, 1>), grid, block, 0, s, F[F_U], F[F_UBAR],
                           F[F_VX], F[F_VY], F[F_VZ], F[F_VBX], F[F_VBY], F[F_VBZ], F[F_PX],
                           F[F_PY], F[F_PZ], F[F_QXX], F[F_QYY], F[F_QZZ], F[F_QXY], F[F_QXZ],
                           F[F_QYZ], f, g, p.tau, p.lambda, p.theta, p.zchunk);
    else
        hipLaunchKernelGGL((primal_sweep_kernel<S, 0>), grid, block, 0, s, F[F_U], F[F_UBAR],
                           F[F_VX], F[F_VY], F[F_VZ], F[F_VBX], F[F_VBY], F[F_VBZ], F[F_PX],
                           F[F_PY], F[F_PZ], F[F_QXX], F[F_QYY], F[F_QZZ], F[F_QXY], F[F_QXZ],
                           F[F_QYZ], f, g, p.tau, p.lambda, p.theta, p.zchunk);
    return hipGetLastError();
}

}  // namespace

hipError_t launch_dual_sweep(hipStream_t s, int storage, void* const* fields, Geometry g,
                             const SweepParams& p) {
    return storage == 1 ? dual_impl<StoreBF16>(s, fields, g, p) : dual_impl<StoreF32>(s, fields, g, p);
}

hipError_t launch_primal_sweep(hipStream_t s, int storage, void* const* fields, const float* f,
                               Geometry g, const SweepParams& p) {
    return storage == 1 ? primal_impl<StoreBF16>(s, fields, f, g, p)
                        : primal_impl<StoreF32>(s, fields, f, g, p);
}

// Fields are allocated as one block: field k starts at fields[0] + k * stride.
// init writes u = u_bar = f and zeroes the 15 fields from v_x on (A3:
// reference tgv_common.cu:36-63, tgv2_common.cu:357-371).
hipError_t launch_init(hipStream_t s, int storage, void* const* fields, const float* f, Geometry g) {
    const size_t n = g.n();
    const char* base = static_cast<const char*>(fields[F_VX]);
    const char* last = static_cast<const char*>(fields[F_QYZ]);
    const size_t elem = storage == 1 ? 2 : 4;
    const size_t zero_count = size_t(last - base) / elem + n;
    const int blocks = 2048;
    if (storage == 1)
        hipLaunchKernelGGL(init_kernel<StoreBF16>, dim3(blocks), dim3(256), 0, s, nullptr,
                           static_cast<uint16_t*>(fields[F_U]), static_cast<uint16_t*>(fields[F_UBAR]),
                           static_cast<uint16_t*>(fields[F_VX]), zero_count, f, n);
    else
        hipLaunchKernelGGL(init_kernel<StoreF32>, dim3(blocks), dim3(256), 0, s, nullptr,
                           static_cast<float*>(fields[F_U]), static_cast<float*>(fields[F_UBAR]),
                           static_cast<float*>(fields[F_VX]), zero_count, f, n);
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

}  // namespace amb
