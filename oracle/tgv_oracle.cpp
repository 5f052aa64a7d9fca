// CPU oracle for the TGV-DCT deshade hot path.
//
// TEST INFRASTRUCTURE ONLY.  This file restates, on the CPU, the algorithm of
// josefkoller/ambient_image_processor's TGV2 primal-dual iteration and DCT
// Poisson bias solve.  Only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg may load it, and only as the checker (or the timed CPU
// baseline); the product path (ambient_image_processor_amd) never does.
//
// Two numeric modes, selected by the template type T:
//   float  - "mirror order": the reference's FP32 operation sequence, with
//            every multiply-add that the reference writes as fmaf, or that
//            nvcc contracts by default, written as an explicit fma; every
//            other op rounded separately (built with -ffp-contract=off).
//            Division and sqrt are IEEE correctly rounded.  The DCT keeps
//            the reference's FP64 cosine * FP32 accumulator numerics.
//   double - the same formulas in FP64 throughout (the FP64 oracle).
//
// Layout (reference source/modules/itk/ITKImage.cpp:426-429): x fastest,
// idx = z*W*H + y*W + x; every field a separate dense array.
//
// Parity pinning (see tests/test_oracle_*.py):
//   stencils          - brute-force sparse matrices and the transpose identities
//   projections/prox  - closed forms and special cases
//   iteration         - invariants (constant image, translation, L2 scaling,
//                       zero iterations, callback stop) + FP32-vs-FP64 drift
//   DCT-II/III        - scipy.fft.dctn/idctn golden fixtures (tests/golden)
//   Poisson divide    - known answer: Neumann Laplacian of DCT modes
//   energy/gap        - Fenchel-Young non-negativity, closed forms on tiny inputs
// The FP32 rounding sequence of the CUDA device math the reference calls
// (norm3df, cospi, cospif) cannot be pinned without running the reference:
// parity unpinned at the last-ulp level for those calls (error-bound tests).

#include <cmath>
#include <cstdint>
#include <cstring>
#include <cstdlib>
#include <vector>
#include <algorithm>
#include <omp.h>

namespace {

typedef long long i64;

struct Dims {
    i64 w, h, d;
    i64 n() const { return w * h * d; }
    i64 stride(int axis) const { return axis == 0 ? 1 : (axis == 1 ? w : w * h); }
    i64 extent(int axis) const { return axis == 0 ? w : (axis == 1 ? h : d); }
};

inline i64 coord(const Dims& s, i64 i, int axis) {
    if (axis == 0) return i % s.w;
    if (axis == 1) return (i / s.w) % s.h;
    return i / (s.w * s.h);
}

// BF16 field storage (the GPU's amb_storage BF16): a field store keeps the
// value rounded to bfloat16, nearest with ties to even, the integer form
// (bits + 0x7fff + bit 16) >> 16 (NaN passed through).  Arithmetic stays FP32.
inline float bf16_rne(float x) {
    uint32_t b;
    std::memcpy(&b, &x, 4);
    if ((b & 0x7fffffffu) > 0x7f800000u) return x;
    b = (b + 0x7fffu + ((b >> 16) & 1u)) & 0xffff0000u;
    std::memcpy(&x, &b, 4);
    return x;
}
// The value a field holds after a store: FP32 / FP64 modes keep it, the BF16
// mode (float only) rounds it.
inline float stored(float x, int bf16) { return bf16 ? bf16_rne(x) : x; }
inline double stored(double x, int) { return x; }

inline float fma_t(float a, float b, float c) { return std::fmaf(a, b, c); }
inline double fma_t(double a, double b, double c) { return std::fma(a, b, c); }
inline float sqrt_t(float a) { return std::sqrt(a); }
inline double sqrt_t(double a) { return std::sqrt(a); }

// ---------------------------------------------------------------------------
// A4: forward difference, reference tgv_cuda/tgv_common.cu:66-130
//   (D+ a)[i] = a[i+e] - a[i] for i < n-1 ; 0 at i = n-1 ("neumann").
// A6: backward difference, reference tgv_cuda/tgv_common.cu:133-197
//   (B a)[0] = -a[0] ; (B a)[i] = -a[i] + a[i-1]   (note the sign).
// ---------------------------------------------------------------------------
template <typename T>
void forward_difference(const T* a, T* out, const Dims& s, int axis) {
    const i64 st = s.stride(axis), n = s.extent(axis);
#pragma omp parallel for schedule(static)
    for (i64 i = 0; i < s.n(); ++i) {
        const i64 c = coord(s, i, axis);
        out[i] = (c < n - 1) ? a[i + st] - a[i] : T(0);
    }
}

template <typename T>
void backward_difference(const T* a, T* out, const Dims& s, int axis) {
    const i64 st = s.stride(axis);
#pragma omp parallel for schedule(static)
    for (i64 i = 0; i < s.n(); ++i) {
        const i64 c = coord(s, i, axis);
        out[i] = (c > 0) ? -a[i] + a[i - st] : -a[i];
    }
}

// ---------------------------------------------------------------------------
// Iteration state: u, u_bar, v[3], v_bar[3], p[3], q[6] (q order xx,yy,zz,xy,xz,yz)
// ---------------------------------------------------------------------------
template <typename T>
struct State {
    T* u; T* ubar;
    T* v[3]; T* vbar[3];
    T* p[3];
    T* q[6];
};

struct Params {
    double lambda, alpha0, alpha1, tau, sigma, theta;
    int data_term;  // 0 = L1, 1 = L2
    int planar;     // depth 1: the 2-D loop (no z backward differences in eps_B(v_bar))
    int bf16 = 0;   // float only: fields stored as bfloat16 (every field store rounded)
};

template <typename T>
void fill_state_pointers(State<T>& st, T** fields) {
    st.u = fields[0]; st.ubar = fields[1];
    for (int c = 0; c < 3; ++c) { st.v[c] = fields[2 + c]; st.vbar[c] = fields[5 + c]; st.p[c] = fields[8 + c]; }
    for (int c = 0; c < 6; ++c) st.q[c] = fields[11 + c];
}

// A3: initialisation, reference tgv_cuda/tgv_common.cu:36-63,
// tgv_cuda/tgv2_common.cu:357-371, launched tgv_deshade/tgv2_l1_deshade.cu:153-167
//   u = u_bar = f ; p = v = v_bar = q = 0.
// bf16: u = u_bar = f rounded to bfloat16 (f itself stays FP32).
template <typename T>
void tgv2_init(State<T>& st, const T* f, const Dims& s, int bf16 = 0) {
#pragma omp parallel for schedule(static)
    for (i64 i = 0; i < s.n(); ++i) {
        st.u[i] = st.ubar[i] = stored(f[i], bf16);
        for (int c = 0; c < 3; ++c) st.v[c][i] = st.vbar[c][i] = st.p[c][i] = T(0);
        for (int c = 0; c < 6; ++c) st.q[c][i] = T(0);
    }
}

// One iteration of the reference loop, tgv_deshade/tgv2_l1_deshade.cu:169-264
// (and tgv_cuda/tgv2_l2.cu:194-… with the L2 u-update), kernel by kernel,
// with the reference's temporaries.
// active (may be null): masked deshade, reference tgv_deshade_masked/tgv2_l1_deshade_masked.cu:161-272
// - the update kernels run only over in-mask voxels (mask >= 1,
// TGVDeshadeMaskedProcessor.cpp:337-372); the stencils are the unmasked ones
// (the edge lists are the volume faces), so out-of-mask voxels keep their
// initial values and are read as neighbours.
// wbar0 (may be null): TGV-k's q-update q + sigma*(eps_B(v_bar) - w_bar_0)
// (tgv3/tgv3_common.cu:28-86, launched tgvk_deshade/tgvk_l1_deshade.cu:200-214).
// prm.bf16: the stores of the reference's update kernels - p (tgv2_common.cu:124-128),
// q and v_prev (:262-275), v and v_bar (:338-347), u and u_bar (tgv2_l1.cu:39-50,
// tgv2_l2.cu:44-46) - keep the bfloat16 rounding of the value; u_bar and v_bar are
// formed from the unrounded new u and v, and u_prev / v_prev are the stored u, v.
template <typename T>
void tgv2_iteration(State<T>& st, const T* f, const Dims& s, const Params& prm,
                    std::vector<T>* tmp /* 9 arrays of n */, const unsigned char* active,
                    T* const* wbar0 = nullptr) {
    const i64 n = s.n();
    const T tau = T(prm.tau), sigma = T(prm.sigma), theta = T(prm.theta);
    const T lambda = T(prm.lambda), alpha0 = T(prm.alpha0), alpha1 = T(prm.alpha1);
    const int rb = prm.bf16;
    T* gx = tmp[0].data(); T* gy = tmp[1].data(); T* gz = tmp[2].data();
    T* uprev = tmp[3].data();
    T* vprev[3] = {tmp[4].data(), tmp[5].data(), tmp[6].data()};
    T* t1 = tmp[7].data(); T* t2 = tmp[8].data();

    // A4 on u_bar (tgv2_l1_deshade.cu:171-177)
    forward_difference(st.ubar, gx, s, 0);
    forward_difference(st.ubar, gy, s, 1);
    forward_difference(st.ubar, gz, s, 2);

    // A5 tgv_kernel_part22, tgv_cuda/tgv2_common.cu:96-131
#pragma omp parallel for schedule(static)
    for (i64 i = 0; i < n; ++i) {
        if (active && !active[i]) continue;
        T px = fma_t(sigma, gx[i] - st.vbar[0][i], st.p[0][i]);
        T py = fma_t(sigma, gy[i] - st.vbar[1][i], st.p[1][i]);
        T pz = fma_t(sigma, gz[i] - st.vbar[2][i], st.p[2][i]);
        T n2 = px * px;
        n2 = fma_t(py, py, n2);
        n2 = fma_t(pz, pz, n2);
        const T norm = sqrt_t(n2);
        const T nrm = std::max(T(1), norm / alpha1);
        st.p[0][i] = stored(px / nrm, rb); st.p[1][i] = stored(py / nrm, rb);
        st.p[2][i] = stored(pz / nrm, rb);
        uprev[i] = st.u[i];
    }

    // A6 on p (tgv2_l1_deshade.cu:186-193)
    backward_difference(st.p[0], gx, s, 0);
    backward_difference(st.p[1], gy, s, 1);
    backward_difference(st.p[2], gz, s, 2);

    // A7 tgv_kernel_part4_tgv2_l1, tgv_cuda/tgv2_l1.cu:27-60
    // A8 tgv_kernel_part4_tgv2_l2, tgv_cuda/tgv2_l2.cu:27-56 (depth==1 drops B_z p_z)
    const T tl = tau * lambda;
#pragma omp parallel for schedule(static)
    for (i64 i = 0; i < n; ++i) {
        if (active && !active[i]) continue;
        T div = (s.d > 1 || prm.data_term == 0) ? (gx[i] + gy[i]) + gz[i] : gx[i] + gy[i];
        T u = fma_t(-tau, div, st.u[i]);
        if (prm.data_term == 0) {
            const T umf = u - f[i];
            if (umf > tl) u = u - tl;
            else if (umf < -tl) u = u + tl;
            else u = f[i];
        } else {
            u = fma_t(tl, f[i], u) / (T(1) + tl);
        }
        st.u[i] = stored(u, rb);
        st.ubar[i] = stored(fma_t(theta, u - uprev[i], u), rb);
    }

    // A9 tgv_launch_gradient2, tgv_cuda/tgv2_common.cu:174-219 (+ addAndHalf :31-40)
    // epsilon_B(v_bar) into the six temporaries named q_*2 by the caller.
    // Stored into tmp slots reusing gx..t2 is not enough room, so compute
    // the six components and apply A10 in one pass per voxel below.
    std::vector<T> e[6];
    for (int c = 0; c < 6; ++c) e[c].assign(n, T(0));
    backward_difference(st.vbar[0], e[0].data(), s, 0);            // xx
    backward_difference(st.vbar[1], e[1].data(), s, 1);            // yy
    backward_difference(st.vbar[1], t1, s, 0);                     // B_x v_y
    backward_difference(st.vbar[0], t2, s, 1);                     // B_y v_x
#pragma omp parallel for schedule(static)
    for (i64 i = 0; i < n; ++i) e[3][i] = (t1[i] + t2[i]) * T(0.5);     // xy
    // The 2-D loop (tgv_cuda/tgv2_common_2d.cu:151-172, launched by
    // tgv_deshade/tgv2_l1_deshade_2d.cu:112-252) has no z backward differences:
    // planar replaces every B_z term by 0, so e_zz, e_xz, e_yz and with them
    // q_z*, v_z, p_z stay 0 and the 3-D kernels reduce to the 2-D ones.
    auto bz = [&](const T* a, T* out) {
        if (prm.planar) std::fill(out, out + n, T(0));
        else backward_difference(a, out, s, 2);
    };
    bz(st.vbar[2], e[2].data());                                   // zz
    backward_difference(st.vbar[2], t1, s, 0);                     // B_x v_z
    bz(st.vbar[0], t2);                                            // B_z v_x
#pragma omp parallel for schedule(static)
    for (i64 i = 0; i < n; ++i) e[4][i] = (t1[i] + t2[i]) * T(0.5);     // xz
    backward_difference(st.vbar[2], t1, s, 1);                     // B_y v_z
    bz(st.vbar[1], t2);                                            // B_z v_y
#pragma omp parallel for schedule(static)
    for (i64 i = 0; i < n; ++i) e[5][i] = (t1[i] + t2[i]) * T(0.5);     // yz

    // A10 tgv_kernel_part5, tgv_cuda/tgv2_common.cu:222-276
    // (caller swaps the argument names, tgv2_l1_deshade.cu:224-232)
#pragma omp parallel for schedule(static)
    for (i64 i = 0; i < n; ++i) {
        if (active && !active[i]) continue;
        T ec[6];
        for (int c = 0; c < 6; ++c) ec[c] = wbar0 ? e[c][i] - wbar0[c][i] : e[c][i];
        T qxx = fma_t(sigma, ec[0], st.q[0][i]);
        T qyy = fma_t(sigma, ec[1], st.q[1][i]);
        T qxy = fma_t(sigma, ec[3], st.q[3][i]);
        T qzz = fma_t(sigma, ec[2], st.q[2][i]);
        T qxz = fma_t(sigma, ec[4], st.q[4][i]);
        T qyz = fma_t(sigma, ec[5], st.q[5][i]);
        T n2 = qxx * qxx;
        n2 = fma_t(qyy, qyy, n2);
        n2 = fma_t(T(2) * qxy, qxy, n2);
        n2 = fma_t(qzz, qzz, n2);
        n2 = fma_t(T(2) * qxz, qxz, n2);
        n2 = fma_t(T(2) * qyz, qyz, n2);
        const T nrm = std::max(T(1), sqrt_t(n2) / alpha0);
        st.q[0][i] = stored(qxx / nrm, rb); st.q[1][i] = stored(qyy / nrm, rb);
        st.q[3][i] = stored(qxy / nrm, rb); st.q[2][i] = stored(qzz / nrm, rb);
        st.q[4][i] = stored(qxz / nrm, rb); st.q[5][i] = stored(qyz / nrm, rb);
        vprev[2][i] = st.v[2][i]; vprev[0][i] = st.v[0][i]; vprev[1][i] = st.v[1][i];
    }

    // A11 tgv_launch_divergence2, tgv_cuda/tgv2_common.cu:279-320 (+ add :44-53)
    //   w_x = (D+x q_xx + D+y q_xy) + D+z q_xz
    //   w_y = (D+y q_yy + D+x q_xy) + D+z q_yz
    //   w_z = (D+z q_zz + D+x q_xz) + D+y q_yz
    T* w[3] = {e[0].data(), e[1].data(), e[2].data()};
    forward_difference(st.q[0], w[0], s, 0);
    forward_difference(st.q[3], t1, s, 1);
#pragma omp parallel for schedule(static)
    for (i64 i = 0; i < n; ++i) w[0][i] = w[0][i] + t1[i];
    forward_difference(st.q[1], w[1], s, 1);
    forward_difference(st.q[3], t1, s, 0);
#pragma omp parallel for schedule(static)
    for (i64 i = 0; i < n; ++i) w[1][i] = w[1][i] + t1[i];
    forward_difference(st.q[4], t1, s, 2);
#pragma omp parallel for schedule(static)
    for (i64 i = 0; i < n; ++i) w[0][i] = w[0][i] + t1[i];
    forward_difference(st.q[5], t1, s, 2);
#pragma omp parallel for schedule(static)
    for (i64 i = 0; i < n; ++i) w[1][i] = w[1][i] + t1[i];
    forward_difference(st.q[2], w[2], s, 2);
    forward_difference(st.q[4], t1, s, 0);
#pragma omp parallel for schedule(static)
    for (i64 i = 0; i < n; ++i) w[2][i] = w[2][i] + t1[i];
    forward_difference(st.q[5], t1, s, 1);
#pragma omp parallel for schedule(static)
    for (i64 i = 0; i < n; ++i) w[2][i] = w[2][i] + t1[i];

    // A12 tgv_kernel_part6, tgv_cuda/tgv2_common.cu:324-354
#pragma omp parallel for schedule(static)
    for (i64 i = 0; i < n; ++i) {
        if (active && !active[i]) continue;
        for (int c = 0; c < 3; ++c) {
            const T vn = fma_t(-tau, w[c][i] - st.p[c][i], st.v[c][i]);
            st.v[c][i] = stored(vn, rb);
            st.vbar[c][i] = stored(fma_t(theta, vn - vprev[c][i], vn), rb);
        }
    }
}

// ---------------------------------------------------------------------------
// TGV-k (k >= 3) deshade, reference tgvk_deshade/tgvk_l1_deshade.cu:152-351.
// Level i (0 <= i < k-2) holds the symmetric tensors w_i, w_bar_i and the dual
// r_i, components (x, y, z, xy, xz, yz).  One iteration: the TGV2 iteration with
// alpha1 = alpha[0] (p), alpha0 = alpha[1] (q) and q shifted by -w_bar_0, then
// per level i:
//   r2  = G_i(w_bar_i)                 (i odd: D+, tgv3/tgv3_common.cu:214-262;
//                                       i even: B, tgvk/tgvk_common.cu:89-135)
//   r_i = proj_{alpha[i+2]}(r_i + sigma*(r2 - w_bar_{i+1}))   (last level: no
//         w_bar_{i+1}; tgv3_common.cu:311-360, tgvk_common.cu:28-86)
//   r2  = H_i(r_i)                     (i odd: B, tgv3_common.cu:265-309;
//                                       i even: D+, tgvk_common.cu:138-182)
//   w_i = w_i - tau*(r2 - q~), q~ = q (i = 0) or r_{i-1} (updated this iteration);
//   w_bar_i = w_i + theta*(w_i - w_i_old)            (tgv3_common.cu:362-405)
// active (may be null): the masked loop, tgvk_deshade_masked/tgvk_l1_deshade_masked.cu:204-520 -
// every update only at voxels with active != 0, the stencils unmasked (as in tgv2_iteration).
// G: diagonal component a differenced along a, off-diagonal (a, b) as
// (d_a + d_b) * 0.5; H: diagonal d_a, off-diagonal d_a + d_b.
// ---------------------------------------------------------------------------
template <typename T>
struct Level { T* w[6]; T* wbar[6]; T* r[6]; };

constexpr int TENSOR_AXES[6][2] = {{0, -1}, {1, -1}, {2, -1}, {0, 1}, {0, 2}, {1, 2}};

template <typename T>
void tensor_op(T* const* in, T* const* out, const Dims& s, bool forward, bool halve, T* tmp) {
    const i64 n = s.n();
    for (int c = 0; c < 6; ++c) {
        const int a = TENSOR_AXES[c][0], b = TENSOR_AXES[c][1];
        if (forward) forward_difference(in[c], out[c], s, a);
        else backward_difference(in[c], out[c], s, a);
        if (b < 0) continue;
        if (forward) forward_difference(in[c], tmp, s, b);
        else backward_difference(in[c], tmp, s, b);
#pragma omp parallel for schedule(static)
        for (i64 i = 0; i < n; ++i)
            out[c][i] = halve ? (out[c][i] + tmp[i]) * T(0.5) : out[c][i] + tmp[i];
    }
}

template <typename T>
void tgvk_iteration(State<T>& st, std::vector<Level<T>>& lv, const T* f, const Dims& s,
                    const Params& prm, const std::vector<double>& alpha,
                    std::vector<T>* tmp /* 9 */, std::vector<T>* r2buf /* 6 */,
                    const unsigned char* active) {
    const i64 n = s.n();
    const int L = int(lv.size());
    const T tau = T(prm.tau), sigma = T(prm.sigma), theta = T(prm.theta);
    tgv2_iteration(st, f, s, prm, tmp, active, lv[0].wbar);
    T* r2[6];
    for (int c = 0; c < 6; ++c) r2[c] = r2buf[c].data();
    T* t = tmp[7].data();
    std::vector<T> wprev[6];
    for (int c = 0; c < 6; ++c) wprev[c].assign(size_t(n), T(0));
    for (int li = 0; li < L; ++li) {
        Level<T>& l = lv[li];
        const bool odd = (li % 2) == 1;
        const bool last = li == L - 1;
        const T a = T(alpha[size_t(li) + 2]);
        tensor_op(l.wbar, r2, s, /*forward=*/odd, /*halve=*/true, t);
        T* const* wnext = last ? nullptr : lv[li + 1].wbar;
#pragma omp parallel for schedule(static)
        for (i64 i = 0; i < n; ++i) {
            if (active && !active[i]) continue;
            T rc[6];
            for (int c = 0; c < 6; ++c)
                rc[c] = wnext ? fma_t(sigma, r2[c][i] - wnext[c][i], l.r[c][i])
                              : fma_t(sigma, r2[c][i], l.r[c][i]);
            T n2 = rc[0] * rc[0];
            n2 = fma_t(rc[1], rc[1], n2);
            n2 = fma_t(T(2) * rc[3], rc[3], n2);
            n2 = fma_t(rc[2], rc[2], n2);
            n2 = fma_t(T(2) * rc[4], rc[4], n2);
            n2 = fma_t(T(2) * rc[5], rc[5], n2);
            const T nrm = std::max(T(1), sqrt_t(n2) / a);
            for (int c = 0; c < 6; ++c) { l.r[c][i] = rc[c] / nrm; wprev[c][i] = l.w[c][i]; }
        }
        tensor_op(l.r, r2, s, /*forward=*/!odd, /*halve=*/false, t);
        T* const* qt = li == 0 ? st.q : lv[li - 1].r;
#pragma omp parallel for schedule(static)
        for (i64 i = 0; i < n; ++i) {
            if (active && !active[i]) continue;
            for (int c = 0; c < 6; ++c) {
                const T wn = fma_t(-tau, r2[c][i] - qt[c][i], l.w[c][i]);
                l.w[c][i] = wn;
                l.wbar[c][i] = fma_t(theta, wn - wprev[c][i], wn);
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Cubic B-spline resampling as ResizeProcessor::process(image, w, h, d, BSpline3)
// runs it (resize/ResizeProcessor.cpp:41-150) through ITK 4.8.1 (README.md:37):
// itk::ResampleImageFilter with itk::BSplineInterpolateImageFunction (order 3)
// on coefficients from itk::BSplineDecompositionImageFilter.  ITK is not part of
// the reference tree; its published algorithm is restated, in FP64:
//  - decomposition (M. Unser, "Splines: a perfect fit for signal and image
//    processing", IEEE SPM 1999, Box 2): along x, then y, then z (an axis of
//    length 1 is left alone); gain (1 - z)(1 - 1/z), pole z = sqrt(3) - 2;
//    mirror-boundary causal initialisation, truncated to the first
//    ceil(log(1e-10) / log|z|) = 18 terms when the line is longer and exact
//    otherwise; c[n] += z c[n-1]; c[N-1] = z / (z^2 - 1) (z c[N-2] + c[N-1]);
//    c[n] = z (c[n+1] - c[n]);
//  - geometry: origins 0, unit input spacing, output spacing 1 / (float(new) /
//    float(old)) per axis: output voxel j samples the continuous input index
//    j * spacing; samples outside [-0.5, N - 0.5) on any axis are 0;
//  - evaluation: support floor(x) - 1 .. floor(x) + 2 with the cubic B-spline
//    weights of w = x - floor(x), indices mirrored about 0 and N - 1 (all 0
//    when N = 1; see mirror_index), the 64 products (w_x w_y w_z) c summed x
//    fastest, the sum rounded to float;
//  - ResizeProcessor's upsampling quirk (:57-66, :146-151): when any extent
//    grows, every extent gets 2 more voxels in the resampling geometry and the
//    first (w, h, d) voxels are kept.
// ---------------------------------------------------------------------------
struct Spline3 {
    double z, gain, pole_factor;
    i64 horizon;
    Spline3() {
        z = std::sqrt(3.0) - 2.0;
        gain = (1.0 - z) * (1.0 - 1.0 / z);
        pole_factor = z / (z * z - 1.0);
        horizon = i64(std::ceil(std::log(1e-10) / std::log(std::fabs(z))));
    }
};

// One line of length N >= 2 (stride st) in place.
void spline3_decompose_line(double* c, i64 N, i64 st, const Spline3& sp) {
    const double z = sp.z;
    for (i64 n = 0; n < N; ++n) c[n * st] *= sp.gain;
    if (sp.horizon < N) {
        double zn = z, sum = c[0];
        for (i64 n = 1; n < sp.horizon; ++n) {
            sum += zn * c[n * st];
            zn *= z;
        }
        c[0] = sum;
    } else {
        double zn = z;
        const double iz = 1.0 / z;
        double z2n = std::pow(z, double(N - 1));
        double sum = c[0] + z2n * c[(N - 1) * st];
        z2n *= z2n * iz;
        for (i64 n = 1; n <= N - 2; ++n) {
            sum += (zn + z2n) * c[n * st];
            zn *= z;
            z2n *= iz;
        }
        c[0] = sum / (1.0 - zn * zn);
    }
    for (i64 n = 1; n < N; ++n) c[n * st] += z * c[(n - 1) * st];
    c[(N - 1) * st] = sp.pole_factor * (z * c[(N - 2) * st] + c[(N - 1) * st]);
    for (i64 n = N - 2; n >= 0; --n) c[n * st] = z * (c[(n + 1) * st] - c[n * st]);
}

void spline3_coefficients(const float* in, double* c, const Dims& s) {
    const Spline3 sp;
    const i64 n = s.n();
    for (i64 i = 0; i < n; ++i) c[i] = double(in[i]);
    for (int axis = 0; axis < 3; ++axis) {
        const i64 N = s.extent(axis), st = s.stride(axis);
        if (N == 1) continue;
        const i64 lines = n / N;
#pragma omp parallel for schedule(static)
        for (i64 l = 0; l < lines; ++l) {
            i64 base;
            if (axis == 0) base = l * s.w;
            else if (axis == 1) base = (l / s.w) * s.w * s.h + (l % s.w);
            else base = l;
            spline3_decompose_line(c + base, N, st, sp);
        }
    }
}

inline void spline3_weights(double x, i64& i1, double w4[4]) {
    i1 = i64(std::floor(x));
    const double w = x - double(i1);
    w4[3] = (1.0 / 6.0) * w * w * w;
    w4[0] = (1.0 / 6.0) + 0.5 * w * (w - 1.0) - w4[3];
    w4[2] = w + w4[0] - 2.0 * w4[3];
    w4[1] = 1.0 - w4[0] - w4[2] - w4[3];
}

// Whole-sample mirror about 0 and N - 1 (period 2N - 2).  ITK 4.8 reflects once
// at each end, which is the same for N >= 3; for N = 2 its upper reflection
// sends index 3 (x in [1, 1.5)) to -1, outside the coefficient buffer, so the
// folded mirror (the spline's own extension) is used for every N.
inline i64 mirror_index(i64 k, i64 N) {
    if (N == 1) return 0;
    const i64 period = 2 * (N - 1);
    k %= period;
    if (k < 0) k += period;
    return k >= N ? period - k : k;
}

// Output extents out (the voxels kept); the resampling geometry per the quirk.
void resize_spline3(const float* in, const Dims& si, float* out, const Dims& so) {
    std::vector<double> c(size_t(si.n()));
    spline3_coefficients(in, c.data(), si);
    const bool up = so.w > si.w || so.h > si.h || so.d > si.d;
    const i64 pad = up ? 2 : 0;
    const i64 geo[3] = {so.w + pad, so.h + pad, so.d + pad};
    const i64 ext_in[3] = {si.w, si.h, si.d};
    double spacing[3];
    for (int a = 0; a < 3; ++a) {
        const float sf = float(geo[a]) / float(ext_in[a]);
        spacing[a] = 1.0 / double(sf);
    }
#pragma omp parallel for schedule(static)
    for (i64 o = 0; o < so.n(); ++o) {
        const i64 oc[3] = {o % so.w, (o / so.w) % so.h, o / (so.w * so.h)};
        double x[3];
        bool inside = true;
        for (int a = 0; a < 3; ++a) {
            x[a] = double(oc[a]) * spacing[a];
            if (!(x[a] >= -0.5 && x[a] < double(ext_in[a]) - 0.5)) inside = false;
        }
        if (!inside) { out[o] = 0.0f; continue; }
        i64 idx[3][4];
        double wt[3][4];
        for (int a = 0; a < 3; ++a) {
            i64 i1;
            spline3_weights(x[a], i1, wt[a]);
            for (int k = 0; k < 4; ++k) idx[a][k] = mirror_index(i1 - 1 + k, ext_in[a]);
        }
        double acc = 0.0;
        for (int p = 0; p < 64; ++p) {
            const int k0 = p & 3, k1 = (p >> 2) & 3, k2 = p >> 4;
            double w = 1.0;
            w *= wt[0][k0];
            w *= wt[1][k1];
            w *= wt[2][k2];
            acc += w * c[size_t((idx[2][k2] * si.h + idx[1][k1]) * si.w + idx[0][k0])];
        }
        out[o] = float(acc);
    }
}

// ---------------------------------------------------------------------------
// cos(pi * t) from the mathematics: exact reduction of t modulo 2 in double,
// octant folding, then libm cos/sin of pi*r with |r| <= 1/4.
// ---------------------------------------------------------------------------
double cospi_d(double t) {
    t = std::fabs(t);
    double r = std::fmod(t, 2.0);          // exact
    if (r > 1.0) r = 2.0 - r;              // cos(pi r) symmetric about 1 (exact)
    double sgn = 1.0;
    if (r > 0.5) { r = 1.0 - r; sgn = -1.0; }   // cos(pi(1-r)) = -cos(pi r) (exact)
    if (r == 0.5) return 0.0;
    double v = (r <= 0.25) ? std::cos(M_PI * r) : std::sin(M_PI * (0.5 - r));
    return sgn * v;
}

// ---------------------------------------------------------------------------
// A15: separable DCT-II, reference cuda_image_operations/cosine_transform.cu:142-223
//   X[k] = sum_n x[n] * cospi((n + 0.5) * k / N) along x, then y, then z; *8 at the end.
//   FP32 mode: (n + 0.5) promotes to double, so each term is an FP64 product
//   added to the FP32 output element (rounded back to float each step).
// A17: separable DCT-III, reference inverse_cosine_transform.cu:79-164
//   y[n] = (sum_{k>=1} X[k] cospi((n + 0.5) k / N)) * 2 + X[0]; the z pass
//   divides by W*H*D*8.0 (double) at the end.
// ---------------------------------------------------------------------------
template <typename T>
void dct_pass(const T* in, T* out, const Dims& s, int axis, bool inverse, bool last) {
    const i64 st = s.stride(axis), n = s.extent(axis);
    const double size = double(T(n));
    // cosine table c[k*n + m] = cospi((m + 0.5) * k / size)
    std::vector<double> c(size_t(n * n));
    for (i64 k = 0; k < n; ++k)
        for (i64 m = 0; m < n; ++m)
            c[size_t(k * n + m)] = cospi_d((double(m) + 0.5) * double(k) / size);
    const double vol8 = double(s.w * s.h * s.d) * 8.0;
#pragma omp parallel for schedule(static)
    for (i64 i = 0; i < s.n(); ++i) {
        const i64 k = coord(s, i, axis);
        const i64 base = i - k * st;
        T acc = T(0);
        if (!inverse) {
            for (i64 m = 0; m < n; ++m)
                acc = T(std::fma(double(in[base + m * st]), c[size_t(k * n + m)], double(acc)));
            if (last) acc = acc * T(8);
        } else {
            // output index k plays the role of the spatial index x
            for (i64 m = 1; m < n; ++m)
                acc = T(std::fma(double(in[base + m * st]), c[size_t(m * n + k)], double(acc)));
            acc = acc * T(2);
            acc = acc + in[base];
            if (last) acc = T(double(acc) / vol8);
        }
        out[i] = acc;
    }
}

template <typename T>
void cosine_transform(const T* in, T* out, const Dims& s, bool inverse) {
    std::vector<T> t1(size_t(s.n())), t2(size_t(s.n()));
    dct_pass(in, t1.data(), s, 0, inverse, false);
    dct_pass(t1.data(), t2.data(), s, 1, inverse, false);
    dct_pass(t2.data(), out, s, 2, inverse, true);
}

// ---------------------------------------------------------------------------
// A16: Poisson divide, reference solve_poisson_in_cosine_domain.cu:27-54
//   X / (6.0 - 2.0*cospif(x/W) - 2.0*cospif(y/H) - 2.0*cospif(z/D)); (0,0,0) -> 0.
//   FP32 mode: coordinates are floats, x/W is an FP32 division, cospif is FP32,
//   the denominator and the quotient are double, the store is float.
//   index_quirk = 1 reproduces the reference's coordinate recovery through
//   float arithmetic (exact only for linear indices below 2^24).
//   Depth 1: the 2-D kernel, X / (4.0 - 2.0*cospif(x/W) - 2.0*cospif(y/H))
//   (solve_poisson_in_cosine_domain.cu:57-84, chosen at :107).
// ---------------------------------------------------------------------------
void poisson_coords(const Dims& s, i64 index, int quirk, double& xd, double& yd, double& zd) {
    if (!quirk) {
        xd = double(index % s.w); yd = double((index / s.w) % s.h); zd = double(index / (s.w * s.h));
        return;
    }
    const unsigned wh = unsigned(s.w * s.h);
    const float z = std::floor(float(unsigned(index) / wh));
    int rest = int(float(int(index)) - z * float(wh));
    const float y = std::floor(float(unsigned(rest) / unsigned(s.w)));
    rest = int(float(rest) - y * float(unsigned(s.w)));
    const float x = float(rest);
    xd = x; yd = y; zd = z;
}

template <typename T>
void solve_poisson(const T* in, T* out, const Dims& s, int quirk) {
#pragma omp parallel for schedule(static)
    for (i64 i = 0; i < s.n(); ++i) {
        double x, y, z;
        poisson_coords(s, i, quirk, x, y, z);
        if (x == 0 && y == 0 && z == 0) { out[i] = T(0); continue; }
        double den;
        if (sizeof(T) == 4) {
            const float cx = float(cospi_d(double(float(x) / float(s.w))));
            const float cy = float(cospi_d(double(float(y) / float(s.h))));
            const float cz = float(cospi_d(double(float(z) / float(s.d))));
            den = s.d == 1 ? 4.0 - 2.0 * double(cx) - 2.0 * double(cy)
                           : 6.0 - 2.0 * double(cx) - 2.0 * double(cy) - 2.0 * double(cz);
        } else {
            den = s.d == 1 ? 4.0 - 2.0 * cospi_d(x / double(s.w)) - 2.0 * cospi_d(y / double(s.h))
                           : 6.0 - 2.0 * cospi_d(x / double(s.w)) - 2.0 * cospi_d(y / double(s.h))
                                 - 2.0 * cospi_d(z / double(s.d));
        }
        out[i] = T(double(in[i]) / den);
    }
}

// A14: Poisson right-hand side, reference cuda_image_operations/divergence.cu:26-56
//   g = (B_x v_x + B_y v_y) + B_z v_z  (the B_z term only when depth > 1)
template <typename T>
void divergence(const T* vx, const T* vy, const T* vz, T* out, const Dims& s) {
    std::vector<T> a(size_t(s.n())), b(size_t(s.n())), c(size_t(s.n()));
    backward_difference(vx, a.data(), s, 0);
    backward_difference(vy, b.data(), s, 1);
    if (s.d > 1) backward_difference(vz, c.data(), s, 2);
#pragma omp parallel for schedule(static)
    for (i64 i = 0; i < s.n(); ++i) {
        T g = a[i] + b[i];
        if (s.d > 1) g = g + c[i];
        out[i] = g;
    }
}

// A18: integrate_image_gradients_poisson_cosine_transform,
// reference tgv_deshade/TGVDeshadeProcessor.cpp:386-408
template <typename T>
void integrate_poisson(const T* vx, const T* vy, const T* vz, T* out, const Dims& s, int quirk) {
    std::vector<T> g(size_t(s.n())), G(size_t(s.n()));
    divergence(vx, vy, vz, g.data(), s);
    cosine_transform(g.data(), G.data(), s, false);
    solve_poisson(G.data(), g.data(), s, quirk);
    cosine_transform(g.data(), out, s, true);
}

// ---------------------------------------------------------------------------
// A21 (new): FP64 energy and primal-dual gap of the TGV2 saddle problem with
// the operators the iteration uses, grad = D+ and eps = eps_B, and their TRUE
// transposes (the reference's B is not the exact transpose of D+).
//   primal   E = G(u) + a1*sum|D+u - v|_2 + a0*sum|eps_B v|_w
//   b1 = G(u) - <u-f, s> + G*(s) - <f, s>, s = -(D+)^T p
//        L2: (1/(2 lambda)) |lambda (u-f) - s|^2 ;  L1: lambda|u-f|_1 - <u-f, s>
//   b2 = a1 sum|D+u - v| - <D+u - v, p> ;  b3 = a0 sum|eps v|_w - <eps v, q>_w
//   gap = b1 + b2 + b3 (>= 0; L1 needs |s| <= lambda, excess in infeasibility)
//   v_residual = || eps_B^T_w q - p ||_2 (stationarity in v)
// out: [E, E_data, E_first, E_second, gap, b1, b2, b3, infeasibility, v_residual]
// ---------------------------------------------------------------------------
template <typename T>
void tgv2_energy(const T* f, const T* u, T* const* v, T* const* p, T* const* q,
                 const Dims& s, const Params& prm, double* out) {
    const i64 n = s.n();
    const double lam = prm.lambda, a0 = prm.alpha0, a1 = prm.alpha1;
    auto at = [&](const T* a, i64 i) -> double { return double(a[i]); };
    double Ed = 0, E1 = 0, E2 = 0, b1 = 0, b2 = 0, b3 = 0, inf = 0, vr = 0;
    const int na = prm.planar ? 2 : 3;   // planar: no z terms in eps_B and its adjoint
#pragma omp parallel for schedule(static) reduction(+:Ed,E1,E2,b1,b2,b3,inf,vr)
    for (i64 i = 0; i < n; ++i) {
        const i64 xc[3] = {coord(s, i, 0), coord(s, i, 1), coord(s, i, 2)};
        double du[3], bv[3][3], tq[3];
        for (int a = 0; a < 3; ++a) {
            const i64 st = s.stride(a), ext = s.extent(a);
            du[a] = (xc[a] < ext - 1) ? at(u, i + st) - at(u, i) : 0.0;
            for (int c = 0; c < 3; ++c)   // B_a v_c
                bv[a][c] = a >= na ? 0.0 : (xc[a] > 0) ? at(v[c], i - st) - at(v[c], i) : -at(v[c], i);
        }
        // s = -(D+)^T p : ((D+)^T b)[j] = b[j-1]*[j>=1] - b[j]*[j<n-1]
        double sdiv = 0;
        for (int a = 0; a < 3; ++a) {
            const i64 st = s.stride(a), ext = s.extent(a);
            double t = 0;
            if (xc[a] >= 1) t += at(p[a], i - st);
            if (xc[a] < ext - 1) t -= at(p[a], i);
            sdiv -= t;
        }
        const double umf = at(u, i) - at(f, i);
        if (prm.data_term == 0) {
            Ed += lam * std::fabs(umf);
            b1 += lam * std::fabs(umf) - umf * sdiv;
            inf += std::max(0.0, std::fabs(sdiv) - lam);
        } else {
            Ed += 0.5 * lam * umf * umf;
            const double r = lam * umf - sdiv;
            b1 += r * r / (2.0 * lam);
        }
        double g[3], gn = 0, gp = 0;
        for (int a = 0; a < 3; ++a) { g[a] = du[a] - at(v[a], i); gn += g[a] * g[a]; gp += g[a] * at(p[a], i); }
        gn = std::sqrt(gn);
        E1 += a1 * gn;
        b2 += a1 * gn - gp;
        const double e[6] = {bv[0][0], bv[1][1], bv[2][2], 0.5 * (bv[0][1] + bv[1][0]),
                             0.5 * (bv[0][2] + bv[2][0]), 0.5 * (bv[1][2] + bv[2][1])};
        double en = 0, eq = 0;
        for (int c = 0; c < 6; ++c) {
            const double wgt = c < 3 ? 1.0 : 2.0;
            en += wgt * e[c] * e[c];
            eq += wgt * e[c] * at(q[c], i);
        }
        en = std::sqrt(en);
        E2 += a0 * en;
        b3 += a0 * en - eq;
        // (eps_B^T_w q)_c = sum_a B_a^T applied to the q component pairing (a,c),
        // B^T = D+ - E_last :  (B^T b)[j] = b[j+1]*[j<n-1] - b[j]
        // <eps v, q>_w = sum_c v_c * sum_a (B_a^T qq(a,c)), qq(a,a)=q_aa, qq(a,c)=q_ac (a!=c)
        static const int qi[3][3] = {{0, 3, 4}, {3, 1, 5}, {4, 5, 2}};
        for (int c = 0; c < na; ++c) {
            double t = 0;
            for (int a = 0; a < na; ++a) {
                const i64 st = s.stride(a), ext = s.extent(a);
                const T* qq = q[qi[a][c]];
                double bt = -at(qq, i);
                if (xc[a] < ext - 1) bt += at(qq, i + st);
                t += bt;
            }
            tq[c] = t - at(p[c], i);
            vr += tq[c] * tq[c];
        }
    }
    out[0] = Ed + E1 + E2; out[1] = Ed; out[2] = E1; out[3] = E2;
    out[4] = b1 + b2 + b3; out[5] = b1; out[6] = b2; out[7] = b3;
    out[8] = inf; out[9] = std::sqrt(vr);
}

}  // namespace

// ===========================================================================
// C exports (ctypes).  Suffix _f32 = mirror order, _f64 = FP64.
// ===========================================================================
#define ORC_EXPORT extern "C" __attribute__((visibility("default")))

#define ORC_DEFINE_TYPED(T, SUF)                                                              \
ORC_EXPORT void orc_forward_difference_##SUF(const T* a, T* out, i64 w, i64 h, i64 d, int axis) { \
    forward_difference<T>(a, out, Dims{w, h, d}, axis); }                                       \
ORC_EXPORT void orc_backward_difference_##SUF(const T* a, T* out, i64 w, i64 h, i64 d, int axis) { \
    backward_difference<T>(a, out, Dims{w, h, d}, axis); }                                      \
ORC_EXPORT void orc_tgv2_init_##SUF(T** fields, const T* f, i64 w, i64 h, i64 d) {           \
    State<T> st; fill_state_pointers(st, fields); tgv2_init(st, f, Dims{w, h, d}); }           \
ORC_EXPORT int orc_tgv2_iterate_##SUF(T** fields, const T* f, i64 w, i64 h, i64 d,            \
        const double* prm6, int data_term, int iterations, const unsigned char* active,         \
        int planar) {                                                                           \
    State<T> st; fill_state_pointers(st, fields);                                               \
    Params prm{prm6[0], prm6[1], prm6[2], prm6[3], prm6[4], prm6[5], data_term, planar};        \
    const Dims s{w, h, d};                                                                      \
    std::vector<T> tmp[9];                                                                      \
    for (auto& t : tmp) t.assign(size_t(s.n()), T(0));                                          \
    for (int it = 0; it < iterations; ++it) tgv2_iteration(st, f, s, prm, tmp, active);         \
    return iterations; }                                                                        \
ORC_EXPORT void orc_divergence_##SUF(const T* vx, const T* vy, const T* vz, T* out,           \
        i64 w, i64 h, i64 d) { divergence<T>(vx, vy, vz, out, Dims{w, h, d}); }               \
ORC_EXPORT void orc_cosine_transform_##SUF(const T* in, T* out, i64 w, i64 h, i64 d) {        \
    cosine_transform<T>(in, out, Dims{w, h, d}, false); }                                       \
ORC_EXPORT void orc_inverse_cosine_transform_##SUF(const T* in, T* out, i64 w, i64 h, i64 d) { \
    cosine_transform<T>(in, out, Dims{w, h, d}, true); }                                        \
ORC_EXPORT void orc_dct_pass_##SUF(const T* in, T* out, i64 w, i64 h, i64 d, int axis,        \
        int inverse, int last) { dct_pass<T>(in, out, Dims{w, h, d}, axis, inverse, last); }   \
ORC_EXPORT void orc_solve_poisson_##SUF(const T* in, T* out, i64 w, i64 h, i64 d, int quirk) { \
    solve_poisson<T>(in, out, Dims{w, h, d}, quirk); }                                          \
ORC_EXPORT void orc_integrate_poisson_##SUF(const T* vx, const T* vy, const T* vz, T* out,    \
        i64 w, i64 h, i64 d, int quirk) { integrate_poisson<T>(vx, vy, vz, out, Dims{w, h, d}, quirk); } \
ORC_EXPORT void orc_tgv2_energy_##SUF(const T* f, T** fields, i64 w, i64 h, i64 d,            \
        const double* prm6, int data_term, int planar, double* out) {                           \
    State<T> st; fill_state_pointers(st, fields);                                               \
    Params prm{prm6[0], prm6[1], prm6[2], prm6[3], prm6[4], prm6[5], data_term, planar};        \
    tgv2_energy<T>(f, st.u, st.v, st.p, st.q, Dims{w, h, d}, prm, out); }

ORC_DEFINE_TYPED(float, f32)
ORC_DEFINE_TYPED(double, f64)

// BF16 field storage (FP32 arithmetic, every field store rounded to bfloat16).
ORC_EXPORT void orc_bf16_rne(const float* in, float* out, i64 n) {
    for (i64 i = 0; i < n; ++i) out[i] = bf16_rne(in[i]); }
ORC_EXPORT void orc_tgv2_init_bf16(float** fields, const float* f, i64 w, i64 h, i64 d) {
    State<float> st; fill_state_pointers(st, fields); tgv2_init(st, f, Dims{w, h, d}, 1); }
ORC_EXPORT int orc_tgv2_iterate_bf16(float** fields, const float* f, i64 w, i64 h, i64 d,
        const double* prm6, int data_term, int iterations, const unsigned char* active,
        int planar) {
    State<float> st; fill_state_pointers(st, fields);
    Params prm{prm6[0], prm6[1], prm6[2], prm6[3], prm6[4], prm6[5], data_term, planar, 1};
    const Dims s{w, h, d};
    std::vector<float> tmp[9];
    for (auto& t : tmp) t.assign(size_t(s.n()), 0.0f);
    for (int it = 0; it < iterations; ++it) tgv2_iteration(st, f, s, prm, tmp, active);
    return iterations; }

// TGV-k: fields = the 17 TGV2 fields, then per level w[6], w_bar[6], r[6].
// prm4 = (lambda, tau, sigma, theta); alpha[order].
#define ORC_DEFINE_TGVK(T, SUF)                                                               \
ORC_EXPORT int orc_tgvk_iterate_##SUF(T** fields, int order, const T* f, i64 w, i64 h, i64 d, \
        const double* prm4, const double* alpha, int iterations, int planar,                    \
        const unsigned char* active) {                                                          \
    if (order < 3) return -1;                                                                   \
    State<T> st; fill_state_pointers(st, fields);                                               \
    const int L = order - 2;                                                                    \
    std::vector<Level<T>> lv(static_cast<size_t>(L));                                           \
    for (int li = 0; li < L; ++li)                                                              \
        for (int c = 0; c < 6; ++c) {                                                           \
            lv[li].w[c] = fields[17 + 18 * li + c];                                             \
            lv[li].wbar[c] = fields[17 + 18 * li + 6 + c];                                      \
            lv[li].r[c] = fields[17 + 18 * li + 12 + c];                                        \
        }                                                                                       \
    Params prm{prm4[0], alpha[1], alpha[0], prm4[1], prm4[2], prm4[3], 0, planar};              \
    std::vector<double> al(alpha, alpha + order);                                               \
    const Dims s{w, h, d};                                                                      \
    std::vector<T> tmp[9], r2[6];                                                               \
    for (auto& t : tmp) t.assign(size_t(s.n()), T(0));                                          \
    for (auto& t : r2) t.assign(size_t(s.n()), T(0));                                           \
    for (int it = 0; it < iterations; ++it) tgvk_iteration(st, lv, f, s, prm, al, tmp, r2, active); \
    return iterations; }

ORC_DEFINE_TGVK(float, f32)
ORC_DEFINE_TGVK(double, f64)

// The level operators G (halve = 1) and H (halve = 0) on six FP64 arrays.
ORC_EXPORT void orc_tensor_op_f64(double** in, double** out, i64 w, i64 h, i64 d, int forward,
                                  int halve) {
    const Dims s{w, h, d};
    std::vector<double> t(size_t(s.n()));
    tensor_op<double>(in, out, s, forward != 0, halve != 0, t.data());
}

ORC_EXPORT double orc_cospi(double t) { return cospi_d(t); }
ORC_EXPORT void orc_spline3_coefficients(const float* in, double* c, i64 w, i64 h, i64 d) {
    spline3_coefficients(in, c, Dims{w, h, d});
}
ORC_EXPORT void orc_resize_spline3(const float* in, i64 w, i64 h, i64 d, float* out, i64 ow,
                                   i64 oh, i64 od) {
    resize_spline3(in, Dims{w, h, d}, out, Dims{ow, oh, od});
}
ORC_EXPORT int orc_num_threads(void) { return omp_get_max_threads(); }
ORC_EXPORT void orc_set_num_threads(int n) { omp_set_num_threads(n); }

// A19 epilogue, reference tgv_deshade/TGVDeshadeProcessor.cpp:143-165 (CLI overload:
// mask multiplies r only) and :71-95 + :275-279 (callback overload: mask multiplies
// l and r; optional add-back of u*(1-mask)).
//   r = u - l ; if mask: [l *= m if mask_shading], r *= m ; if clamp: r<0 -> 0 ;
//   if add_background_back and mask: r = r + u*(1 - m)
//   else if mask_again and mask:     r = r * m
// mask_again is the masked processor's last step without add-back
// (tgv_deshade_masked/TGVDeshadeMaskedProcessor.cpp:201-209).
ORC_EXPORT void orc_deshade_epilogue_f32(const float* u, float* l, const float* mask, float* r,
                                         i64 n, int mask_shading, int set_negative_zero,
                                         int add_background_back, int mask_again) {
#pragma omp parallel for schedule(static)
    for (i64 i = 0; i < n; ++i) {
        float ri = u[i] - l[i];
        if (mask) {
            if (mask_shading) l[i] = l[i] * mask[i];
            ri = ri * mask[i];
        }
        if (set_negative_zero && ri < 0) ri = 0.0f;
        if (add_background_back && mask) {
            const float bg = float(1.0 - double(mask[i]));
            ri = ri + u[i] * bg;
        } else if (mask_again && mask) {
            ri = ri * mask[i];
        }
        r[i] = ri;
    }
}
