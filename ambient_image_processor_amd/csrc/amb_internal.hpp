// Internal interface between the HIP kernel files and the C ABI (capi.hip).
// Written for gfx950 (MI355X): wave64, 256 CUs in 8 XCDs, 160 KB LDS per CU.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstddef>

namespace amb {

// ---------------------------------------------------------------------------
// Field storage.  Arithmetic is always FP32; StoreBF16 keeps the iteration
// fields as bfloat16 (round to nearest even on store).
// ---------------------------------------------------------------------------
// Element address base + i.  With a 32-bit index the byte offset is formed in
// 32-bit arithmetic, so the load/store can use a scalar base register plus a
// 32-bit vector offset instead of a 64-bit vector address per field (callers
// pick uint32_t only when every byte offset fits in 32 bits).
template <class T, class Idx>
__device__ __forceinline__ T* elem(T* base, Idx i) {
    if constexpr (sizeof(Idx) == 4)
        return reinterpret_cast<T*>(reinterpret_cast<char*>(base) + uint32_t(i * uint32_t(sizeof(T))));
    else
        return base + i;
}
template <class T, class Idx>
__device__ __forceinline__ const T* elem(const T* base, Idx i) {
    if constexpr (sizeof(Idx) == 4)
        return reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) +
                                          uint32_t(i * uint32_t(sizeof(T))));
    else
        return base + i;
}

struct StoreF32 {
    typedef float T;
    template <class Idx>
    static __device__ __forceinline__ float ld(const float* p, Idx i) { return *elem(p, i); }
    template <class Idx>
    static __device__ __forceinline__ void st(float* p, Idx i, float v) { *elem(p, i) = v; }
    // value after a store/load round trip
    static __device__ __forceinline__ float rt(float v) { return v; }
    template <int VX>
    static __device__ __forceinline__ void rt_all(float (&)[VX]) {}
};

struct StoreBF16 {
    typedef uint16_t T;
    template <class Idx>
    static __device__ __forceinline__ float ld(const uint16_t* p, Idx i) {
        return __uint_as_float(uint32_t(*elem(p, i)) << 16);
    }
    // Round to nearest even with gfx950's v_cvt_pk_bf16_f32.  For every
    // non-NaN input (denormals and overflow to infinity included) it gives the
    // bits of the integer form (u + 0x7fff + bit 16 of u) >> 16, checked over
    // all 2^32 patterns; a NaN stays a NaN (its payload bits may differ).
    static __device__ __forceinline__ uint16_t rne(float v) {
        return __builtin_bit_cast(uint16_t, static_cast<__bf16>(v));
    }
    // two values in one instruction: a in the low half, b in the high half
    static __device__ __forceinline__ uint32_t rne2(float a, float b) {
        typedef float f2 __attribute__((ext_vector_type(2)));
        typedef __bf16 h2 __attribute__((ext_vector_type(2)));
        const f2 v = {a, b};
        return __builtin_bit_cast(uint32_t, __builtin_convertvector(v, h2));
    }
    template <class Idx>
    static __device__ __forceinline__ void st(uint16_t* p, Idx i, float v) { *elem(p, i) = rne(v); }
    static __device__ __forceinline__ float rt(float v) { return __uint_as_float(uint32_t(rne(v)) << 16); }
    // rt of VX values, in pairs (one packed conversion per pair) when VX is even
    template <int VX>
    static __device__ __forceinline__ void rt_all(float (&a)[VX]) {
        if constexpr (VX == 1) {
            a[0] = rt(a[0]);
        } else {
            static_assert(VX % 2 == 0, "pairs");
#pragma unroll
            for (int e = 0; e < VX; e += 2) {
                const uint32_t pk = rne2(a[e], a[e + 1]);
                a[e] = __uint_as_float(pk << 16);
                a[e + 1] = __uint_as_float(pk & 0xffff0000u);
            }
        }
    }
};

// Block id -> tile id such that each of the 8 XCDs (workgroups are dealt to
// them round-robin: block b runs on XCD b % 8) receives one contiguous range
// of tiles, so tiles that share an input panel share that XCD's L2.  A
// bijection on [0, nblocks) for any nblocks: XCD x holds the
// nblocks / 8 + (x < nblocks % 8) tiles that follow those of XCDs 0 .. x - 1.
__device__ __forceinline__ unsigned xcd_contiguous(unsigned bid, unsigned nblocks) {
    const unsigned x = bid & 7u, k = bid >> 3, q = nblocks >> 3, r = nblocks & 7u;
    return x * q + (x < r ? x : r) + k;
}

// Field indices inside the 17-field state block (C ABI order).
enum Field {
    F_U = 0, F_UBAR = 1, F_VX = 2, F_VY = 3, F_VZ = 4, F_VBX = 5, F_VBY = 6, F_VBZ = 7,
    F_PX = 8, F_PY = 9, F_PZ = 10, F_QXX = 11, F_QYY = 12, F_QZZ = 13, F_QXY = 14,
    F_QXZ = 15, F_QYZ = 16, N_FIELDS = 17
};

struct Geometry {
    int w, h, d;
    __host__ __device__ size_t n() const { return size_t(w) * size_t(h) * size_t(d); }
};

// Launch-plan overrides for tests and tuning (amb_context_set_plan_override):
// held by the context, resolved into its plans when an iteration begins; the
// product defaults when unset (0, or -1 for the switches).  Results do not
// depend on them: every plan computes the same values.
struct PlanOverrides {
    int tile = -1;               // sweep tile shape id 0..7 (tile_shape)
    int zchunk = 0;              // planes per block of the two sweeps
    int sweep_pf = -1;           // next-plane prefetch of the vector sweeps (0/1)
    int fused = -1;              // 0: no fused TGV2 sweep
    int fused_vx = 0, fused_ty = 0, fused_zchunk = 0, fused_tx = 0;
    int fused_pf = -1;           // 0/1
    int fused_image = -1;        // 0: no row-marched 2-D sweep
    int fused_image_tx = -1, fused_image_zchunk = 0;
    int index64 = 0;             // 1: 64-bit-offset kernel forms on any volume
    int tgvk_zchunk = 0;
    int tgvk_fused = -1;         // 0: no fused TGV-k sweep
    int tgvk_fused_tx = 0, tgvk_fused_ty = 0, tgvk_fused_zchunk = 0;
    int adjoint_vx = 0;          // 1: one voxel per thread in the reverse phases
};

// 64-bit offsets: volumes of 2^30 voxels and more (a 4-byte field's byte
// offsets no longer fit 32 bits), or forced by the overrides.
inline bool needs_index64(const Geometry& g, const PlanOverrides& ov) {
    return g.n() * 4 >= (size_t(1) << 32) || ov.index64 > 0;
}

struct SweepParams {
    float lambda, alpha0, alpha1, tau, sigma, theta;
    int data_term;   // 0 L1, 1 L2
    int zchunk;      // planes per block along z
    int tile;        // tile shape id (tile_shape())
    // masked deshade: updates only where active[i] != 0 (nullptr: every voxel)
    const uint8_t* active;
    // depth 1 only: the reference's 2-D algebra (no z backward differences in
    // eps_B(v_bar); the z components of v, p, q stay 0)
    int planar;
    // planar, and every copy of v_z, p_z, q_zz, q_xz, q_yz holds +0: the fused
    // sweep may take them as +0 without loading or storing them
    int planar_zero_z;
    // fused sweep only: non-null = the launch also reduces the FP64 energy and
    // gap of the finished iteration into one row of 10 doubles per block here
    // (the plan from fused_energy_plan)
    double* energy_partials;
    // resolved when the iteration begins: 64-bit-offset kernel forms, and the
    // vector sweeps' next-plane prefetch (-1: the storage's default)
    int index64;
    int sweep_pf;
    // masked fused sweep only: one byte per tile of its plan, 0 = no voxel the
    // tile owns is active, so the tile skips (valid once both copies of the
    // state agree outside the active set); nullptr = no skipping
    const uint8_t* tile_active;
};

// Column-tile shapes of the sweep kernels: TX x TY threads, each thread owning
// VX consecutive x voxels (ids 0-3: one voxel per thread; ids 4-7: four voxels
// per thread with vector loads, used when the width is a multiple of 4).
inline void tile_shape(int id, int& tx, int& ty, int& vx) {
    vx = id >= 4 ? 4 : 1;
    switch (id) {
        case 1: tx = 64; ty = 8; break;
        case 2: tx = 32; ty = 8; break;
        case 3: tx = 64; ty = 2; break;
        case 4: tx = 64; ty = 4; break;
        case 5: tx = 64; ty = 2; break;
        case 6: tx = 64; ty = 8; break;
        case 7: tx = 32; ty = 8; break;
        default: tx = 64; ty = 4; break;
    }
}
constexpr int TILE_SCALAR_DEFAULT = 0;
constexpr int TILE_VECTOR_DEFAULT = 4;

// Launchers (stream-ordered, asynchronous).  `fields` points at 17 arrays of
// the storage type (void* so both storages share the signature).
hipError_t launch_init(hipStream_t s, int storage, void* const* fields, const float* f, Geometry g);
// zero `bytes` (a multiple of 4, 4-byte aligned) at p with a fill kernel
hipError_t launch_zero_fill(hipStream_t s, void* p, size_t bytes);
// Raises a kernel's dynamic LDS limit on the current device to at least
// `bytes`, once per (device, kernel, size): setting it before every launch
// costs a gap between back-to-back launches.  Thread-safe.
hipError_t ensure_dynamic_lds(const void* kernel, size_t bytes);
// Fused sweep (primal half of iteration k + dual half of iteration k + 1):
// reads u, v, p, q through the 17-field table `src` and writes them to the 13
// fields `dst13` (order u, vx, vy, vz, px, py, pz, qxx, qyy, qzz, qxy, qxz,
// qyz); u_bar and v_bar are neither read nor written.  Needs an even W.
struct FusedPlan {
    int ty;       // owned y rows per block
    int zchunk;   // planes per block
    int pf;       // 1: next-plane loads issued one step ahead
    int vx;       // voxels per thread along x (2 or 4)
    int tx;       // 0: whole rows; else owned x columns per block (x tiles, vx 2)
};
bool fused_plan(const Geometry& g, int storage, int device, const PlanOverrides& ov,
                FusedPlan* plan);
size_t fused_lds_bytes(const Geometry& g, const FusedPlan& plan);
// Plan of the fused sweep's energy form (SweepParams::energy_partials) derived
// from the iteration plan; false when the form does not cover the geometry.
bool fused_energy_plan(const Geometry& g, int device, const PlanOverrides& ov,
                       const FusedPlan& base, FusedPlan* plan);
// blocks of a fused launch with this plan (= rows of energy partials)
size_t fused_grid_blocks(const Geometry& g, const FusedPlan& plan);
// out[t] = 1 if tile t of the fused plan owns an active voxel (fused_grid_blocks bytes)
hipError_t launch_tile_activity(hipStream_t s, const uint8_t* act, Geometry g,
                                const FusedPlan& plan, uint8_t* out);
hipError_t launch_fused_sweep(hipStream_t s, int storage, void* const* src, void* const* dst13,
                              const float* f, Geometry g, const SweepParams& p,
                              const FusedPlan& plan);
constexpr int N_FUSED_FIELDS = 13;
extern const int FUSED_FIELD_MAP[N_FUSED_FIELDS];
// The fused sweep out of place with u_bar and v_bar stored as well: src and dst
// are full 17-field tables (state layout); dst receives u, u_bar, v, v_bar of the
// primal half and p, q of the next dual half.  FP32 fields, 32-bit offsets, the
// 3-D loop without a mask, the L2 data term and a two- or one-voxel plan
// (fused_bars_supported); for the adjoint's segment recompute.
bool fused_bars_supported(int storage, Geometry g, const SweepParams& p, const FusedPlan& plan);
// The fused sweep divides |p| / alpha1 and |q| / alpha0 through host
// reciprocals, exact for alphas within [2^-60, 2^36]; outside that range the
// iteration runs the two sweeps.
bool fused_divisors_ok(float alpha0, float alpha1);
hipError_t launch_fused_sweep_bars(hipStream_t s, void* const* src, void* const* dst,
                                   const float* f, Geometry g, const SweepParams& p,
                                   const FusedPlan& plan);
// The 2-D loop (planar, FP32 fields, z components +0 in both copies) as a
// fused sweep marched along the image's rows; same src / dst13 convention as
// launch_fused_sweep.
bool fused_image_plan(const Geometry& g, int device, const PlanOverrides& ov, FusedPlan* plan);
hipError_t launch_fused_image(hipStream_t s, void* const* src, void* const* dst13, const float* f,
                              Geometry g, const SweepParams& p, const FusedPlan& plan);

hipError_t launch_dual_sweep(hipStream_t s, int storage, void* const* fields, Geometry g,
                             const SweepParams& p);
hipError_t launch_primal_sweep(hipStream_t s, int storage, void* const* fields, const float* f,
                               Geometry g, const SweepParams& p);
hipError_t launch_convert(hipStream_t s, int storage, const void* src, void* dst, size_t n,
                          int to_storage);

// DCT / Poisson / divergence / epilogue (dct.hip)
struct CosTables;   // device-resident cosine tables for one geometry
hipError_t launch_divergence(hipStream_t s, const float* vx, const float* vy, const float* vz,
                             float* out, Geometry g);
hipError_t launch_dct_reference_pass(hipStream_t s, const float* in, float* out,
                                     const double* table, Geometry g, int axis, int inverse,
                                     int last);
hipError_t launch_dct_mfma_pass(hipStream_t s, const float* in, float* out, const float* table,
                                const uint16_t* table_hi, const uint16_t* table_lo, Geometry g,
                                int axis, int inverse, int last, int precision,
                                const float* eig_x, const float* eig_y, const float* eig_z,
                                int fuse_poisson, int quirk);
int dct_mfma_max_extent();
hipError_t launch_poisson_divide(hipStream_t s, const float* in, float* out, const float* cx,
                                 const float* cy, const float* cz, Geometry g, int quirk);
hipError_t launch_epilogue(hipStream_t s, const float* u, float* l, const float* mask, float* r,
                           size_t n, int mask_shading, int set_negative_zero,
                           int add_background_back, int mask_again);
hipError_t launch_active_from_mask(hipStream_t s, const float* mask, uint8_t* act, size_t n);

// TGV-k (k >= 3) sweeps (tgvk.hip).  fields_dev: device table of the 17 state
// fields; levels_dev: device table of 18 pointers per level (w[6], w_bar[6],
// r[6]); alpha_dev: the k weights; active (may be null): one byte per voxel,
// updates only where nonzero.  FP32 storage.
// ov: the z chunk and 64-bit-offset overrides (resolved when the iteration begins).
hipError_t launch_tgvk_dual(hipStream_t s, float* const* fields_dev, float* const* levels_dev,
                            const float* alpha_dev, const uint8_t* active, Geometry g, int levels,
                            float sigma, int planar, const PlanOverrides& ov);
hipError_t launch_tgvk_primal(hipStream_t s, float* const* fields_dev, float* const* levels_dev,
                              const float* f, const uint8_t* active, Geometry g, int levels,
                              float tau, float lambda, float theta, const PlanOverrides& ov);
// Fused TGV-k sweep for one level (k = 3): primal half of iteration k and
// dual half of iteration k + 1, reading the 25 fields `src` (order u, vx, vy,
// vz, px, py, pz, qxx, qyy, qzz, qxy, qxz, qyz, wx, wy, wz, wxy, wxz, wyz, rx,
// ry, rz, rxy, rxz, ryz) and writing the same fields to `dst`; u_bar, v_bar
// and w_bar are neither read nor written.  dst[k] - src[k] must be the same
// for the first 13 fields and the same for the last 12.  alpha: the three
// weights (host).
struct TgvkFusedPlan {
    int tx;       // owned x columns per block (W: whole rows)
    int ty;       // owned y rows per block
    int zchunk;   // planes per block
};
constexpr int N_TGVK_FUSED_FIELDS = 25;
bool tgvk_fused_plan(const Geometry& g, int levels, int device, const PlanOverrides& ov,
                     TgvkFusedPlan* plan);
hipError_t launch_tgvk_fused(hipStream_t s, const float* const* src, float* const* dst,
                             const float* f, const uint8_t* active, Geometry g, float tau,
                             float sigma, float lambda, float theta, const float* alpha,
                             int planar, const TgvkFusedPlan& plan);

// Cubic B-spline resampling (resize.hip): in (gi) to out (go, the kept voxels);
// coef: gi.n() doubles of scratch.
hipError_t launch_resize_spline3(hipStream_t s, const float* in, Geometry gi, double* coef,
                                 float* out, Geometry go);
// Elementwise tail of the downsampled deshade (resize.hip); mask may be null.
hipError_t launch_downsampled_tail(hipStream_t s, const float* f, const float* mask, float* l,
                                   float* r, size_t n, int masked_variant, int add_background_back,
                                   int set_negative_zero);
hipError_t launch_multiply(hipStream_t s, const float* a, const float* b, float* out, size_t n);

// FP64 energy / gap (energy.hip).  fields: HOST array of the 17 device field
// pointers (passed to the kernel by value); partials: at least
// energy_partials_size(g) doubles.
size_t energy_partials_size(Geometry g);
hipError_t launch_energy(hipStream_t s, int storage, const void* const* fields, const float* f,
                         Geometry g, double lambda, double alpha0, double alpha1, int data_term,
                         int planar, double* partials, double* out10);
// Final fixed-order FP64 sum of `rows` rows of 10 partial sums into out10
// (slot 9 is the square root of its sum).
hipError_t launch_energy_finish(hipStream_t s, const double* partials, int rows, double* out10);

// Adjoint of the unrolled L2 iteration (adjoint.hip).  adj, xk, xk1: 17-field
// tables (adjoint of x_{k+1} updated in place to the adjoint of x_k); slots:
// adjoint_slots_size(g) doubles, zeroed before the first step, in which every
// block accumulates its (alpha0, alpha1, lambda) terms over the steps;
// launch_adjoint_reduce sums them (fixed order) into out3.
size_t adjoint_slots_size(Geometry g);
// The reverse pass's buffers and constants: adj (the 17 adjoint fields), pbuf
// (3 fields: phase A's p adjoint), ast (1 field), wst (3 fields), grad_f, slots.
struct AdjointArgs {
    void* const* adj;
    float* pbuf;
    float* ast;
    float* wst;
    float* grad_f;
    const float* f;
    double* slots;
    Geometry g;
    float tau, sigma, theta, lambda, alpha0, alpha1;
    int vx;   // voxels per thread of phases B and C (adjoint_vx)
};
// Voxels per thread of the reverse phases: 2 (paired 8-byte loads) when the
// width is even, else 1; want = 1 (plan override adjoint_vx) forces 1.
int adjoint_vx(Geometry g, int want);
// Iteration k in reverse: launch_adjoint_a (phase A of the last iteration only,
// u_next = u of x_N), then per k: launch_adjoint_b (xk = x_k), then
// launch_adjoint_ca (phase C of k with phase A of k - 1, u_next = u of x_k) or,
// for k = 0, launch_adjoint_c.
hipError_t launch_adjoint_a(hipStream_t s, const AdjointArgs& a, const float* u_next);
hipError_t launch_adjoint_b(hipStream_t s, const AdjointArgs& a, void* const* xk);
hipError_t launch_adjoint_ca(hipStream_t s, const AdjointArgs& a, const float* u_next);
hipError_t launch_adjoint_c(hipStream_t s, const AdjointArgs& a);
hipError_t launch_adjoint_reduce(hipStream_t s, const double* slots, Geometry g, double* out3);
hipError_t launch_adjoint_finish(hipStream_t s, void* const* adj, float* grad_f, Geometry g);

// Host-side cos(pi t) from the mathematics (exact reduction modulo 2, octant
// folding, libm cos/sin on |pi r| <= pi/4).
double host_cospi(double t);

}  // namespace amb
