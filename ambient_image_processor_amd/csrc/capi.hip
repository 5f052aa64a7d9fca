// C ABI of the MI355X TGV-DCT deshade path (declared in include/ambient_amd.h).
//
// Host orchestration of the reference launchers' work, device-resident:
//   tgv2_l1_deshade_launch  (reference tgv_deshade/tgv2_l1_deshade.cu:33-303)
//   tgv2_l1/l2_launch       (tgv_cuda/tgv2_l1.cu:64-321, tgv2_l2.cu:60-316)
//   DCT Poisson solve       (tgv_deshade/TGVDeshadeProcessor.cpp:386-408)
//   deshade composition     (TGVDeshadeProcessor.cpp:71-95, 122-166, 224-282)
// The reference allocates 32 device arrays per call and copies every
// intermediate back to the host; here a context keeps 17 state fields + f +
// 5 scratch volumes resident and all steps run on the context's stream.
#include "amb_internal.hpp"
#include "../../include/ambient_amd.h"

#include <cmath>
#include <cstdio>
#include <cstring>
#include <exception>
#include <new>
#include <algorithm>
#include <cstdlib>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

using namespace amb;

namespace {

thread_local std::string g_last_error;

amb_status fail(amb_status st, const std::string& msg) {
    g_last_error = msg;
    return st;
}

amb_status fail_hip(hipError_t e, const char* what, int line) {
    char buf[512];
    std::snprintf(buf, sizeof(buf), "HIP error %d (%s) in %s [capi.hip:%d]", int(e),
                  hipGetErrorString(e), what, line);
    g_last_error = buf;
    return e == hipErrorOutOfMemory ? AMB_ERR_OUT_OF_MEMORY : AMB_ERR_HIP;
}

#define AMB_HIP(call)                                                     \
    do {                                                                  \
        hipError_t e_ = (call);                                           \
        if (e_ != hipSuccess) return fail_hip(e_, #call, __LINE__);       \
    } while (0)

#define AMB_TRY(call)                                                     \
    do {                                                                  \
        amb_status s_ = (call);                                           \
        if (s_ != AMB_OK) return s_;                                      \
    } while (0)

constexpr int N_SCRATCH = 5;
constexpr int N_PROF = 15;
enum { P_DUAL = 0, P_PRIMAL = 1, P_DCT = 2, P_POISSON = 3, P_DIV = 4, P_EPI = 5, P_INIT = 6,
       P_ENERGY = 7, P_ADJOINT = 8, P_FUSED = 9, P_TGVK_DUAL = 10, P_TGVK_PRIMAL = 11,
       P_RESIZE = 12, P_TGVK_FUSED = 13, P_FUSED_ENERGY = 14 };

struct AxisTables {
    int n = 0;
    double* fwd64 = nullptr;     // [m][k] cospi((m+0.5)k/n)
    double* inv64 = nullptr;     // [m][k] cospi((k+0.5)m/n)
    float* fwd32 = nullptr;      // FP32 copies for the MFMA passes, [k][m] layout
    float* inv32 = nullptr;
    uint16_t* fwd_hi = nullptr;  // bf16 splits, [k][m]
    uint16_t* fwd_lo = nullptr;
    uint16_t* inv_hi = nullptr;
    uint16_t* inv_lo = nullptr;
    float* eig = nullptr;        // float(cospi(float(k)/float(n))), n + 1 entries
};

}  // namespace

struct amb_context {
    int device = 0;
    hipStream_t stream = nullptr;
    // state
    Geometry g{0, 0, 0};
    int storage = 0;
    size_t cap_bytes = 0;      // bytes of the fields block
    size_t cap_n = 0;          // voxels the float buffers can hold
    char* block = nullptr;     // fields
    void* fields[N_FIELDS] = {};
    void** fields_dev = nullptr;   // device copy of the field pointer table
    float* f = nullptr;
    float* scratch[N_SCRATCH] = {};
    char* fblock = nullptr;    // f + scratch
    double* partials = nullptr;
    size_t partials_cap = 0;
    double* energy_out = nullptr;
    bool have_state = false;
    bool z_zero = false;   // v_z, p_z, q_zz, q_xz, q_yz are +0 (from begin(), not set_state)
    SweepParams sp{};
    // launch-plan overrides (tests, tuning); read when an iteration begins.
    // kov: the ones the TGV-k loop of the current state began with
    PlanOverrides ov{};
    PlanOverrides kov{};
    amb_tgv_params params{};
    // cosine tables per axis, keyed by extent
    AxisTables axes[3];
    // profiling
    bool prof_on = false;
    uint32_t prof_every = 1;          // time one launch in prof_every per class
    uint64_t prof_seen[N_PROF] = {};  // launches while profiling, timed or not
    std::vector<hipEvent_t> pool;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> pending[N_PROF];
    double prof_ms[N_PROF] = {};
    uint64_t prof_n[N_PROF] = {};
    // tiles of the 3-D fused sweep (class 9) launched while profiling: swept and in
    // the grid (a masked run skips the tiles that own no active voxel)
    uint64_t prof_tiles_swept = 0, prof_tiles_total = 0;
    // host staging for callbacks
    std::vector<float> cb_u, cb_v[3];
    // adjoint workspace (grow-only): checkpoints, recompute slots, adjoint state,
    // A*, W*, dL/df, partials, scalar accumulators
    char* adj_block = nullptr;
    size_t adj_cap = 0;
    // masked iteration: one byte per voxel, nonzero = update (grow-only), and
    // one byte per tile of the fused plan, nonzero = the tile owns an active voxel
    uint8_t* act = nullptr;
    size_t act_cap = 0;
    uint8_t* tact = nullptr;
    size_t tact_cap = 0;
    size_t tact_active = 0;   // active tiles of the masked run's plan
    // fused sweep: second copy of u, v, p, q (grow-only) and the launch plan
    char* alt_block = nullptr;
    size_t alt_cap = 0;
    void* alt[N_FUSED_FIELDS] = {};
    bool fused_ok = false;
    FusedPlan fplan{0, 0};
    FusedPlan fplan_run{0, 0};   // the plan of the call in progress (masked: its own z chunk)
    bool iplan_ok = false;   // the 2-D loop marched along the image's rows
    FusedPlan iplan{0, 0};
    // energy / gap trace: every energy_interval iterations of a call (0 = off)
    // and after its last one; in the fused sweep's energy form where that
    // iteration's primal half runs fused, else the standalone pass
    uint32_t energy_interval = 0;
    bool fen_ok = false;
    FusedPlan fen_plan{0, 0};
    double* etrace_dev = nullptr;     // rows of 10 doubles
    size_t etrace_cap = 0;            // rows
    std::vector<double> etrace;       // rows of 11: iteration count, 10 values
    std::vector<uint32_t> etrace_it;
    // TGV-k: 18 fields per level (w[6], w_bar[6], r[6]; grow-only), their
    // device pointer table and the k weights
    char* kblock = nullptr;
    size_t kcap = 0;
    float** klv_dev = nullptr;
    size_t klv_cap = 0;
    float* kalpha_dev = nullptr;
    size_t kalpha_cap = 0;
    std::vector<float*> klv_host;
    std::vector<float> kalpha_host;
    int klevels = 0;
    // fused TGV-k sweep (one level): second copy of w_0 and r_0 (grow-only),
    // the launch plan
    char* kalt_block = nullptr;
    size_t kalt_cap = 0;
    float* kalt[12] = {};
    char* kalt13_block = nullptr;   // its own second copy of the 13 TGV2 fields
    size_t kalt13_cap = 0;
    void* kalt13[N_FUSED_FIELDS] = {};
    bool kfused_ok = false;
    TgvkFusedPlan kfplan{0, 0, 0};
    // resampling and the downsampled deshade: one grow-only workspace
    char* rblock = nullptr;
    size_t rcap = 0;
};

namespace {

size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

amb_status set_device(amb_context* c) {
    AMB_HIP(hipSetDevice(c->device));
    return AMB_OK;
}

hipEvent_t take_event(amb_context* c) {
    if (!c->pool.empty()) {
        hipEvent_t e = c->pool.back();
        c->pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

struct ProfScope {
    amb_context* c;
    int cls;
    hipEvent_t a = nullptr;
    ProfScope(amb_context* c_, int cls_) : c(c_), cls(cls_) {
        if (c->prof_on && (c->prof_seen[cls]++ % c->prof_every) == 0) {
            a = take_event(c);
            if (a) (void)hipEventRecord(a, c->stream);
        }
    }
    ~ProfScope() {
        if (a) {
            hipEvent_t b = take_event(c);
            if (b) {
                (void)hipEventRecord(b, c->stream);
                c->pending[cls].push_back({a, b});
            } else {
                c->pool.push_back(a);
            }
        }
    }
};

amb_status drain_profile(amb_context* c) {
    AMB_HIP(hipStreamSynchronize(c->stream));
    for (int k = 0; k < N_PROF; ++k) {
        for (auto& pr : c->pending[k]) {
            float ms = 0.f;
            AMB_HIP(hipEventElapsedTime(&ms, pr.first, pr.second));
            c->prof_ms[k] += ms;
            c->prof_n[k] += 1;
            c->pool.push_back(pr.first);
            c->pool.push_back(pr.second);
        }
        c->pending[k].clear();
    }
    return AMB_OK;
}

amb_status check_geometry(uint32_t w, uint32_t h, uint32_t d) {
    if (w == 0 || h == 0 || d == 0) return fail(AMB_ERR_INVALID_ARGUMENT, "zero-sized volume");
    const double n = double(w) * double(h) * double(d);
    if (n > 2147483647.0) return fail(AMB_ERR_INVALID_ARGUMENT, "volume exceeds 2^31-1 voxels");
    if (w > 65536 || h > 65536 || d > 65536)
        return fail(AMB_ERR_INVALID_ARGUMENT, "extent exceeds 65536");
    return AMB_OK;
}

amb_status check_params(const amb_tgv_params* p) {
    if (!p) return fail(AMB_ERR_INVALID_ARGUMENT, "params is NULL");
    if (p->data_term > 1) return fail(AMB_ERR_INVALID_ARGUMENT, "data_term must be L1 or L2");
    if (p->storage > 1) return fail(AMB_ERR_INVALID_ARGUMENT, "storage must be F32 or BF16");
    if (p->dct_precision > 2) return fail(AMB_ERR_INVALID_ARGUMENT, "unknown dct_precision");
    if (!(p->alpha0 > 0.f) || !(p->alpha1 > 0.f))
        return fail(AMB_ERR_INVALID_ARGUMENT, "alpha0 and alpha1 must be positive");
    if (!(p->tau > 0.f) || !(p->sigma > 0.f))
        return fail(AMB_ERR_INVALID_ARGUMENT, "tau and sigma must be positive");
    if (!(p->lambda >= 0.f)) return fail(AMB_ERR_INVALID_ARGUMENT, "lambda must be >= 0");
    return AMB_OK;
}

// Allocate (or reuse) the state block for geometry g and the storage type.
amb_status ensure_buffers(amb_context* c, Geometry g, int storage) {
    const size_t n = g.n();
    const size_t elem = storage == 1 ? 2 : 4;
    const size_t fstride = align_up(n * elem, 256);
    const size_t need = fstride * N_FIELDS;
    if (need > c->cap_bytes) {
        if (c->block) AMB_HIP(hipFree(c->block));
        c->block = nullptr;
        c->cap_bytes = 0;
        AMB_HIP(hipMalloc(reinterpret_cast<void**>(&c->block), need));
        c->cap_bytes = need;
    }
    for (int k = 0; k < N_FIELDS; ++k) c->fields[k] = c->block + size_t(k) * fstride;
    const size_t fl = align_up(n * 4, 256);
    if (n > c->cap_n) {
        if (c->fblock) AMB_HIP(hipFree(c->fblock));
        c->fblock = nullptr;
        c->cap_n = 0;
        AMB_HIP(hipMalloc(reinterpret_cast<void**>(&c->fblock), fl * (1 + N_SCRATCH)));
        c->cap_n = n;
    }
    c->f = reinterpret_cast<float*>(c->fblock);
    for (int k = 0; k < N_SCRATCH; ++k)
        c->scratch[k] = reinterpret_cast<float*>(c->fblock + fl * size_t(1 + k));
    if (!c->fields_dev) AMB_HIP(hipMalloc(reinterpret_cast<void**>(&c->fields_dev), sizeof(void*) * N_FIELDS));
    AMB_HIP(hipMemcpyAsync(c->fields_dev, c->fields, sizeof(void*) * N_FIELDS,
                           hipMemcpyHostToDevice, c->stream));
    size_t pn = energy_partials_size(g);
    FusedPlan base{0, 0}, en{0, 0};
    if (fused_plan(g, storage, c->device, c->ov, &base) &&
        fused_energy_plan(g, c->device, c->ov, base, &en))
        pn = std::max(pn, fused_grid_blocks(g, en) * 10 + 16);
    if (pn > c->partials_cap) {
        if (c->partials) AMB_HIP(hipFree(c->partials));
        c->partials = nullptr;
        AMB_HIP(hipMalloc(reinterpret_cast<void**>(&c->partials), pn * sizeof(double)));
        c->partials_cap = pn;
    }
    if (!c->energy_out) AMB_HIP(hipMalloc(reinterpret_cast<void**>(&c->energy_out), 16 * sizeof(double)));
    c->g = g;
    c->storage = storage;
    return AMB_OK;
}

void free_axis(AxisTables& t) {
    void* ptrs[] = {t.fwd64, t.inv64, t.fwd32, t.inv32, t.fwd_hi, t.fwd_lo, t.inv_hi, t.inv_lo, t.eig};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    t = AxisTables();
}

uint16_t host_bf16_rne(float v) {
    uint32_t u;
    std::memcpy(&u, &v, 4);
    u += 0x7fffu + ((u >> 16) & 1u);
    return uint16_t(u >> 16);
}
float host_bf16_to_float(uint16_t h) {
    uint32_t u = uint32_t(h) << 16;
    float v;
    std::memcpy(&v, &u, 4);
    return v;
}

// Largest extent of the DCT passes: the reference-numerics passes read an
// n x n FP64 table per direction (16 n^2 bytes of device memory per axis,
// 4 GiB at n = 16384); the reference evaluates its cosines on the fly and has
// no such limit.
constexpr int DCT_MAX_EXTENT = 16384;

// Cosine tables for one axis of extent n (built once per extent, host FP64,
// uploaded in row chunks so that host memory stays O(n)): the FP64 [m][k]
// tables of the reference-numerics passes always, the FP32 and bf16-split
// [k][m] tables of the MFMA passes only for n <= dct_mfma_max_extent().
amb_status ensure_axis(amb_context* c, int axis, int n) {
    AxisTables& t = c->axes[axis];
    if (t.n == n) return AMB_OK;
    if (n > DCT_MAX_EXTENT)
        return fail(AMB_ERR_NOT_SUPPORTED,
                    "DCT Poisson solve: extent " + std::to_string(n) +
                        " exceeds 16384 (the cosine tables would exceed 4 GiB per axis)");
    free_axis(t);
    const size_t nn = size_t(n) * n;
    const bool mfma = n <= dct_mfma_max_extent();
    const double size = double(float(n));
    AMB_HIP(hipMalloc(reinterpret_cast<void**>(&t.fwd64), nn * 8));
    AMB_HIP(hipMalloc(reinterpret_cast<void**>(&t.inv64), nn * 8));
    if (mfma) {
        AMB_HIP(hipMalloc(reinterpret_cast<void**>(&t.fwd32), nn * 4));
        AMB_HIP(hipMalloc(reinterpret_cast<void**>(&t.inv32), nn * 4));
        AMB_HIP(hipMalloc(reinterpret_cast<void**>(&t.fwd_hi), nn * 2));
        AMB_HIP(hipMalloc(reinterpret_cast<void**>(&t.fwd_lo), nn * 2));
        AMB_HIP(hipMalloc(reinterpret_cast<void**>(&t.inv_hi), nn * 2));
        AMB_HIP(hipMalloc(reinterpret_cast<void**>(&t.inv_lo), nn * 2));
    }
    AMB_HIP(hipMalloc(reinterpret_cast<void**>(&t.eig), (size_t(n) + 1) * 4));
    const size_t rows = std::max<size_t>(1, (size_t(1) << 21) / size_t(n));   // ~2M entries
    std::vector<double> a64, b64;
    std::vector<float> a32, b32;
    std::vector<uint16_t> ah, al, bh, bl;
    for (size_t r0 = 0; r0 < size_t(n); r0 += rows) {
        const size_t nr = std::min(rows, size_t(n) - r0), cnt = nr * size_t(n);
        // FP64 reference tables, row r = m: [m][k] cospi((m+0.5)k/n), cospi((k+0.5)m/n)
        a64.resize(cnt);
        b64.resize(cnt);
        for (size_t r = 0; r < nr; ++r) {
            const int m = int(r0 + r);
            for (int k = 0; k < n; ++k) {
                a64[r * n + k] = host_cospi((double(m) + 0.5) * double(k) / size);
                b64[r * n + k] = host_cospi((double(k) + 0.5) * double(m) / size);
            }
        }
        AMB_HIP(hipMemcpy(t.fwd64 + r0 * n, a64.data(), cnt * 8, hipMemcpyHostToDevice));
        AMB_HIP(hipMemcpy(t.inv64 + r0 * n, b64.data(), cnt * 8, hipMemcpyHostToDevice));
        if (!mfma) continue;
        // MFMA tables, row r = k: [k][m] (contiguous over the summed index); the
        // inverse folds the reference's "X[0] + 2 sum_{m>=1}" into weights 1, 2
        a32.resize(cnt); b32.resize(cnt);
        ah.resize(cnt); al.resize(cnt); bh.resize(cnt); bl.resize(cnt);
        for (size_t r = 0; r < nr; ++r) {
            const int k = int(r0 + r);
            for (int m = 0; m < n; ++m) {
                const double cf = host_cospi((double(m) + 0.5) * double(k) / size);
                const double ci = (m == 0 ? 1.0 : 2.0) * host_cospi((double(k) + 0.5) * double(m) / size);
                const size_t o = r * n + size_t(m);
                a32[o] = float(cf);
                b32[o] = float(ci);
                const uint16_t h1 = host_bf16_rne(float(cf));
                ah[o] = h1;
                al[o] = host_bf16_rne(float(cf - double(host_bf16_to_float(h1))));
                const uint16_t h2 = host_bf16_rne(float(ci));
                bh[o] = h2;
                bl[o] = host_bf16_rne(float(ci - double(host_bf16_to_float(h2))));
            }
        }
        AMB_HIP(hipMemcpy(t.fwd32 + r0 * n, a32.data(), cnt * 4, hipMemcpyHostToDevice));
        AMB_HIP(hipMemcpy(t.inv32 + r0 * n, b32.data(), cnt * 4, hipMemcpyHostToDevice));
        AMB_HIP(hipMemcpy(t.fwd_hi + r0 * n, ah.data(), cnt * 2, hipMemcpyHostToDevice));
        AMB_HIP(hipMemcpy(t.fwd_lo + r0 * n, al.data(), cnt * 2, hipMemcpyHostToDevice));
        AMB_HIP(hipMemcpy(t.inv_hi + r0 * n, bh.data(), cnt * 2, hipMemcpyHostToDevice));
        AMB_HIP(hipMemcpy(t.inv_lo + r0 * n, bl.data(), cnt * 2, hipMemcpyHostToDevice));
    }
    std::vector<float> eig(size_t(n) + 1);
    for (int k = 0; k <= n; ++k) eig[k] = float(host_cospi(double(float(k) / float(n))));
    AMB_HIP(hipMemcpy(t.eig, eig.data(), (size_t(n) + 1) * 4, hipMemcpyHostToDevice));
    t.n = n;
    return AMB_OK;
}

amb_status ensure_tables(amb_context* c, Geometry g) {
    AMB_TRY(ensure_axis(c, 0, g.w));
    AMB_TRY(ensure_axis(c, 1, g.h));
    AMB_TRY(ensure_axis(c, 2, g.d));
    return AMB_OK;
}

hipMemcpyKind kind_in(amb_memory mem) {
    return mem == AMB_MEM_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
}
hipMemcpyKind kind_out(amb_memory mem) {
    return mem == AMB_MEM_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
}

// Sweep tile shape: 64x4 threads unless the tile override selects another
// shape id (tile_shape() in amb_internal.hpp), for tuning.
int choose_tile(const Geometry& g, const PlanOverrides& ov) {
    int t = g.w % 4 == 0 ? TILE_VECTOR_DEFAULT : TILE_SCALAR_DEFAULT;
    if (ov.tile >= 0 && ov.tile <= 7) t = ov.tile;
    if (t >= 4 && g.w % 4 != 0) t = TILE_SCALAR_DEFAULT;   // vector tiles need W % 4 == 0
    return t;
}

// Planes per block: 4 for FP32 vector tiles and 16 for BF16 vector tiles
// (measured on MI355X at 256^3), else >= 4096 blocks with chunks of 8..64.
int choose_zchunk(const Geometry& g, int tile, int storage, const PlanOverrides& ov) {
    if (ov.zchunk > 0) return ov.zchunk;
    if (tile >= 4) return std::min(storage == 1 ? 16 : 4, g.d);
    int tx, ty, vx;
    tile_shape(tile, tx, ty, vx);
    const long nxy = long((g.w + tx * vx - 1) / (tx * vx)) * long((g.h + ty - 1) / ty);
    // aim for >= 4096 blocks (16 per CU) while keeping chunks >= 8 planes
    long zc = (long(g.d) * nxy) / 4096;
    if (zc < 8) zc = 8;
    if (zc > 64) zc = 64;
    if (zc > g.d) zc = g.d;
    return int(zc);
}

// u (and optionally v) of the state as FP32 into dst (device), converting BF16.
amb_status state_field_to_float(amb_context* c, int field, float* dst) {
    if (c->storage == 0) {
        AMB_HIP(hipMemcpyAsync(dst, c->fields[field], c->g.n() * 4, hipMemcpyDeviceToDevice,
                               c->stream));
    } else {
        AMB_HIP(launch_convert(c->stream, 1, c->fields[field], dst, c->g.n(), 0));
    }
    return AMB_OK;
}

const float* state_field_f32_view(amb_context* c, int field, int scratch_idx) {
    if (c->storage == 0) return static_cast<const float*>(c->fields[field]);
    if (launch_convert(c->stream, 1, c->fields[field], c->scratch[scratch_idx], c->g.n(), 0) !=
        hipSuccess)
        return nullptr;
    return c->scratch[scratch_idx];
}

amb_status copy_out(amb_context* c, amb_memory mem, float* dst, const float* src_dev) {
    if (!dst) return AMB_OK;
    AMB_HIP(hipMemcpyAsync(dst, src_dev, c->g.n() * 4, kind_out(mem), c->stream));
    return AMB_OK;
}

amb_status copy_field_out(amb_context* c, amb_memory mem, float* dst, int field) {
    if (!dst) return AMB_OK;
    if (c->storage == 0) {
        AMB_HIP(hipMemcpyAsync(dst, c->fields[field], c->g.n() * 4, kind_out(mem), c->stream));
    } else if (mem == AMB_MEM_DEVICE) {
        AMB_HIP(launch_convert(c->stream, 1, c->fields[field], dst, c->g.n(), 0));
    } else {
        AMB_HIP(launch_convert(c->stream, 1, c->fields[field], c->scratch[4], c->g.n(), 0));
        AMB_HIP(hipMemcpyAsync(dst, c->scratch[4], c->g.n() * 4, hipMemcpyDeviceToHost, c->stream));
    }
    return AMB_OK;
}

amb_status run(amb_context* c, uint32_t iterations, uint32_t interval, amb_iteration_callback cb,
               void* user, bool with_v);

amb_status begin(amb_context* c, amb_memory mem, const float* f, uint32_t w, uint32_t h,
                 uint32_t d, const amb_tgv_params* p) {
    AMB_TRY(check_geometry(w, h, d));
    AMB_TRY(check_params(p));
    if (!f) return fail(AMB_ERR_INVALID_ARGUMENT, "f is NULL");
    if ((p->flags & AMB_FLAG_PLANAR) && d != 1)
        return fail(AMB_ERR_INVALID_ARGUMENT, "AMB_FLAG_PLANAR needs depth 1");
    AMB_TRY(set_device(c));
    const Geometry g{int(w), int(h), int(d)};
    AMB_TRY(ensure_buffers(c, g, int(p->storage)));
    AMB_HIP(hipMemcpyAsync(c->f, f, g.n() * 4, kind_in(mem), c->stream));
    {
        ProfScope ps(c, P_INIT);
        AMB_HIP(launch_init(c->stream, c->storage, c->fields, c->f, g));
    }
    c->params = *p;
    c->sp.lambda = p->lambda; c->sp.alpha0 = p->alpha0; c->sp.alpha1 = p->alpha1;
    c->sp.tau = p->tau; c->sp.sigma = p->sigma; c->sp.theta = p->theta;
    c->sp.data_term = int(p->data_term);
    c->sp.tile = choose_tile(g, c->ov);
    c->sp.zchunk = choose_zchunk(g, c->sp.tile, int(p->storage), c->ov);
    c->sp.index64 = needs_index64(g, c->ov) ? 1 : 0;
    c->sp.sweep_pf = c->ov.sweep_pf;
    c->sp.energy_partials = nullptr;
    c->sp.tile_active = nullptr;
    c->sp.active = nullptr;
    c->sp.planar = (p->flags & AMB_FLAG_PLANAR) ? 1 : 0;
    const bool fdiv = fused_divisors_ok(c->sp.alpha0, c->sp.alpha1);
    c->fused_ok = fdiv && fused_plan(g, c->storage, c->device, c->ov, &c->fplan);
    c->iplan_ok = fdiv && c->sp.planar && c->storage == 0 &&
                  fused_image_plan(g, c->device, c->ov, &c->iplan);
    c->fen_ok = c->fused_ok && !c->sp.planar &&
                fused_energy_plan(g, c->device, c->ov, c->fplan, &c->fen_plan) &&
                fused_grid_blocks(g, c->fen_plan) * 10 + 16 <= c->partials_cap;
    c->have_state = true;
    c->z_zero = true;
    return AMB_OK;
}

// Second copy of the 13 ping-ponged fields for the fused TGV2 sweep.  Field k
// of the state lives either in the state block or in the second block;
// c->alt[k] is set to the other one.  The second block is packed in the
// sweep's field order, so the two copies of different fields lie at different
// distances: a second block with the state block's layout (every field's
// copies one distance apart) made the fused sweep 2.4% slower on MI355X at
// 256^3 (384 vs 375 us, same box), also with the block offset by 4 KB or 1 MB.
amb_status ensure_alt(amb_context* c) {
    const size_t elem = c->storage == 1 ? 2 : 4;
    const size_t fstride = align_up(c->g.n() * elem, 256);
    const size_t need = fstride * N_FUSED_FIELDS;
    auto main_at = [&](int k) { return static_cast<void*>(c->block + size_t(FUSED_FIELD_MAP[k]) * fstride); };
    if (need > c->alt_cap) {
        for (int k = 0; k < N_FUSED_FIELDS; ++k)   // state must be in the state block
            if (c->fields[FUSED_FIELD_MAP[k]] != main_at(k))
                return fail(AMB_ERR_NOT_SUPPORTED, "fused-sweep buffers out of step with the state");
        if (c->alt_block) AMB_HIP(hipFree(c->alt_block));
        c->alt_block = nullptr;
        c->alt_cap = 0;
        AMB_HIP(hipMalloc(reinterpret_cast<void**>(&c->alt_block), need));
        c->alt_cap = need;
    }
    for (int k = 0; k < N_FUSED_FIELDS; ++k) {
        void* a = c->alt_block + size_t(k) * fstride;
        c->alt[k] = c->fields[FUSED_FIELD_MAP[k]] == main_at(k) ? a : main_at(k);
    }
    return AMB_OK;
}

// Primal half of iteration k + dual half of iteration k + 1 into the second
// copy, then swap the copies (the state table points at the new values).
// energy_out (device, 10 doubles; may be null): the launch also reduces the
// energy and gap of the state after iteration k (the fused sweep's energy form).
amb_status fused_step(amb_context* c, double* energy_out = nullptr) {
    if (energy_out) {
        ProfScope ps(c, P_FUSED_ENERGY);
        SweepParams sp = c->sp;
        sp.energy_partials = c->partials;
        AMB_HIP(launch_fused_sweep(c->stream, c->storage, c->fields, c->alt, c->f, c->g, sp,
                                   c->fen_plan));
        AMB_HIP(launch_energy_finish(c->stream, c->partials,
                                     int(fused_grid_blocks(c->g, c->fen_plan)), energy_out));
    } else {
        ProfScope ps(c, P_FUSED);
        if (c->sp.planar_zero_z && c->iplan_ok)
            AMB_HIP(launch_fused_image(c->stream, c->fields, c->alt, c->f, c->g, c->sp, c->iplan));
        else {
            AMB_HIP(launch_fused_sweep(c->stream, c->storage, c->fields, c->alt, c->f, c->g, c->sp,
                                       c->fplan_run));
            if (c->prof_on) {
                const size_t nt = fused_grid_blocks(c->g, c->fplan_run);
                c->prof_tiles_total += nt;
                c->prof_tiles_swept += c->sp.tile_active ? c->tact_active : nt;
            }
        }
    }
    for (int k = 0; k < N_FUSED_FIELDS; ++k) std::swap(c->fields[FUSED_FIELD_MAP[k]], c->alt[k]);
    return AMB_OK;
}

// The standalone FP64 energy pass over the current state into out (device).
amb_status energy_pass(amb_context* c, double* out) {
    ProfScope ps(c, P_ENERGY);
    AMB_HIP(launch_energy(c->stream, c->storage, c->fields, c->f, c->g, c->params.lambda,
                          c->params.alpha0, c->params.alpha1, int(c->params.data_term),
                          c->sp.planar, c->partials, out));
    return AMB_OK;
}

// Device slot of the next energy trace row (grow-only buffer; rows already
// written are copied into the larger one).
amb_status etrace_slot(amb_context* c, double** slot) {
    const size_t row = c->etrace_it.size();
    if (row >= c->etrace_cap) {
        const size_t cap = std::max<size_t>(64, 2 * c->etrace_cap);
        double* nb = nullptr;
        AMB_HIP(hipMalloc(reinterpret_cast<void**>(&nb), cap * 10 * sizeof(double)));
        if (c->etrace_dev) {
            AMB_HIP(hipMemcpyAsync(nb, c->etrace_dev, row * 10 * sizeof(double),
                                   hipMemcpyDeviceToDevice, c->stream));
            AMB_HIP(hipStreamSynchronize(c->stream));
            AMB_HIP(hipFree(c->etrace_dev));
        }
        c->etrace_dev = nb;
        c->etrace_cap = cap;
    }
    *slot = c->etrace_dev + row * 10;
    return AMB_OK;
}

// Copies the trace rows of the call to the host (iteration count + 10 values each).
amb_status etrace_collect(amb_context* c) {
    const size_t rows = c->etrace_it.size();
    std::vector<double> vals(rows * 10);
    if (rows) {
        AMB_HIP(hipMemcpyAsync(vals.data(), c->etrace_dev, rows * 10 * sizeof(double),
                               hipMemcpyDeviceToHost, c->stream));
        AMB_HIP(hipStreamSynchronize(c->stream));
    }
    c->etrace.assign(rows * 11, 0.0);
    for (size_t r = 0; r < rows; ++r) {
        c->etrace[r * 11] = double(c->etrace_it[r]);
        for (int k = 0; k < 10; ++k) c->etrace[r * 11 + 1 + k] = vals[r * 10 + k];
    }
    return AMB_OK;
}

// Active-set buffer for the masked iteration (n bytes, 256-B aligned so the
// vector sweeps can read four bytes at a time).
amb_status ensure_active(amb_context* c) {
    const size_t need = align_up(c->g.n(), 256);
    if (need > c->act_cap) {
        if (c->act) AMB_HIP(hipFree(c->act));
        c->act = nullptr;
        c->act_cap = 0;
        AMB_HIP(hipMalloc(reinterpret_cast<void**>(&c->act), need));
        c->act_cap = need;
    }
    return AMB_OK;
}

// Runs the iteration with updates restricted to c->act; the restriction is
// dropped again on every exit path.
amb_status run_masked(amb_context* c, uint32_t iterations, uint32_t interval,
                      amb_iteration_callback cb, void* user) {
    struct Reset {
        amb_context* c;
        ~Reset() { c->sp.active = nullptr; c->sp.tile_active = nullptr; }
    } reset{c};
    c->sp.active = c->act;
    return run(c, iterations, interval, cb, user, true);
}

amb_status sweep(amb_context* c, int which) {
    if (which == 0) {
        ProfScope ps(c, P_DUAL);
        AMB_HIP(launch_dual_sweep(c->stream, c->storage, c->fields, c->g, c->sp));
    } else {
        ProfScope ps(c, P_PRIMAL);
        AMB_HIP(launch_primal_sweep(c->stream, c->storage, c->fields, c->f, c->g, c->sp));
    }
    return AMB_OK;
}

// The reference loop with its callback placement (tgv2_l1_deshade.cu:169-272,
// callback rule tgv2_common.cu:381-383).
//
// With the fused sweep the loop runs dual_0, fused(primal_k + dual_k+1) for
// k = 0 .. N - 2, primal_N-1: the same updates in the same order.  An iteration
// whose end is observed by the callback ends with the primal sweep alone (the
// next one then starts with the dual sweep), so the callback and an early stop
// see exactly the state of the two-sweep loop.  u_bar and v_bar are stored
// only by the primal sweep; the last iteration is always a primal sweep.
amb_status run(amb_context* c, uint32_t iterations, uint32_t interval, amb_iteration_callback cb,
               void* user, bool with_v) {
    const bool image = c->iplan_ok && c->z_zero;
    const bool fused = (c->fused_ok || image) && iterations > 1;
    if (fused) AMB_TRY(ensure_alt(c));
    // the 2-D loop keeps v_z, p_z, q_zz, q_xz, q_yz at +0 in the state; with
    // their second copies cleared too, the fused sweep skips them
    c->sp.planar_zero_z = fused && c->sp.planar && c->storage == 0 && c->z_zero ? 1 : 0;
    if (c->sp.planar_zero_z) {
        const int zf[5] = {3, 6, 9, 11, 12};   // vz, pz, qzz, qxz, qyz in the fused order
        for (int k : zf) AMB_HIP(launch_zero_fill(c->stream, c->alt[k], c->g.n() * 4));
    }
    // masked 3-D loop: tiles of the fused plan that own no active voxel skip,
    // from the second fused launch of the call on (the first one writes every
    // voxel of the second copy, after which both copies agree outside the
    // active set: the masked sweeps store only active voxels)
    uint8_t* tflags = nullptr;
    c->sp.tile_active = nullptr;
    c->fplan_run = c->fplan;
    if (fused && c->sp.active && !image && !c->sp.planar_zero_z) {
        // the z chunk from the mask: the iteration plan's, 16 or 8 planes, whichever
        // gives the fewest rounds of active tiles (one resident block per CU) x
        // (planes + 5) (finer chunks isolate more inactive voxels and spread the
        // active ones over more CUs; each tile also marches one halo plane and
        // pays a ramp).  Measured at 256^3 per launch: ball of 51% of the voxels
        // 362 us with 64 planes, 376 with 16; 18%: 318 / 256 us; 6.5%: 276 / 165 us.
        FusedPlan cand[3] = {c->fplan, c->fplan, c->fplan};
        cand[1].zchunk = std::min(c->fplan.zchunk, 16);
        cand[2].zchunk = std::min(c->fplan.zchunk, 8);
        const int ncand = c->ov.fused_zchunk > 0 ? 1 : 3;   // a forced z chunk stays
        size_t need = 0;
        for (int k = 0; k < ncand; ++k) need += align_up(fused_grid_blocks(c->g, cand[k]), 256);
        if (need > c->tact_cap) {
            if (c->tact) AMB_HIP(hipFree(c->tact));
            c->tact = nullptr;
            c->tact_cap = 0;
            AMB_HIP(hipMalloc(reinterpret_cast<void**>(&c->tact), need));
            c->tact_cap = need;
        }
        size_t off[3] = {0, 0, 0};
        for (int k = 0; k < ncand; ++k) {
            off[k] = k ? off[k - 1] + align_up(fused_grid_blocks(c->g, cand[k - 1]), 256) : 0;
            AMB_HIP(launch_tile_activity(c->stream, c->sp.active, c->g, cand[k], c->tact + off[k]));
        }
        std::vector<uint8_t> h(need);
        AMB_HIP(hipMemcpyAsync(h.data(), c->tact, need, hipMemcpyDeviceToHost, c->stream));
        AMB_HIP(hipStreamSynchronize(c->stream));
        size_t active[3] = {0, 0, 0};
        for (int k = 0; k < ncand; ++k) {
            const size_t nt = fused_grid_blocks(c->g, cand[k]);
            for (size_t t = 0; t < nt; ++t) active[k] += h[off[k] + t] != 0;
        }
        int best = 0;
        int cus = 0;
        if (ncand > 1 &&
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->device) ==
                hipSuccess && cus > 0) {
            double best_cost = 0.0;
            for (int k = 0; k < ncand; ++k) {
                const double rounds = double((active[k] + size_t(cus) - 1) / size_t(cus));
                const double cost = rounds * double(cand[k].zchunk + 5);
                if (k == 0 || cost < best_cost) { best = k; best_cost = cost; }
            }
        }
        c->tact_active = active[best];
        c->fplan_run = cand[best];
        tflags = c->tact + off[best];
    }
    bool need_dual = true, fused_ran = false;
    const uint32_t every = c->energy_interval;
    const bool en_fused = fused && !image && c->fen_ok && !c->sp.planar_zero_z;
    c->etrace_it.clear();
    for (uint32_t it = 0; it < iterations; ++it) {
        const bool observed = cb && interval > 0 && it > 0 && it % interval == 0;
        // energy after this iteration: every `every` iterations and after the last
        const bool want_e = every > 0 && ((it + 1) % every == 0 || it + 1 == iterations);
        if (need_dual) AMB_TRY(sweep(c, 0));
        if (fused && !observed && it + 1 < iterations && (!want_e || en_fused)) {
            double* slot = nullptr;
            if (want_e) AMB_TRY(etrace_slot(c, &slot));
            c->sp.tile_active = fused_ran ? tflags : nullptr;
            AMB_TRY(fused_step(c, slot));
            fused_ran = true;
            need_dual = false;
        } else {
            AMB_TRY(sweep(c, 1));
            need_dual = true;
            if (want_e) {
                double* slot = nullptr;
                AMB_TRY(etrace_slot(c, &slot));
                AMB_TRY(energy_pass(c, slot));
            }
        }
        if (want_e) c->etrace_it.push_back(it + 1);
        if (observed) {
            const size_t n = c->g.n();
            c->cb_u.resize(n);
            AMB_TRY(copy_field_out(c, AMB_MEM_HOST, c->cb_u.data(), F_U));
            if (with_v)
                for (int k = 0; k < 3; ++k) {
                    c->cb_v[k].resize(n);
                    AMB_TRY(copy_field_out(c, AMB_MEM_HOST, c->cb_v[k].data(), F_VX + k));
                }
            AMB_HIP(hipStreamSynchronize(c->stream));
            const int stop = cb(it, iterations, c->cb_u.data(), with_v ? c->cb_v[0].data() : nullptr,
                                with_v ? c->cb_v[1].data() : nullptr,
                                with_v ? c->cb_v[2].data() : nullptr, user);
            if (stop) {
                if (every > 0 && !want_e) {   // the energy where the loop stopped
                    double* slot = nullptr;
                    AMB_TRY(etrace_slot(c, &slot));
                    AMB_TRY(energy_pass(c, slot));
                    c->etrace_it.push_back(it + 1);
                }
                break;
            }
        }
    }
    c->sp.tile_active = nullptr;
    if (every > 0) AMB_TRY(etrace_collect(c));
    return AMB_OK;
}

// TGV-k (k >= 3): the TGV2 state from begin() plus k - 2 levels of w, w_bar, r,
// all zero at the start (tgvk_deshade/tgvk_l1_deshade.cu:116-151).
amb_status tgvk_begin(amb_context* c, amb_memory mem, const float* f, uint32_t w, uint32_t h,
                      uint32_t d, const amb_tgv_params* p, uint32_t order, const float* alpha) {
    if (!p) return fail(AMB_ERR_INVALID_ARGUMENT, "params is NULL");
    if (order < 3 || order > 1000) return fail(AMB_ERR_INVALID_ARGUMENT, "order must be 3 .. 1000");
    if (!alpha) return fail(AMB_ERR_INVALID_ARGUMENT, "alpha is NULL");
    for (uint32_t k = 0; k < order; ++k)
        if (!(alpha[k] > 0.f)) return fail(AMB_ERR_INVALID_ARGUMENT, "alpha weights must be positive");
    if (p->data_term != AMB_DATA_L1)
        return fail(AMB_ERR_NOT_SUPPORTED, "the TGV-k deshade has the L1 data term only");
    if (p->storage != AMB_STORAGE_F32)
        return fail(AMB_ERR_NOT_SUPPORTED, "the TGV-k deshade runs on FP32 fields");
    if (h > 4u * 65535u) return fail(AMB_ERR_NOT_SUPPORTED, "height above 262140");
    amb_tgv_params pk = *p;   // the TGV2 part's weights come from alpha
    pk.alpha0 = alpha[1];
    pk.alpha1 = alpha[0];
    AMB_TRY(begin(c, mem, f, w, h, d, &pk));
    const int L = int(order) - 2;
    const size_t fstride = align_up(c->g.n() * 4, 256);
    const size_t need = fstride * 18 * size_t(L);
    if (need > c->kcap) {
        if (c->kblock) AMB_HIP(hipFree(c->kblock));
        c->kblock = nullptr;
        c->kcap = 0;
        AMB_HIP(hipMalloc(reinterpret_cast<void**>(&c->kblock), need));
        c->kcap = need;
    }
    AMB_HIP(launch_zero_fill(c->stream, c->kblock, need));
    const size_t nt = 18 * size_t(L);
    if (nt > c->klv_cap) {
        if (c->klv_dev) AMB_HIP(hipFree(c->klv_dev));
        c->klv_dev = nullptr;
        c->klv_cap = 0;
        AMB_HIP(hipMalloc(reinterpret_cast<void**>(&c->klv_dev), nt * sizeof(float*)));
        c->klv_cap = nt;
    }
    if (order > c->kalpha_cap) {
        if (c->kalpha_dev) AMB_HIP(hipFree(c->kalpha_dev));
        c->kalpha_dev = nullptr;
        c->kalpha_cap = 0;
        AMB_HIP(hipMalloc(reinterpret_cast<void**>(&c->kalpha_dev), order * sizeof(float)));
        c->kalpha_cap = order;
    }
    // the tables stay in the context until the copies have run (next sync)
    AMB_HIP(hipStreamSynchronize(c->stream));
    c->klv_host.resize(nt);
    for (size_t j = 0; j < nt; ++j)
        c->klv_host[j] = reinterpret_cast<float*>(c->kblock + j * fstride);
    c->kalpha_host.assign(alpha, alpha + order);
    AMB_HIP(hipMemcpyAsync(c->klv_dev, c->klv_host.data(), nt * sizeof(float*),
                           hipMemcpyHostToDevice, c->stream));
    AMB_HIP(hipMemcpyAsync(c->kalpha_dev, c->kalpha_host.data(), order * sizeof(float),
                           hipMemcpyHostToDevice, c->stream));
    AMB_HIP(hipMemcpyAsync(c->fields_dev, c->fields, sizeof(void*) * N_FIELDS,
                           hipMemcpyHostToDevice, c->stream));
    c->klevels = L;
    c->kov = c->ov;
    c->kfused_ok = tgvk_fused_plan(c->g, L, c->device, c->kov, &c->kfplan);
    return AMB_OK;
}

// Level-table slots of w_0 (0..5) and r_0 (12..17), the level fields the
// fused TGV-k sweep ping-pongs; c->kalt[m] is the other copy of slot KLV_PING[m].
constexpr int KLV_PING[12] = {0, 1, 2, 3, 4, 5, 12, 13, 14, 15, 16, 17};

// The fused TGV-k sweep reaches every destination as its source pointer plus
// one distance per buffer (25 destination pointers would not fit the scalar
// registers), so its second copies mirror the first copies' layouts: the 13
// TGV2 fields in a block of the state block's layout (the u_bar and v_bar
// slots unused), the level fields in a block of the level block's layout.
// (Measured on MI355X at 256^3, same box: 808 us per launch this way; 829 us
// with each field's second copy in the slot of another field, which the TGV2
// sweep prefers - see ensure_alt.)

amb_status ensure_kalt13(amb_context* c) {
    const size_t fstride = align_up(c->g.n() * 4, 256);
    const size_t need = fstride * N_FIELDS;
    auto main_at = [&](int k) { return static_cast<void*>(c->block + size_t(FUSED_FIELD_MAP[k]) * fstride); };
    if (need > c->kalt13_cap) {
        for (int k = 0; k < N_FUSED_FIELDS; ++k)   // state must be in the state block
            if (c->fields[FUSED_FIELD_MAP[k]] != main_at(k))
                return fail(AMB_ERR_NOT_SUPPORTED, "fused TGV-k buffers out of step with the state");
        if (c->kalt13_block) AMB_HIP(hipFree(c->kalt13_block));
        c->kalt13_block = nullptr;
        c->kalt13_cap = 0;
        AMB_HIP(hipMalloc(reinterpret_cast<void**>(&c->kalt13_block), need));
        c->kalt13_cap = need;
    }
    for (int k = 0; k < N_FUSED_FIELDS; ++k) {
        void* a = c->kalt13_block + size_t(FUSED_FIELD_MAP[k]) * fstride;
        c->kalt13[k] = c->fields[FUSED_FIELD_MAP[k]] == main_at(k) ? a : main_at(k);
    }
    return AMB_OK;
}

amb_status ensure_kalt(amb_context* c) {
    const size_t fstride = align_up(c->g.n() * 4, 256);
    const size_t need = fstride * 18;
    auto main_at = [&](int j) { return reinterpret_cast<float*>(c->kblock + size_t(j) * fstride); };
    if (need > c->kalt_cap) {
        for (int m = 0; m < 12; ++m)   // state must be in the level block
            if (c->klv_host[KLV_PING[m]] != main_at(KLV_PING[m]))
                return fail(AMB_ERR_NOT_SUPPORTED, "fused TGV-k buffers out of step with the state");
        if (c->kalt_block) AMB_HIP(hipFree(c->kalt_block));
        c->kalt_block = nullptr;
        c->kalt_cap = 0;
        AMB_HIP(hipMalloc(reinterpret_cast<void**>(&c->kalt_block), need));
        c->kalt_cap = need;
    }
    for (int m = 0; m < 12; ++m) {
        const int j = KLV_PING[m];
        float* a = reinterpret_cast<float*>(c->kalt_block + size_t(j) * fstride);
        c->kalt[m] = c->klv_host[j] == main_at(j) ? a : main_at(j);
    }
    return AMB_OK;
}

// Primal half of TGV-k iteration k + dual half of iteration k + 1 into the
// second copies, then swap the copies in the host tables.
amb_status tgvk_fused_step(amb_context* c) {
    const float* src[N_TGVK_FUSED_FIELDS];
    float* dst[N_TGVK_FUSED_FIELDS];
    for (int k = 0; k < N_FUSED_FIELDS; ++k) {
        src[k] = static_cast<const float*>(c->fields[FUSED_FIELD_MAP[k]]);
        dst[k] = static_cast<float*>(c->kalt13[k]);
    }
    for (int m = 0; m < 12; ++m) {
        src[N_FUSED_FIELDS + m] = c->klv_host[KLV_PING[m]];
        dst[N_FUSED_FIELDS + m] = c->kalt[m];
    }
    // every field's other copy lies at one distance per block (ensure_kalt13, ensure_kalt)
    for (int k = 1; k < N_TGVK_FUSED_FIELDS; ++k) {
        const int k0 = k < N_FUSED_FIELDS ? 0 : N_FUSED_FIELDS;
        if (reinterpret_cast<const char*>(dst[k]) - reinterpret_cast<const char*>(src[k]) !=
            reinterpret_cast<const char*>(dst[k0]) - reinterpret_cast<const char*>(src[k0]))
            return fail(AMB_ERR_NOT_SUPPORTED, "fused TGV-k buffers out of step with the state");
    }
    {
        ProfScope ps(c, P_TGVK_FUSED);
        AMB_HIP(launch_tgvk_fused(c->stream, src, dst, c->f, c->sp.active, c->g, c->sp.tau,
                                  c->sp.sigma, c->sp.lambda, c->sp.theta, c->kalpha_host.data(),
                                  c->sp.planar, c->kfplan));
    }
    for (int k = 0; k < N_FUSED_FIELDS; ++k) std::swap(c->fields[FUSED_FIELD_MAP[k]], c->kalt13[k]);
    for (int m = 0; m < 12; ++m) std::swap(c->klv_host[KLV_PING[m]], c->kalt[m]);
    return AMB_OK;
}

// Device copies of the field and level tables after fused steps moved the
// state; synchronises, so the host tables may change again afterwards.
amb_status tgvk_upload_tables(amb_context* c) {
    AMB_HIP(hipMemcpyAsync(c->fields_dev, c->fields, sizeof(void*) * N_FIELDS,
                           hipMemcpyHostToDevice, c->stream));
    AMB_HIP(hipMemcpyAsync(c->klv_dev, c->klv_host.data(), c->klv_host.size() * sizeof(float*),
                           hipMemcpyHostToDevice, c->stream));
    AMB_HIP(hipStreamSynchronize(c->stream));
    return AMB_OK;
}

// The TGV-k loop: per iteration one dual launch (p, q, every r_i) and one
// primal launch (u, v, every w_i); callback rule as in run().  With the fused
// sweep (one level) the loop runs dual_0, fused(primal_k + dual_k+1) for
// k = 0 .. N - 2, primal_N-1, and an observed iteration ends with the primal
// launch alone, exactly as run() does for TGV2.  The dual and primal launches
// read the state through the device tables, which are brought up to date
// after fused steps have swapped the copies.
amb_status tgvk_run(amb_context* c, uint32_t iterations, uint32_t interval,
                    amb_iteration_callback cb, void* user) {
    float* const* F = reinterpret_cast<float* const*>(c->fields_dev);
    const bool fused = c->kfused_ok && iterations > 1;
    if (fused) {
        AMB_TRY(ensure_kalt13(c));
        AMB_TRY(ensure_kalt(c));
        // earlier uploads from the host tables have run before they change
        AMB_HIP(hipStreamSynchronize(c->stream));
    }
    bool need_dual = true, stale = false;
    for (uint32_t it = 0; it < iterations; ++it) {
        const bool observed = cb && interval > 0 && it > 0 && it % interval == 0;
        if (need_dual) {
            ProfScope ps(c, P_TGVK_DUAL);
            AMB_HIP(launch_tgvk_dual(c->stream, F, c->klv_dev, c->kalpha_dev, c->sp.active, c->g,
                                     c->klevels, c->sp.sigma, c->sp.planar, c->kov));
        }
        if (fused && !observed && it + 1 < iterations) {
            AMB_TRY(tgvk_fused_step(c));
            need_dual = false;
            stale = true;
        } else {
            if (stale) AMB_TRY(tgvk_upload_tables(c));
            stale = false;
            ProfScope ps(c, P_TGVK_PRIMAL);
            AMB_HIP(launch_tgvk_primal(c->stream, F, c->klv_dev, c->f, c->sp.active, c->g,
                                       c->klevels, c->sp.tau, c->sp.lambda, c->sp.theta,
                                       c->kov));
            need_dual = true;
        }
        if (observed) {
            const size_t n = c->g.n();
            c->cb_u.resize(n);
            AMB_TRY(copy_field_out(c, AMB_MEM_HOST, c->cb_u.data(), F_U));
            for (int k = 0; k < 3; ++k) {
                c->cb_v[k].resize(n);
                AMB_TRY(copy_field_out(c, AMB_MEM_HOST, c->cb_v[k].data(), F_VX + k));
            }
            AMB_HIP(hipStreamSynchronize(c->stream));
            if (cb(it, iterations, c->cb_u.data(), c->cb_v[0].data(), c->cb_v[1].data(),
                   c->cb_v[2].data(), user))
                break;
        }
    }
    return AMB_OK;
}

// The TGV-k loop with updates restricted to c->act (dropped again on exit).
amb_status tgvk_run_masked(amb_context* c, uint32_t iterations, uint32_t interval,
                           amb_iteration_callback cb, void* user) {
    struct Reset {
        amb_context* c;
        ~Reset() { c->sp.active = nullptr; }
    } reset{c};
    c->sp.active = c->act;
    return tgvk_run(c, iterations, interval, cb, user);
}

amb_status check_precision(const Geometry& g, int precision) {
    if (precision == AMB_DCT_REFERENCE) return AMB_OK;
    const int mx = dct_mfma_max_extent();
    if (g.w > mx || g.h > mx || g.d > mx)
        return fail(AMB_ERR_NOT_SUPPORTED,
                    "MFMA DCT passes support extents up to 4096; use AMB_DCT_REFERENCE");
    return AMB_OK;
}

// One separable 3-pass transform from src into dst (src is overwritten as
// scratch).  poisson_flags >= 0 fuses the eigenvalue divide into the last
// forward pass (MFMA precisions only).
amb_status transform3(amb_context* c, float* src, float* tmp, float* dst, int inverse,
                      int precision, int poisson_flags = -1) {
    const Geometry g = c->g;
    AMB_TRY(check_precision(g, precision));
    AMB_TRY(ensure_tables(c, g));
    float* bufs[4] = {src, tmp, src, dst};   // x: src->tmp, y: tmp->src, z: src->dst
    for (int axis = 0; axis < 3; ++axis) {
        ProfScope ps(c, P_DCT);
        const AxisTables& t = c->axes[axis];
        const int last = axis == 2;
        if (precision == AMB_DCT_REFERENCE) {
            AMB_HIP(launch_dct_reference_pass(c->stream, bufs[axis], bufs[axis + 1],
                                              inverse ? t.inv64 : t.fwd64, g, axis, inverse, last));
        } else {
            const int fuse = (last && !inverse && poisson_flags >= 0) ? 1 : 0;
            const int quirk = (poisson_flags > 0 && (poisson_flags & AMB_FLAG_POISSON_INDEX_QUIRK)) ? 1 : 0;
            AMB_HIP(launch_dct_mfma_pass(c->stream, bufs[axis], bufs[axis + 1],
                                         inverse ? t.inv32 : t.fwd32, inverse ? t.inv_hi : t.fwd_hi,
                                         inverse ? t.inv_lo : t.fwd_lo, g, axis, inverse, last,
                                         precision, c->axes[0].eig, c->axes[1].eig, c->axes[2].eig,
                                         fuse, quirk));
        }
    }
    return AMB_OK;
}

amb_status poisson(amb_context* c, const float* in, float* out, uint32_t flags) {
    AMB_TRY(ensure_tables(c, c->g));
    ProfScope ps(c, P_POISSON);
    AMB_HIP(launch_poisson_divide(c->stream, in, out, c->axes[0].eig, c->axes[1].eig,
                                  c->axes[2].eig, c->g, (flags & AMB_FLAG_POISSON_INDEX_QUIRK) ? 1 : 0));
    return AMB_OK;
}

// l = iDCT(P(DCT(B-div(vx, vy, vz)))) into scratch[1]; vx.. are device FP32.
amb_status integrate(amb_context* c, const float* vx, const float* vy, const float* vz,
                     int precision, uint32_t flags, float** result) {
    float* s0 = c->scratch[0];
    float* s1 = c->scratch[1];
    float* s2 = c->scratch[2];
    {
        ProfScope ps(c, P_DIV);
        AMB_HIP(launch_divergence(c->stream, vx, vy, vz, s0, c->g));
    }
    if (precision == AMB_DCT_REFERENCE) {
        AMB_TRY(transform3(c, s0, s1, s2, 0, precision));  // DCT -> s2
        AMB_TRY(poisson(c, s2, s0, flags));                 // -> s0
        AMB_TRY(transform3(c, s0, s2, s1, 1, precision));  // iDCT -> s1
    } else {
        // eigenvalue divide fused into the forward z pass
        AMB_TRY(transform3(c, s0, s1, s2, 0, precision, int(flags & AMB_FLAG_POISSON_INDEX_QUIRK)));
        AMB_TRY(transform3(c, s2, s0, s1, 1, precision));  // iDCT -> s1
    }
    *result = s1;
    return AMB_OK;
}

// Upload three host/device volumes into scratch[2..4] (device FP32 views).
amb_status stage3(amb_context* c, amb_memory mem, const float* a, const float* b, const float* d,
                  const float** out) {
    const float* in[3] = {a, b, d};
    for (int k = 0; k < 3; ++k) {
        if (!in[k]) {
            if (k == 2 && c->g.d == 1) {
                out[k] = c->scratch[2 + k];   // unused for depth 1
                continue;
            }
            return fail(AMB_ERR_INVALID_ARGUMENT, "gradient component is NULL");
        }
        if (mem == AMB_MEM_DEVICE) {
            out[k] = in[k];
        } else {
            AMB_HIP(hipMemcpyAsync(c->scratch[2 + k], in[k], c->g.n() * 4, hipMemcpyHostToDevice,
                                   c->stream));
            out[k] = c->scratch[2 + k];
        }
    }
    return AMB_OK;
}

amb_status geometry_only(amb_context* c, uint32_t w, uint32_t h, uint32_t d) {
    AMB_TRY(check_geometry(w, h, d));
    AMB_TRY(set_device(c));
    const Geometry g{int(w), int(h), int(d)};
    // keep the iteration state if the geometry matches, else (re)allocate FP32
    if (!(c->have_state && c->g.w == g.w && c->g.h == g.h && c->g.d == g.d)) {
        AMB_TRY(ensure_buffers(c, g, c->have_state ? c->storage : 0));
        c->have_state = false;
    }
    return AMB_OK;
}

// Every amb_status entry point is a function-try-block ending in
// guard_exception(): no C++ exception (a std::bad_alloc of a host vector, say)
// crosses the C ABI; it becomes a status code and amb_last_error() text.
amb_status guard_exception() noexcept {
    try {
        throw;
    } catch (const std::bad_alloc&) {
        return fail(AMB_ERR_OUT_OF_MEMORY, "host memory allocation failed");
    } catch (const std::exception& e) {
        return fail(AMB_ERR_HIP, std::string("internal error: ") + e.what());
    } catch (...) {
        return fail(AMB_ERR_HIP, "internal error: unknown C++ exception");
    }
}

}  // namespace



// ---------------------------------------------------------------------------
// amb_tgv2_energy_noalloc: the energy pass for callers that cannot allocate.
// Its block partials live in `scratch`, a fixed device array: 4096 rows of 10
// doubles (the energy grid is capped at 4096 blocks, so every volume fits) plus
// 16 spare, 327,808 bytes per device, held until the process ends.  Each block
// writes its whole row and the final sum reads only the rows written, so
// nothing has to start at zero.  In host code `scratch` names a host-side
// stand-in: its device address is taken once per device with
// hipGetSymbolAddress (that device current) and kept in g_scratch_ptr.
// ---------------------------------------------------------------------------
constexpr size_t NOALLOC_SCRATCH_DOUBLES = 4096 * 10 + 16;
static __device__ __align__(256) double scratch[NOALLOC_SCRATCH_DOUBLES];

namespace {

std::mutex g_scratch_mutex;
std::vector<double*> g_scratch_ptr;   // per device; null until resolved

amb_status scratch_ptr(int device, double** out) {
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (device >= int(g_scratch_ptr.size())) g_scratch_ptr.resize(size_t(device) + 1, nullptr);
    if (!g_scratch_ptr[device]) {
        void* p = nullptr;
        AMB_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return fail(AMB_ERR_HIP, "hipGetSymbolAddress returned no address for scratch");
        g_scratch_ptr[device] = static_cast<double*>(p);
    }
    *out = g_scratch_ptr[device];
    return AMB_OK;
}

}  // namespace

// ===========================================================================
extern "C" {

AMB_API const char* amb_last_error(void) { return g_last_error.c_str(); }
AMB_API const char* amb_version(void) { return "ambient_amd 0.1 gfx950"; }

AMB_API void amb_tgv_params_default(amb_tgv_params* p) {
    if (!p) return;
    std::memset(p, 0, sizeof(*p));
    p->lambda = 1.0f;
    p->alpha0 = 2.0f;
    p->alpha1 = 1.0f;
    p->tau = float(1.0 / std::sqrt(20.0));   // tgv2_l1_deshade.cu:148 (Pixel tau = 1.0 / sqrt(20.0))
    p->sigma = p->tau;
    p->theta = 1.0f;
    p->iterations = 10000;
    p->data_term = AMB_DATA_L1;
    p->storage = AMB_STORAGE_F32;
    p->dct_precision = AMB_DCT_REFERENCE;
    p->callback_interval = 0;
    p->flags = AMB_FLAG_SET_NEGATIVE_ZERO;
}

AMB_API amb_status amb_context_create(int device, amb_context** out) try {
    if (!out) return fail(AMB_ERR_INVALID_ARGUMENT, "ctx is NULL");
    *out = nullptr;
    int count = 0;
    AMB_HIP(hipGetDeviceCount(&count));
    if (device < 0 || device >= count) return fail(AMB_ERR_INVALID_ARGUMENT, "no such device");
    AMB_HIP(hipSetDevice(device));
    amb_context* c = new amb_context();
    c->device = device;
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete c;
        return fail_hip(e, "hipStreamCreateWithFlags", __LINE__);
    }
    *out = c;
    return AMB_OK;
} catch (...) {
    return guard_exception();
}

AMB_API amb_status amb_context_destroy(amb_context* c) try {
    if (!c) return AMB_OK;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    for (int k = 0; k < N_PROF; ++k)
        for (auto& pr : c->pending[k]) {
            (void)hipEventDestroy(pr.first);
            (void)hipEventDestroy(pr.second);
        }
    for (hipEvent_t e : c->pool) (void)hipEventDestroy(e);
    for (auto& t : c->axes) free_axis(t);
    if (c->block) (void)hipFree(c->block);
    if (c->fblock) (void)hipFree(c->fblock);
    if (c->fields_dev) (void)hipFree(c->fields_dev);
    if (c->partials) (void)hipFree(c->partials);
    if (c->energy_out) (void)hipFree(c->energy_out);
    if (c->etrace_dev) (void)hipFree(c->etrace_dev);
    if (c->adj_block) (void)hipFree(c->adj_block);
    if (c->act) (void)hipFree(c->act);
    if (c->tact) (void)hipFree(c->tact);
    if (c->alt_block) (void)hipFree(c->alt_block);
    if (c->kblock) (void)hipFree(c->kblock);
    if (c->klv_dev) (void)hipFree(c->klv_dev);
    if (c->kalpha_dev) (void)hipFree(c->kalpha_dev);
    if (c->kalt_block) (void)hipFree(c->kalt_block);
    if (c->kalt13_block) (void)hipFree(c->kalt13_block);
    if (c->rblock) (void)hipFree(c->rblock);
    (void)hipStreamDestroy(c->stream);
    delete c;
    return AMB_OK;
} catch (...) {
    return guard_exception();
}

AMB_API amb_status amb_device_alloc(amb_context* c, size_t bytes, void** ptr) try {
    if (!c || !ptr) return fail(AMB_ERR_INVALID_ARGUMENT, "NULL argument");
    AMB_TRY(set_device(c));
    AMB_HIP(hipMalloc(ptr, bytes));
    return AMB_OK;
} catch (...) {
    return guard_exception();
}

AMB_API amb_status amb_device_free(amb_context* c, void* ptr) try {
    if (!c) return fail(AMB_ERR_INVALID_ARGUMENT, "ctx is NULL");
    AMB_TRY(set_device(c));
    AMB_HIP(hipStreamSynchronize(c->stream));
    AMB_HIP(hipFree(ptr));
    return AMB_OK;
} catch (...) {
    return guard_exception();
}

AMB_API amb_status amb_memcpy(amb_context* c, void* dst, const void* src, size_t bytes, int kind) try {
    if (!c || (!dst && bytes) || (!src && bytes)) return fail(AMB_ERR_INVALID_ARGUMENT, "NULL argument");
    AMB_TRY(set_device(c));
    const hipMemcpyKind k = kind == 0 ? hipMemcpyHostToDevice
                          : kind == 1 ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice;
    AMB_HIP(hipMemcpyAsync(dst, src, bytes, k, c->stream));
    AMB_HIP(hipStreamSynchronize(c->stream));
    return AMB_OK;
} catch (...) {
    return guard_exception();
}

AMB_API amb_status amb_synchronize(amb_context* c) try {
    if (!c) return fail(AMB_ERR_INVALID_ARGUMENT, "ctx is NULL");
    AMB_TRY(set_device(c));
    AMB_HIP(hipStreamSynchronize(c->stream));
    return AMB_OK;
} catch (...) {
    return guard_exception();
}

AMB_API void* amb_context_stream(amb_context* c) { return c ? static_cast<void*>(c->stream) : nullptr; }

AMB_API amb_status amb_context_set_plan_override(amb_context* c, const char* name,
                                                 int32_t value) try {
    if (!c || !name) return fail(AMB_ERR_INVALID_ARGUMENT, "NULL argument");
    PlanOverrides& o = c->ov;
    const PlanOverrides unset{};
    struct Key {
        const char* name;
        int PlanOverrides::*field;
    };
    static const Key keys[] = {
        {"tile", &PlanOverrides::tile}, {"zchunk", &PlanOverrides::zchunk},
        {"sweep_pf", &PlanOverrides::sweep_pf}, {"fused", &PlanOverrides::fused},
        {"fused_vx", &PlanOverrides::fused_vx}, {"fused_ty", &PlanOverrides::fused_ty},
        {"adjoint_vx", &PlanOverrides::adjoint_vx},
        {"fused_zchunk", &PlanOverrides::fused_zchunk}, {"fused_tx", &PlanOverrides::fused_tx},
        {"fused_pf", &PlanOverrides::fused_pf}, {"fused_image", &PlanOverrides::fused_image},
        {"fused_image_tx", &PlanOverrides::fused_image_tx},
        {"fused_image_zchunk", &PlanOverrides::fused_image_zchunk},
        {"index64", &PlanOverrides::index64}, {"tgvk_zchunk", &PlanOverrides::tgvk_zchunk},
        {"tgvk_fused", &PlanOverrides::tgvk_fused},
        {"tgvk_fused_tx", &PlanOverrides::tgvk_fused_tx},
        {"tgvk_fused_ty", &PlanOverrides::tgvk_fused_ty},
        {"tgvk_fused_zchunk", &PlanOverrides::tgvk_fused_zchunk}};
    for (const Key& k : keys)
        if (std::strcmp(k.name, name) == 0) {
            o.*k.field = value < 0 ? unset.*k.field : value;
            return AMB_OK;
        }
    return fail(AMB_ERR_INVALID_ARGUMENT, std::string("unknown plan override: ") + name);
} catch (...) {
    return guard_exception();
}

AMB_API amb_status amb_context_clear_plan_overrides(amb_context* c) try {
    if (!c) return fail(AMB_ERR_INVALID_ARGUMENT, "ctx is NULL");
    c->ov = PlanOverrides{};
    return AMB_OK;
} catch (...) {
    return guard_exception();
}

AMB_API amb_status amb_tgv2_begin(amb_context* c, amb_memory mem, const float* f, uint32_t w,
                                  uint32_t h, uint32_t d, const amb_tgv_params* p) try {
    if (!c) return fail(AMB_ERR_INVALID_ARGUMENT, "ctx is NULL");
    AMB_TRY(begin(c, mem, f, w, h, d, p));
    AMB_HIP(hipStreamSynchronize(c->stream));
    return AMB_OK;
} catch (...) {
    return guard_exception();
}

AMB_API amb_status amb_tgv2_set_state(amb_context* c, amb_memory mem, const float* const* fields) try {
    if (!c || !fields) return fail(AMB_ERR_INVALID_ARGUMENT, "NULL argument");
    if (!c->have_state) return fail(AMB_ERR_INVALID_ARGUMENT, "no iteration state: call amb_tgv2_begin");
    AMB_TRY(set_device(c));
    const size_t n = c->g.n();
    c->z_zero = false;
    for (int k = 0; k < N_FIELDS; ++k) {
        if (!fields[k]) return fail(AMB_ERR_INVALID_ARGUMENT, "field pointer is NULL");
        if (c->storage == 0) {
            AMB_HIP(hipMemcpyAsync(c->fields[k], fields[k], n * 4, kind_in(mem), c->stream));
        } else {
            const float* src = fields[k];
            if (mem == AMB_MEM_HOST) {
                AMB_HIP(hipMemcpyAsync(c->scratch[4], fields[k], n * 4, hipMemcpyHostToDevice, c->stream));
                src = c->scratch[4];
            }
            AMB_HIP(launch_convert(c->stream, 1, src, c->fields[k], n, 1));
        }
    }
    AMB_HIP(hipStreamSynchronize(c->stream));
    return AMB_OK;
} catch (...) {
    return guard_exception();
}

AMB_API amb_status amb_tgv2_get_state(amb_context* c, amb_memory mem, float* const* fields) try {
    if (!c || !fields) return fail(AMB_ERR_INVALID_ARGUMENT, "NULL argument");
    if (!c->have_state) return fail(AMB_ERR_INVALID_ARGUMENT, "no iteration state");
    AMB_TRY(set_device(c));
    for (int k = 0; k < N_FIELDS; ++k) {
        AMB_TRY(copy_field_out(c, mem, fields[k], k));
        if (c->storage == 1 && mem == AMB_MEM_HOST) AMB_HIP(hipStreamSynchronize(c->stream));
    }
    AMB_HIP(hipStreamSynchronize(c->stream));
    return AMB_OK;
} catch (...) {
    return guard_exception();
}

AMB_API amb_status amb_tgv2_iterate(amb_context* c, uint32_t iterations) try {
    if (!c) return fail(AMB_ERR_INVALID_ARGUMENT, "ctx is NULL");
    if (!c->have_state) return fail(AMB_ERR_INVALID_ARGUMENT, "no iteration state: call amb_tgv2_begin");
    AMB_TRY(set_device(c));
    AMB_TRY(run(c, iterations, 0, nullptr, nullptr, false));
    AMB_HIP(hipStreamSynchronize(c->stream));
    return AMB_OK;
} catch (...) {
    return guard_exception();
}

AMB_API amb_status amb_tgv2_sweep(amb_context* c, int which) try {
    if (!c) return fail(AMB_ERR_INVALID_ARGUMENT, "ctx is NULL");
    if (!c->have_state) return fail(AMB_ERR_INVALID_ARGUMENT, "no iteration state: call amb_tgv2_begin");
    if (which != 0 && which != 1) return fail(AMB_ERR_INVALID_ARGUMENT, "which must be 0 or 1");
    AMB_TRY(set_device(c));
    AMB_TRY(sweep(c, which));
    AMB_HIP(hipStreamSynchronize(c->stream));
    return AMB_OK;
} catch (...) {
    return guard_exception();
}

AMB_API amb_status amb_tgv2_deshade(amb_context* c, amb_memory mem, const float* f, uint32_t w,
                                    uint32_t h, uint32_t d, const amb_tgv_params* p, float* u,
                                    float* vx, float* vy, float* vz, amb_iteration_callback cb,
                                    void* user) try {
    if (!c) return fail(AMB_ERR_INVALID_ARGUMENT, "ctx is NULL");
    AMB_TRY(begin(c, mem, f, w, h, d, p));
    AMB_TRY(run(c, p->iterations, p->callback_interval, cb, user, true));
    AMB_TRY(copy_field_out(c, mem, u, F_U));
    AMB_TRY(copy_field_out(c, mem, vx, F_VX));
    AMB_TRY(copy_field_out(c, mem, vy, F_VY));
    AMB_TRY(copy_field_out(c, mem, vz, F_VZ));
    AMB_HIP(hipStreamSynchronize(c->stream));
    return AMB_OK;
} catch (...) {
    return guard_exception();
}

AMB_API amb_status amb_tgv2_deshade_masked(amb_context* c, amb_memory mem, const float* f,
                                           uint32_t w, uint32_t h, uint32_t d,
                                           const amb_tgv_params* p, const uint8_t* active, float* u,
                                           float* vx, float* vy, float* vz,
                                           amb_iteration_callback cb, void* user) try {
    if (!c) return fail(AMB_ERR_INVALID_ARGUMENT, "ctx is NULL");
    if (!active) return fail(AMB_ERR_INVALID_ARGUMENT, "active is NULL");
    AMB_TRY(begin(c, mem, f, w, h, d, p));
    AMB_TRY(ensure_active(c));
    AMB_HIP(hipMemcpyAsync(c->act, active, c->g.n(), kind_in(mem), c->stream));
    AMB_TRY(run_masked(c, p->iterations, p->callback_interval, cb, user));
    AMB_TRY(copy_field_out(c, mem, u, F_U));
    AMB_TRY(copy_field_out(c, mem, vx, F_VX));
    AMB_TRY(copy_field_out(c, mem, vy, F_VY));
    AMB_TRY(copy_field_out(c, mem, vz, F_VZ));
    AMB_HIP(hipStreamSynchronize(c->stream));
    return AMB_OK;
} catch (...) {
    return guard_exception();
}

AMB_API amb_status amb_tgv2_denoise(amb_context* c, amb_memory mem, const float* f, uint32_t w,
                                    uint32_t h, uint32_t d, const amb_tgv_params* p, float* u,
                                    amb_iteration_callback cb, void* user) try {
    if (!c) return fail(AMB_ERR_INVALID_ARGUMENT, "ctx is NULL");
    AMB_TRY(begin(c, mem, f, w, h, d, p));
    AMB_TRY(run(c, p->iterations, p->callback_interval, cb, user, false));
    AMB_TRY(copy_field_out(c, mem, u, F_U));
    AMB_HIP(hipStreamSynchronize(c->stream));
    return AMB_OK;
} catch (...) {
    return guard_exception();
}

AMB_API amb_status amb_divergence(amb_context* c, amb_memory mem, const float* dx, const float* dy,
                                  const float* dz, uint32_t w, uint32_t h, uint32_t d, float* out) try {
    if (!c || !out) return fail(AMB_ERR_INVALID_ARGUMENT, "NULL argument");
    AMB_TRY(geometry_only(c, w, h, d));
    const float* v[3];
    AMB_TRY(stage3(c, mem, dx, dy, dz, v));
    {
        ProfScope ps(c, P_DIV);
        AMB_HIP(launch_divergence(c->stream, v[0], v[1], v[2], c->scratch[0], c->g));
    }
    AMB_TRY(copy_out(c, mem, out, c->scratch[0]));
    AMB_HIP(hipStreamSynchronize(c->stream));
    return AMB_OK;
} catch (...) {
    return guard_exception();
}

static amb_status transform_entry(amb_context* c, amb_memory mem, const float* image, uint32_t w,
                                  uint32_t h, uint32_t d, uint32_t precision, float* out,
                                  int inverse) {
    if (!c || !image || !out) return fail(AMB_ERR_INVALID_ARGUMENT, "NULL argument");
    if (precision > 2) return fail(AMB_ERR_INVALID_ARGUMENT, "unknown dct_precision");
    AMB_TRY(geometry_only(c, w, h, d));
    AMB_HIP(hipMemcpyAsync(c->scratch[0], image, c->g.n() * 4, kind_in(mem), c->stream));
    AMB_TRY(transform3(c, c->scratch[0], c->scratch[1], c->scratch[2], inverse, int(precision)));
    AMB_TRY(copy_out(c, mem, out, c->scratch[2]));
    AMB_HIP(hipStreamSynchronize(c->stream));
    return AMB_OK;
}

AMB_API amb_status amb_cosine_transform(amb_context* c, amb_memory mem, const float* image,
                                        uint32_t w, uint32_t h, uint32_t d, uint32_t precision,
                                        float* out) try {
    return transform_entry(c, mem, image, w, h, d, precision, out, 0);
} catch (...) {
    return guard_exception();
}

AMB_API amb_status amb_inverse_cosine_transform(amb_context* c, amb_memory mem, const float* image,
                                                uint32_t w, uint32_t h, uint32_t d,
                                                uint32_t precision, float* out) try {
    return transform_entry(c, mem, image, w, h, d, precision, out, 1);
} catch (...) {
    return guard_exception();
}

AMB_API amb_status amb_solve_poisson_in_cosine_domain(amb_context* c, amb_memory mem,
                                                      const float* image, uint32_t w, uint32_t h,
                                                      uint32_t d, uint32_t flags, float* out) try {
    if (!c || !image || !out) return fail(AMB_ERR_INVALID_ARGUMENT, "NULL argument");
    AMB_TRY(geometry_only(c, w, h, d));
    AMB_HIP(hipMemcpyAsync(c->scratch[0], image, c->g.n() * 4, kind_in(mem), c->stream));
    AMB_TRY(poisson(c, c->scratch[0], c->scratch[1], flags));
    AMB_TRY(copy_out(c, mem, out, c->scratch[1]));
    AMB_HIP(hipStreamSynchronize(c->stream));
    return AMB_OK;
} catch (...) {
    return guard_exception();
}

AMB_API amb_status amb_integrate_gradients_poisson(amb_context* c, amb_memory mem, const float* vx,
                                                   const float* vy, const float* vz, uint32_t w,
                                                   uint32_t h, uint32_t d, uint32_t precision,
                                                   uint32_t flags, float* shading) try {
    if (!c || !shading) return fail(AMB_ERR_INVALID_ARGUMENT, "NULL argument");
    if (precision > 2) return fail(AMB_ERR_INVALID_ARGUMENT, "unknown dct_precision");
    AMB_TRY(geometry_only(c, w, h, d));
    const float* v[3];
    AMB_TRY(stage3(c, mem, vx, vy, vz, v));
    float* l = nullptr;
    AMB_TRY(integrate(c, v[0], v[1], v[2], int(precision), flags, &l));
    AMB_TRY(copy_out(c, mem, shading, l));
    AMB_HIP(hipStreamSynchronize(c->stream));
    return AMB_OK;
} catch (...) {
    return guard_exception();
}

AMB_API amb_status amb_deshade_epilogue(amb_context* c, amb_memory mem, const float* u,
                                        const float* shading, const float* mask, uint32_t w,
                                        uint32_t h, uint32_t d, uint32_t flags, float* shading_out,
                                        float* deshaded) try {
    if (!c || !u || !shading || !deshaded) return fail(AMB_ERR_INVALID_ARGUMENT, "NULL argument");
    AMB_TRY(geometry_only(c, w, h, d));
    const size_t n = c->g.n();
    // u -> scratch[2], l -> scratch[1] (modified in place), mask -> scratch[3], r -> scratch[4]
    AMB_HIP(hipMemcpyAsync(c->scratch[2], u, n * 4, kind_in(mem), c->stream));
    AMB_HIP(hipMemcpyAsync(c->scratch[1], shading, n * 4, kind_in(mem), c->stream));
    if (mask) AMB_HIP(hipMemcpyAsync(c->scratch[3], mask, n * 4, kind_in(mem), c->stream));
    const bool masked = mask && (flags & AMB_FLAG_MASKED_ITERATION);
    {
        ProfScope ps(c, P_EPI);
        AMB_HIP(launch_epilogue(c->stream, c->scratch[2], c->scratch[1], mask ? c->scratch[3] : nullptr,
                                c->scratch[4], n, (masked || (flags & AMB_FLAG_MASK_SHADING)) ? 1 : 0,
                                (flags & AMB_FLAG_SET_NEGATIVE_ZERO) ? 1 : 0,
                                (flags & AMB_FLAG_ADD_BACKGROUND_BACK) ? 1 : 0, masked ? 1 : 0));
    }
    AMB_TRY(copy_out(c, mem, shading_out, c->scratch[1]));
    AMB_TRY(copy_out(c, mem, deshaded, c->scratch[4]));
    AMB_HIP(hipStreamSynchronize(c->stream));
    return AMB_OK;
} catch (...) {
    return guard_exception();
}

// Poisson integration of the state's v, then the epilogue (A18, A19): shared by
// the TGV2 and TGV-k pipelines.
static amb_status deshade_tail(amb_context* c, amb_memory mem, const float* mask, uint32_t flags,
                               uint32_t precision, bool masked, float* u, float* shading,
                               float* deshaded) {
    // v as FP32 device views (BF16 storage: converted into scratch[2..4])
    const float* v[3];
    for (int k = 0; k < 3; ++k) {
        if (c->storage == 0) {
            v[k] = static_cast<const float*>(c->fields[F_VX + k]);
        } else {
            AMB_HIP(launch_convert(c->stream, 1, c->fields[F_VX + k], c->scratch[2 + k], c->g.n(), 0));
            v[k] = c->scratch[2 + k];
        }
    }
    float* l = nullptr;
    AMB_TRY(integrate(c, v[0], v[1], v[2], int(precision), flags, &l));   // l in scratch[1]
    // u as FP32 (scratch[2] for BF16), mask into scratch[3], r into scratch[4]
    const float* uf = static_cast<const float*>(c->fields[F_U]);
    if (c->storage == 1) {
        AMB_HIP(launch_convert(c->stream, 1, c->fields[F_U], c->scratch[2], c->g.n(), 0));
        uf = c->scratch[2];
    }
    const float* m = nullptr;
    if (mask) {
        if (mem == AMB_MEM_DEVICE) {
            m = mask;
        } else {
            AMB_HIP(hipMemcpyAsync(c->scratch[3], mask, c->g.n() * 4, hipMemcpyHostToDevice, c->stream));
            m = c->scratch[3];
        }
    }
    {
        ProfScope ps(c, P_EPI);
        AMB_HIP(launch_epilogue(c->stream, uf, l, m, c->scratch[4], c->g.n(),
                                (masked || (flags & AMB_FLAG_MASK_SHADING)) ? 1 : 0,
                                (flags & AMB_FLAG_SET_NEGATIVE_ZERO) ? 1 : 0,
                                (flags & AMB_FLAG_ADD_BACKGROUND_BACK) ? 1 : 0,
                                masked ? 1 : 0));
    }
    AMB_TRY(copy_out(c, mem, u, uf));
    AMB_TRY(copy_out(c, mem, shading, l));
    AMB_TRY(copy_out(c, mem, deshaded, c->scratch[4]));
    AMB_HIP(hipStreamSynchronize(c->stream));
    return AMB_OK;
}

AMB_API amb_status amb_tgv2_deshade_pipeline(amb_context* c, amb_memory mem, const float* f,
                                             uint32_t w, uint32_t h, uint32_t d,
                                             const amb_tgv_params* p, const float* mask, float* u,
                                             float* shading, float* deshaded,
                                             amb_iteration_callback cb, void* user) try {
    if (!c) return fail(AMB_ERR_INVALID_ARGUMENT, "ctx is NULL");
    AMB_TRY(begin(c, mem, f, w, h, d, p));
    const bool masked = mask && (p->flags & AMB_FLAG_MASKED_ITERATION);
    // the masked processor takes the 2-D loop for depth-1 images
    // (TGVDeshadeMaskedProcessor.cpp:93-98)
    if (masked && d == 1 && !c->sp.planar) {
        c->sp.planar = 1;
        c->iplan_ok = c->storage == 0 && fused_divisors_ok(c->sp.alpha0, c->sp.alpha1) &&
                      fused_image_plan(c->g, c->device, c->ov, &c->iplan);
    }
    if (masked) {
        // active set = mask >= 1 (TGVDeshadeMaskedProcessor.cpp:348-353); f stays unmasked
        // (:99 clones it before the mask multiply at :109)
        AMB_TRY(ensure_active(c));
        const float* md = mask;
        if (mem != AMB_MEM_DEVICE) {
            AMB_HIP(hipMemcpyAsync(c->scratch[3], mask, c->g.n() * 4, hipMemcpyHostToDevice, c->stream));
            md = c->scratch[3];
        }
        AMB_HIP(launch_active_from_mask(c->stream, md, c->act, c->g.n()));
        AMB_TRY(run_masked(c, p->iterations, p->callback_interval, cb, user));
    } else {
        AMB_TRY(run(c, p->iterations, p->callback_interval, cb, user, true));
    }
    return deshade_tail(c, mem, mask, p->flags, p->dct_precision, masked, u, shading, deshaded);
} catch (...) {
    return guard_exception();
}

AMB_API amb_status amb_tgvk_deshade(amb_context* c, amb_memory mem, const float* f, uint32_t w,
                                    uint32_t h, uint32_t d, const amb_tgv_params* p,
                                    uint32_t order, const float* alpha, float* u, float* vx,
                                    float* vy, float* vz, amb_iteration_callback cb, void* user) try {
    if (!c) return fail(AMB_ERR_INVALID_ARGUMENT, "ctx is NULL");
    AMB_TRY(tgvk_begin(c, mem, f, w, h, d, p, order, alpha));
    AMB_TRY(tgvk_run(c, p->iterations, p->callback_interval, cb, user));
    AMB_TRY(copy_field_out(c, mem, u, F_U));
    AMB_TRY(copy_field_out(c, mem, vx, F_VX));
    AMB_TRY(copy_field_out(c, mem, vy, F_VY));
    AMB_TRY(copy_field_out(c, mem, vz, F_VZ));
    AMB_HIP(hipStreamSynchronize(c->stream));
    return AMB_OK;
} catch (...) {
    return guard_exception();
}

AMB_API amb_status amb_tgvk_deshade_masked(amb_context* c, amb_memory mem, const float* f,
                                           uint32_t w, uint32_t h, uint32_t d,
                                           const amb_tgv_params* p, uint32_t order,
                                           const float* alpha, const uint8_t* active, float* u,
                                           float* vx, float* vy, float* vz,
                                           amb_iteration_callback cb, void* user) try {
    if (!c) return fail(AMB_ERR_INVALID_ARGUMENT, "ctx is NULL");
    if (!active) return fail(AMB_ERR_INVALID_ARGUMENT, "active is NULL");
    AMB_TRY(tgvk_begin(c, mem, f, w, h, d, p, order, alpha));
    AMB_TRY(ensure_active(c));
    AMB_HIP(hipMemcpyAsync(c->act, active, c->g.n(), kind_in(mem), c->stream));
    AMB_TRY(tgvk_run_masked(c, p->iterations, p->callback_interval, cb, user));
    AMB_TRY(copy_field_out(c, mem, u, F_U));
    AMB_TRY(copy_field_out(c, mem, vx, F_VX));
    AMB_TRY(copy_field_out(c, mem, vy, F_VY));
    AMB_TRY(copy_field_out(c, mem, vz, F_VZ));
    AMB_HIP(hipStreamSynchronize(c->stream));
    return AMB_OK;
} catch (...) {
    return guard_exception();
}

AMB_API amb_status amb_tgvk_deshade_pipeline(amb_context* c, amb_memory mem, const float* f,
                                             uint32_t w, uint32_t h, uint32_t d,
                                             const amb_tgv_params* p, uint32_t order,
                                             const float* alpha, const float* mask, float* u,
                                             float* shading, float* deshaded,
                                             amb_iteration_callback cb, void* user) try {
    if (!c) return fail(AMB_ERR_INVALID_ARGUMENT, "ctx is NULL");
    if (!p) return fail(AMB_ERR_INVALID_ARGUMENT, "params is NULL");
    if (!alpha) return fail(AMB_ERR_INVALID_ARGUMENT, "alpha is NULL");
    if (order < 2) return fail(AMB_ERR_INVALID_ARGUMENT, "order must be at least 2");
    // TGVKDeshadeProcessor.cpp:93-100 (and TGVKDeshadeMaskedProcessor.cpp:95-108):
    // depth 1 takes the 2-D loops; the mask multiplies l and r
    // (deshade_poisson_cosine_transform, :155-158)
    const bool masked = mask && (p->flags & AMB_FLAG_MASKED_ITERATION);
    amb_tgv_params pp = *p;
    pp.flags |= AMB_FLAG_MASK_SHADING;
    if (!masked) pp.flags &= ~AMB_FLAG_MASKED_ITERATION;
    if (d == 1) pp.flags |= AMB_FLAG_PLANAR;
    if (order == 2) {   // :132-143, the TGV2 launcher with alpha0 = alpha[1], alpha1 = alpha[0]
        pp.alpha0 = alpha[1];
        pp.alpha1 = alpha[0];
        return amb_tgv2_deshade_pipeline(c, mem, f, w, h, d, &pp, mask, u, shading, deshaded, cb,
                                         user);
    }
    AMB_TRY(tgvk_begin(c, mem, f, w, h, d, &pp, order, alpha));
    if (masked) {
        // active set = mask >= 1 (TGVDeshadeMaskedProcessor.cpp:348-353); f stays unmasked
        // (TGVKDeshadeMaskedProcessor.cpp:110 clones it before the mask multiply at :118)
        AMB_TRY(ensure_active(c));
        const float* md = mask;
        if (mem != AMB_MEM_DEVICE) {
            AMB_HIP(hipMemcpyAsync(c->scratch[3], mask, c->g.n() * 4, hipMemcpyHostToDevice, c->stream));
            md = c->scratch[3];
        }
        AMB_HIP(launch_active_from_mask(c->stream, md, c->act, c->g.n()));
        AMB_TRY(tgvk_run_masked(c, pp.iterations, pp.callback_interval, cb, user));
    } else {
        AMB_TRY(tgvk_run(c, pp.iterations, pp.callback_interval, cb, user));
    }
    return deshade_tail(c, mem, mask, pp.flags, pp.dct_precision, masked, u, shading, deshaded);
} catch (...) {
    return guard_exception();
}

// ---- resampling -------------------------------------------------------------
namespace {

// Carves `count` regions of the given byte sizes out of the context's resize
// workspace (256-B aligned), growing it when needed.
amb_status resize_workspace(amb_context* c, const size_t* bytes, int count, char** out) {
    size_t need = 0;
    for (int k = 0; k < count; ++k) need += align_up(bytes[k], 256);
    if (need > c->rcap) {
        if (c->rblock) AMB_HIP(hipFree(c->rblock));
        c->rblock = nullptr;
        c->rcap = 0;
        AMB_HIP(hipMalloc(reinterpret_cast<void**>(&c->rblock), need));
        c->rcap = need;
    }
    size_t off = 0;
    for (int k = 0; k < count; ++k) {
        out[k] = c->rblock + off;
        off += align_up(bytes[k], 256);
    }
    return AMB_OK;
}

uint32_t resize_extent(uint32_t n, float factor) {
    const float v = std::ceil(float(n) * factor);
    return v < 1.0f ? 1u : uint32_t(v);
}

}  // namespace

AMB_API void amb_resize_extents(uint32_t w, uint32_t h, uint32_t d, float factor,
                                uint32_t out[3]) {
    out[0] = resize_extent(w, factor);
    out[1] = resize_extent(h, factor);
    out[2] = resize_extent(d, factor);
}

AMB_API amb_status amb_resize_bspline3(amb_context* c, amb_memory mem, const float* in, uint32_t w,
                                       uint32_t h, uint32_t d, float* out, uint32_t ow,
                                       uint32_t oh, uint32_t od) try {
    if (!c || !in || !out) return fail(AMB_ERR_INVALID_ARGUMENT, "NULL argument");
    AMB_TRY(check_geometry(w, h, d));
    AMB_TRY(check_geometry(ow, oh, od));
    AMB_TRY(set_device(c));
    const Geometry gi{int(w), int(h), int(d)}, go{int(ow), int(oh), int(od)};
    const size_t sz[3] = {gi.n() * sizeof(double), gi.n() * 4, go.n() * 4};
    char* ws[3];
    AMB_TRY(resize_workspace(c, sz, 3, ws));
    const float* src = in;
    float* dst = out;
    if (mem != AMB_MEM_DEVICE) {
        AMB_HIP(hipMemcpyAsync(ws[1], in, gi.n() * 4, hipMemcpyHostToDevice, c->stream));
        src = reinterpret_cast<float*>(ws[1]);
        dst = reinterpret_cast<float*>(ws[2]);
    }
    {
        ProfScope ps(c, P_RESIZE);
        AMB_HIP(launch_resize_spline3(c->stream, src, gi, reinterpret_cast<double*>(ws[0]), dst, go));
    }
    if (mem != AMB_MEM_DEVICE)
        AMB_HIP(hipMemcpyAsync(out, dst, go.n() * 4, hipMemcpyDeviceToHost, c->stream));
    AMB_HIP(hipStreamSynchronize(c->stream));
    return AMB_OK;
} catch (...) {
    return guard_exception();
}

AMB_API amb_status amb_tgvk_deshade_downsampled(amb_context* c, amb_memory mem, const float* f,
                                                uint32_t w, uint32_t h, uint32_t d,
                                                const amb_tgv_params* p, float factor,
                                                uint32_t order, const float* alpha,
                                                const float* mask, float* u, float* shading,
                                                float* deshaded, amb_iteration_callback cb,
                                                void* user) try {
    if (!c || !f || !p || !alpha) return fail(AMB_ERR_INVALID_ARGUMENT, "NULL argument");
    if (!(factor > 0.f)) return fail(AMB_ERR_INVALID_ARGUMENT, "downsampling factor must be positive");
    AMB_TRY(check_geometry(w, h, d));
    AMB_TRY(set_device(c));
    const bool masked_variant = (p->flags & AMB_FLAG_MASKED_ITERATION) != 0;
    const Geometry g{int(w), int(h), int(d)};
    uint32_t se[3] = {w, h, d};
    if (factor != 1.0f) amb_resize_extents(w, h, d, factor, se);
    const Geometry gs{int(se[0]), int(se[1]), int(se[2])};
    const size_t n = g.n(), ns = gs.n();
    // full: f, mask, f*mask, u, l, r; small: f, mask, u, l, r; coefficients
    const size_t sz[12] = {n * 4, n * 4, n * 4, n * 4, n * 4, n * 4,
                           ns * 4, ns * 4, ns * 4, ns * 4, ns * 4,
                           std::max(n, ns) * sizeof(double)};
    char* ws[12];
    AMB_TRY(resize_workspace(c, sz, 12, ws));
    float* F = reinterpret_cast<float*>(ws[0]);
    float* M = mask ? reinterpret_cast<float*>(ws[1]) : nullptr;
    float* FM = reinterpret_cast<float*>(ws[2]);
    float* U = reinterpret_cast<float*>(ws[3]);
    float* L = reinterpret_cast<float*>(ws[4]);
    float* R = reinterpret_cast<float*>(ws[5]);
    float* Fs = reinterpret_cast<float*>(ws[6]);
    float* Ms = reinterpret_cast<float*>(ws[7]);
    float* Us = reinterpret_cast<float*>(ws[8]);
    float* Ls = reinterpret_cast<float*>(ws[9]);
    float* Rs = reinterpret_cast<float*>(ws[10]);
    double* coef = reinterpret_cast<double*>(ws[11]);
    const hipMemcpyKind kin = kind_in(mem);
    AMB_HIP(hipMemcpyAsync(F, f, n * 4, kin, c->stream));
    if (mask) AMB_HIP(hipMemcpyAsync(M, mask, n * 4, kin, c->stream));
    // TGVKDeshadeDownsampledProcessor.cpp:46-53: the masked input goes down
    const float* src = F;
    if (masked_variant && M) {
        AMB_HIP(launch_multiply(c->stream, F, M, FM, n));
        src = FM;
    }
    const bool resample = factor != 1.0f;
    const float* fs = src;
    const float* ms = M;
    if (resample) {
        ProfScope ps(c, P_RESIZE);
        AMB_HIP(launch_resize_spline3(c->stream, src, g, coef, Fs, gs));
        fs = Fs;
        if (M) {
            AMB_HIP(launch_resize_spline3(c->stream, M, g, coef, Ms, gs));
            ms = Ms;
        }
    }
    // the (masked) TGV-k processor on the resampled volume, device to device
    AMB_TRY(amb_tgvk_deshade_pipeline(c, AMB_MEM_DEVICE, fs, gs.w, gs.h, gs.d, p, order, alpha, ms,
                                      resample ? Us : U, resample ? Ls : L, Rs, cb, user));
    if (resample) {
        ProfScope ps(c, P_RESIZE);
        AMB_HIP(launch_resize_spline3(c->stream, Us, gs, coef, U, g));
        AMB_HIP(launch_resize_spline3(c->stream, Ls, gs, coef, L, g));
    }
    {
        ProfScope ps(c, P_EPI);
        AMB_HIP(launch_downsampled_tail(c->stream, F, M, L, R, n, masked_variant ? 1 : 0,
                                        (p->flags & AMB_FLAG_ADD_BACKGROUND_BACK) ? 1 : 0,
                                        (p->flags & AMB_FLAG_SET_NEGATIVE_ZERO) ? 1 : 0));
    }
    const hipMemcpyKind kout = kind_out(mem);
    if (u) AMB_HIP(hipMemcpyAsync(u, U, n * 4, kout, c->stream));
    if (shading) AMB_HIP(hipMemcpyAsync(shading, L, n * 4, kout, c->stream));
    if (deshaded) AMB_HIP(hipMemcpyAsync(deshaded, R, n * 4, kout, c->stream));
    AMB_HIP(hipStreamSynchronize(c->stream));
    return AMB_OK;
} catch (...) {
    return guard_exception();
}

AMB_API amb_status amb_tgv2_energy(amb_context* c, double out[10]) try {
    if (!c || !out) return fail(AMB_ERR_INVALID_ARGUMENT, "NULL argument");
    if (!c->have_state) return fail(AMB_ERR_INVALID_ARGUMENT, "no iteration state");
    AMB_TRY(set_device(c));
    AMB_TRY(energy_pass(c, c->energy_out));
    AMB_HIP(hipMemcpyAsync(out, c->energy_out, 10 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    AMB_HIP(hipStreamSynchronize(c->stream));
    return AMB_OK;
} catch (...) {
    return guard_exception();
}

AMB_API amb_status amb_tgv2_energy_noalloc(const float* f, const void* const* fields,
                                           uint32_t w, uint32_t h, uint32_t d,
                                           const amb_tgv_params* p, void* stream,
                                           double* out_dev) try {
    if (!f || !fields || !p || !out_dev) return fail(AMB_ERR_INVALID_ARGUMENT, "NULL argument");
    for (int k = 0; k < N_FIELDS; ++k)
        if (!fields[k]) return fail(AMB_ERR_INVALID_ARGUMENT, "NULL field pointer");
    AMB_TRY(check_geometry(w, h, d));
    AMB_TRY(check_params(p));
    if ((p->flags & AMB_FLAG_PLANAR) && d != 1)
        return fail(AMB_ERR_INVALID_ARGUMENT, "AMB_FLAG_PLANAR needs depth 1");
    const Geometry g{int(w), int(h), int(d)};
    if (energy_partials_size(g) > NOALLOC_SCRATCH_DOUBLES)
        return fail(AMB_ERR_NOT_SUPPORTED, "energy partials exceed the 4096-row scratch array");
    int device = 0;
    AMB_HIP(hipGetDevice(&device));
    double* part = nullptr;
    AMB_TRY(scratch_ptr(device, &part));
    AMB_HIP(launch_energy(static_cast<hipStream_t>(stream), int(p->storage), fields, f, g,
                          p->lambda, p->alpha0, p->alpha1, int(p->data_term),
                          (p->flags & AMB_FLAG_PLANAR) ? 1 : 0, part, out_dev));
    return AMB_OK;
} catch (...) {
    return guard_exception();
}

AMB_API amb_status amb_tgv2_set_energy_interval(amb_context* c, uint32_t interval) try {
    if (!c) return fail(AMB_ERR_INVALID_ARGUMENT, "NULL context");
    c->energy_interval = interval;
    c->etrace.clear();
    c->etrace_it.clear();
    return AMB_OK;
} catch (...) {
    return guard_exception();
}

AMB_API amb_status amb_tgv2_energy_trace(amb_context* c, double* out, uint32_t max_rows,
                                         uint32_t* rows) try {
    if (!c || !rows) return fail(AMB_ERR_INVALID_ARGUMENT, "NULL argument");
    const size_t n = c->etrace.size() / 11;
    *rows = uint32_t(n);
    if (out) std::memcpy(out, c->etrace.data(), std::min<size_t>(n, max_rows) * 11 * sizeof(double));
    return AMB_OK;
} catch (...) {
    return guard_exception();
}

AMB_API amb_status amb_tgv2_denoise_adjoint(amb_context* c, amb_memory mem, const float* f,
                                            uint32_t w, uint32_t h, uint32_t d,
                                            const amb_tgv_params* p, const float* dL_du,
                                            float* u_out, float* dL_df, double dL_dparams[3],
                                            uint32_t checkpoint_every) try {
    if (!c || !dL_du || !dL_df || !dL_dparams) return fail(AMB_ERR_INVALID_ARGUMENT, "NULL argument");
    AMB_TRY(check_params(p));
    if (p->data_term != AMB_DATA_L2)
        return fail(AMB_ERR_NOT_SUPPORTED, "the adjoint is implemented for the L2 data term");
    if (p->storage != AMB_STORAGE_F32)
        return fail(AMB_ERR_NOT_SUPPORTED, "the adjoint runs on FP32 fields");
    if (p->flags & AMB_FLAG_PLANAR)
        return fail(AMB_ERR_NOT_SUPPORTED, "the adjoint is implemented for the 3-D loop");
    const uint32_t K = checkpoint_every ? checkpoint_every : 20;
    const uint32_t N = p->iterations;
    AMB_TRY(begin(c, mem, f, w, h, d, p));
    const Geometry g = c->g;
    const size_t n = g.n();
    const size_t fstride = align_up(n * 4, 256);
    const size_t state_bytes = fstride * N_FIELDS;
    const uint32_t nseg = N ? (N + K - 1) / K : 0;
    // the reverse pass rebuilds each segment with the fused sweep's form that
    // also stores u_bar and v_bar, out of place, one launch per iteration (else
    // in place with the two sweeps after a copy of the state).  With a checkpoint
    // at every iteration (K = 1) that form keeps every state instead: the
    // forward runs as those launches into N + 1 slots and nothing is recomputed.
    const bool fb = c->fused_ok && fused_bars_supported(c->storage, g, c->sp, c->fplan);
    const bool resident = fb && K == 1 && N > 0;
    const size_t n_ckpt = resident ? 1 : nseg, n_slots = resident ? size_t(N) + 1 : size_t(K) + 1;
    const size_t part_bytes = align_up(adjoint_slots_size(g) * sizeof(double), 256);
    const size_t need = state_bytes * (n_ckpt + n_slots + 1) + fstride * 8 + part_bytes + 256;
    if (need > c->adj_cap) {
        if (c->adj_block) AMB_HIP(hipFree(c->adj_block));
        c->adj_block = nullptr;
        c->adj_cap = 0;
        AMB_HIP(hipMalloc(reinterpret_cast<void**>(&c->adj_block), need));
        c->adj_cap = need;
    }
    char* base = c->adj_block;
    auto state_at = [&](char* blk, void** tab) {
        for (int k = 0; k < N_FIELDS; ++k) tab[k] = blk + size_t(k) * fstride;
    };
    char* ckpt = base;
    char* slots = ckpt + state_bytes * n_ckpt;
    char* adjs = slots + state_bytes * n_slots;
    float* ast = reinterpret_cast<float*>(adjs + state_bytes);
    float* wst = reinterpret_cast<float*>(adjs + state_bytes + fstride);   // 3 fields
    float* gradf = reinterpret_cast<float*>(adjs + state_bytes + fstride * 4);
    float* pbuf = reinterpret_cast<float*>(adjs + state_bytes + fstride * 5);   // 3 fields
    double* partials = reinterpret_cast<double*>(adjs + state_bytes + fstride * 8);
    double* acc = reinterpret_cast<double*>(adjs + state_bytes + fstride * 8 + part_bytes);
    ProfScope ps(c, P_ADJOINT);
    // forward with a checkpoint of x_k at every k % K == 0: each segment runs the
    // iteration loop (fused sweeps inside, ending with the two-sweep state); the
    // fused sweep ping-pongs fields between two blocks, so the checkpoint copies
    // field by field
    const uint32_t every = c->energy_interval;
    c->energy_interval = 0;
    struct RestoreInterval {
        amb_context* c;
        uint32_t v;
        ~RestoreInterval() { c->energy_interval = v; }
    } restore{c, every};
    void* tabs[2][N_FIELDS];
    // slots[j] = u, u_bar, v, v_bar of x_{k0 + j} with p, q of the dual half of
    // iteration k0 + j + 1: slot 0 from x_{k0} (at cp) by one dual sweep, slot j
    // from slot j - 1 by one fused launch (primal half k0 + j, dual half k0 + j + 1)
    auto fill_slots = [&](const char* cp, uint32_t L) -> amb_status {
        AMB_HIP(hipMemcpyAsync(slots, cp, state_bytes, hipMemcpyDeviceToDevice, c->stream));
        state_at(slots, tabs[0]);
        AMB_HIP(launch_dual_sweep(c->stream, 0, tabs[0], g, c->sp));
        for (uint32_t j = 1; j <= L; ++j) {
            state_at(slots + state_bytes * size_t(j - 1), tabs[0]);
            state_at(slots + state_bytes * size_t(j), tabs[1]);
            AMB_HIP(launch_fused_sweep_bars(c->stream, tabs[0], tabs[1], c->f, g, c->sp, c->fplan));
        }
        return AMB_OK;
    };
    if (resident) {
        for (int k = 0; k < N_FIELDS; ++k)
            AMB_HIP(hipMemcpyAsync(ckpt + size_t(k) * fstride, c->fields[k], n * 4,
                                   hipMemcpyDeviceToDevice, c->stream));
        AMB_TRY(fill_slots(ckpt, N));
        if (u_out)
            AMB_HIP(hipMemcpyAsync(u_out, slots + state_bytes * size_t(N) + size_t(F_U) * fstride,
                                   n * 4, kind_out(mem), c->stream));
    } else {
        for (uint32_t seg = 0; seg < nseg; ++seg) {
            char* cp = ckpt + state_bytes * seg;
            for (int k = 0; k < N_FIELDS; ++k)
                AMB_HIP(hipMemcpyAsync(cp + size_t(k) * fstride, c->fields[k], n * 4,
                                       hipMemcpyDeviceToDevice, c->stream));
            const uint32_t L = std::min(K, N - seg * K);
            if (fb && seg + 1 == nseg) {
                // the last segment straight into the slots: the reverse pass starts
                // with it, so it needs no recompute
                AMB_TRY(fill_slots(cp, L));
                if (u_out)
                    AMB_HIP(hipMemcpyAsync(u_out, slots + state_bytes * size_t(L) +
                                                      size_t(F_U) * fstride,
                                           n * 4, kind_out(mem), c->stream));
            } else {
                AMB_TRY(run(c, L, 0, nullptr, nullptr, false));
            }
        }
        if (u_out && (!fb || nseg == 0)) AMB_TRY(copy_field_out(c, mem, u_out, F_U));
    }
    // reverse: adjoint of x_N = (dL/du, 0, ...)
    void* adj[N_FIELDS];
    state_at(adjs, adj);
    AMB_HIP(launch_zero_fill(c->stream, adjs, state_bytes));
    AMB_HIP(hipMemcpyAsync(adj[F_U], dL_du, n * 4, kind_in(mem), c->stream));
    AMB_HIP(launch_zero_fill(c->stream, gradf, n * 4));
    AMB_HIP(launch_zero_fill(c->stream, partials, adjoint_slots_size(g) * sizeof(double)));
    const AdjointArgs aa{adj, pbuf, ast, wst, gradf, c->f, partials, g, c->sp.tau, c->sp.sigma,
                         c->sp.theta, c->sp.lambda, c->sp.alpha0, c->sp.alpha1,
                         adjoint_vx(g, c->ov.adjoint_vx)};
    // iteration k in reverse: phase B with x_k, then phase C of k with phase A
    // of k - 1 (u of x_k); phase A alone for the last iteration, C alone for k = 0
    bool first = true;
    auto reverse_step = [&](uint32_t k, void* const* xk, void* const* xk1) -> amb_status {
        if (first) {
            AMB_HIP(launch_adjoint_a(c->stream, aa, static_cast<const float*>(xk1[F_U])));
            first = false;
        }
        AMB_HIP(launch_adjoint_b(c->stream, aa, xk));
        if (k > 0) AMB_HIP(launch_adjoint_ca(c->stream, aa, static_cast<const float*>(xk[F_U])));
        else AMB_HIP(launch_adjoint_c(c->stream, aa));
        return AMB_OK;
    };
    // x_{k0 + j} from the filled slots: u_bar, v_bar (and u, v) from slot j, p and
    // q from slot j - 1 (the checkpoint at cp for j = 0); u of x_{k0 + j + 1}
    // from slot j + 1
    auto reverse_slots = [&](char* cp, uint32_t k0, uint32_t L) -> amb_status {
        for (int j = int(L) - 1; j >= 0; --j) {
            state_at(slots + state_bytes * size_t(j), tabs[0]);
            char* pq = j > 0 ? slots + state_bytes * size_t(j - 1) : cp;
            for (int k = F_PX; k <= F_QYZ; ++k) tabs[0][k] = pq + size_t(k) * fstride;
            state_at(slots + state_bytes * size_t(j + 1), tabs[1]);
            AMB_TRY(reverse_step(k0 + uint32_t(j), tabs[0], tabs[1]));
        }
        return AMB_OK;
    };
    if (resident) AMB_TRY(reverse_slots(ckpt, 0, N));
    for (int seg = resident ? -1 : int(nseg) - 1; seg >= 0; --seg) {
        const uint32_t k0 = uint32_t(seg) * K, k1 = std::min(k0 + K, N), L = k1 - k0;
        char* cp = ckpt + state_bytes * size_t(seg);
        if (fb) {
            if (seg + 1 < int(nseg)) AMB_TRY(fill_slots(cp, L));   // the last one is filled
            AMB_TRY(reverse_slots(cp, k0, L));
            continue;
        }
        // slots[0] = x_{k0}; slots[j] = x_{k0 + j} recomputed with the same sweeps
        AMB_HIP(hipMemcpyAsync(slots, cp, state_bytes, hipMemcpyDeviceToDevice, c->stream));
        for (uint32_t j = 1; j <= L; ++j) {
            char* dst = slots + state_bytes * j;
            AMB_HIP(hipMemcpyAsync(dst, dst - state_bytes, state_bytes, hipMemcpyDeviceToDevice,
                                   c->stream));
            state_at(dst, tabs[0]);
            AMB_HIP(launch_dual_sweep(c->stream, 0, tabs[0], g, c->sp));
            AMB_HIP(launch_primal_sweep(c->stream, 0, tabs[0], c->f, g, c->sp));
        }
        for (int j = int(L) - 1; j >= 0; --j) {
            state_at(slots + state_bytes * size_t(j), tabs[0]);
            state_at(slots + state_bytes * size_t(j + 1), tabs[1]);
            AMB_TRY(reverse_step(k0 + uint32_t(j), tabs[0], tabs[1]));
        }
    }
    AMB_HIP(launch_adjoint_reduce(c->stream, partials, g, acc));
    AMB_HIP(launch_adjoint_finish(c->stream, adj, gradf, g));
    AMB_HIP(hipMemcpyAsync(dL_df, gradf, n * 4, kind_out(mem), c->stream));
    AMB_HIP(hipMemcpyAsync(dL_dparams, acc, 3 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    AMB_HIP(hipStreamSynchronize(c->stream));
    return AMB_OK;
} catch (...) {
    return guard_exception();
}

AMB_API amb_status amb_profile_enable(amb_context* c, int enable) try {
    if (!c) return fail(AMB_ERR_INVALID_ARGUMENT, "ctx is NULL");
    c->prof_on = enable != 0;
    c->prof_every = enable > 1 ? uint32_t(enable) : 1u;
    return AMB_OK;
} catch (...) {
    return guard_exception();
}

AMB_API amb_status amb_profile_read(amb_context* c, int cls, double* ms, uint64_t* n) try {
    if (!c || cls < 0 || cls >= N_PROF) return fail(AMB_ERR_INVALID_ARGUMENT, "bad argument");
    AMB_TRY(set_device(c));
    AMB_TRY(drain_profile(c));
    if (ms) *ms = c->prof_ms[cls];
    if (n) *n = c->prof_n[cls];
    return AMB_OK;
} catch (...) {
    return guard_exception();
}

AMB_API amb_status amb_profile_count(amb_context* c, int cls, uint64_t* launches) try {
    if (!c || !launches || cls < 0 || cls >= N_PROF)
        return fail(AMB_ERR_INVALID_ARGUMENT, "bad argument");
    *launches = c->prof_seen[cls];
    return AMB_OK;
} catch (...) {
    return guard_exception();
}

AMB_API amb_status amb_profile_fused_tiles(amb_context* c, uint64_t* swept, uint64_t* total) try {
    if (!c || !swept || !total) return fail(AMB_ERR_INVALID_ARGUMENT, "bad argument");
    *swept = c->prof_tiles_swept;
    *total = c->prof_tiles_total;
    return AMB_OK;
} catch (...) {
    return guard_exception();
}

AMB_API amb_status amb_profile_reset(amb_context* c) try {
    if (!c) return fail(AMB_ERR_INVALID_ARGUMENT, "ctx is NULL");
    AMB_TRY(set_device(c));
    AMB_TRY(drain_profile(c));
    for (int k = 0; k < N_PROF; ++k) { c->prof_ms[k] = 0.0; c->prof_n[k] = 0; c->prof_seen[k] = 0; }
    c->prof_tiles_swept = 0;
    c->prof_tiles_total = 0;
    return AMB_OK;
} catch (...) {
    return guard_exception();
}

}  // extern "C"
