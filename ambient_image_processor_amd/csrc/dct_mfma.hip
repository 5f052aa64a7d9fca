// Separable DCT-II / DCT-III passes as dense per-axis matrix products on the
// gfx950 matrix cores.
//
// Reference: cuda_image_operations/cosine_transform.cu:142-223 and
// inverse_cosine_transform.cu:79-164 evaluate every output with an O(n) loop
// of FP64 cospi calls; here each axis pass is out = C * in along that axis,
// with C the n x n cosine matrix (host FP64, rounded once):
//   forward  C[k][m] = cos(pi (m + 1/2) k / n)
//   inverse  C[k][m] = w_m cos(pi (k + 1/2) m / n), w_0 = 1, w_m = 2 (the
//            reference's "X[0] + 2 sum_{m>=1}" folded into the matrix)
// The last (z) pass applies the reference's *8 (forward) or 1/(W*H*D*8)
// (inverse), and can fuse the Poisson eigenvalue divide of
// solve_poisson_in_cosine_domain.cu:42-53 into its epilogue.
//
// Precisions:
//   F32     v_mfma_f32_16x16x4_f32, FP32 operands, FP32 accumulate
//   BF16X3  v_mfma_f32_16x16x32_bf16 on hi/lo bf16 splits of both operands:
//           C_hi*X_hi + C_hi*X_lo + C_lo*X_hi, FP32 accumulate (the lo*lo
//           term is below the split's own rounding)
//
// Two kernels: the column pass (y, z: the summed index is strided, output
// columns are contiguous) keeps an n x 32 output tile per block and stages
// 32 x 32 input tiles in LDS; the row pass (x: the summed index is
// contiguous) keeps a 32-row x n tile.  4 waves per block split the n
// outputs; operand k-slots are assigned identically for A and B, so the
// products are independent of the hardware's k ordering within a slot group.
#include "amb_internal.hpp"

namespace amb {
namespace {

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef short i16x8 __attribute__((ext_vector_type(8)));

constexpr int KC = 32;     // summed-index chunk
constexpr int NS = 32;     // output columns (column pass) / rows (row pass) per block
constexpr int NT = 256;    // threads per block (4 waves)

__device__ __forceinline__ uint16_t bf16_rne(float v) {
    uint32_t u = __float_as_uint(v);
    u += 0x7fffu + ((u >> 16) & 1u);
    return uint16_t(u >> 16);
}
__device__ __forceinline__ float bf16_f(uint16_t h) { return __uint_as_float(uint32_t(h) << 16); }

// hi/lo bf16 split of 8 floats into MFMA operands
__device__ __forceinline__ void split8(const float (&x)[8], bf16x8& hi, bf16x8& lo) {
    i16x8 h, l;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        const uint16_t a = bf16_rne(x[e]);
        const uint16_t b = bf16_rne(x[e] - bf16_f(a));
        h[e] = short(a);
        l[e] = short(b);
    }
    hi = __builtin_bit_cast(bf16x8, h);
    lo = __builtin_bit_cast(bf16x8, l);
}

// 8 consecutive bf16 matrix entries (zero beyond n_valid); one 16-byte load
// when all 8 are valid and aligned (extent a multiple of 8).
__device__ __forceinline__ bf16x8 load_bf16x8(const uint16_t* p, int n_valid, bool aligned) {
    i16x8 v;
    if (n_valid == 8 && aligned) {
        v = *reinterpret_cast<const i16x8*>(p);
    } else {
#pragma unroll
        for (int e = 0; e < 8; ++e) v[e] = e < n_valid ? short(p[e]) : short(0);
    }
    return __builtin_bit_cast(bf16x8, v);
}

struct PassArgs {
    const float* in;
    float* out;
    const float* ct;        // FP32 matrix, [m][k] layout (transposed)
    const uint16_t* chi;    // BF16 hi split, [k][m] layout
    const uint16_t* clo;    // BF16 lo split, [k][m] layout
    int n;                  // extent of the transformed axis
    int S;                  // column pass: stride of the axis (= inner extent); row pass: unused
    int O;                  // column pass: outer count; row pass: number of rows
    float scale;            // multiplies the result (last pass)
    // fused Poisson divide (forward z pass)
    int poisson;
    int quirk;
    Geometry g;
    const float* cx;
    const float* cy;
    const float* cz;
};

// Poisson eigenvalue divide on an output element at linear index idx
// (the same arithmetic as poisson_kernel in dct.hip).
__device__ __forceinline__ float poisson_apply(float v, size_t idx, const PassArgs& a) {
    const unsigned wh = unsigned(a.g.w) * unsigned(a.g.h);
    int x, y, z;
    if (!a.quirk) {
        x = int(idx % a.g.w); y = int((idx / a.g.w) % a.g.h); z = int(idx / wh);
    } else {
        const float zf = floorf(float(unsigned(idx) / wh));
        int rest = int(float(int(idx)) - zf * float(wh));
        const float yf = floorf(float(unsigned(rest) / unsigned(a.g.w)));
        rest = int(float(rest) - yf * float(unsigned(a.g.w)));
        x = rest; y = int(yf); z = int(zf);
    }
    if (x == 0 && y == 0 && z == 0) return 0.0f;
    const double den = 6.0 - 2.0 * double(a.cx[x]) - 2.0 * double(a.cy[y]) - 2.0 * double(a.cz[z]);
    return float(double(v) / den);
}

// ---------------------------------------------------------------------------
// Column pass: out[o][k][s] = sum_m C[k][m] in[o][m][s]
// MT = 16-row output tiles per wave (n padded to 64 * MT).
// ---------------------------------------------------------------------------
template <int PREC, int MT>
__global__ __launch_bounds__(NT) void dct_col_kernel(PassArgs a) {
    __shared__ float sB[KC][NS + 4];
    const int n = a.n, S = a.S;
    const int nst = (S + NS - 1) / NS;
    const int o = blockIdx.x / nst;
    const int s0 = (blockIdx.x % nst) * NS;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int li = lane & 15, lg = lane >> 4;
    const int kbase = wave * (MT * 16);
    const size_t obase = size_t(o) * size_t(n) * size_t(S);

    f32x4 acc[MT][2];
#pragma unroll
    for (int t = 0; t < MT; ++t)
#pragma unroll
        for (int u = 0; u < 2; ++u) acc[t][u] = f32x4{0.f, 0.f, 0.f, 0.f};

    for (int m0 = 0; m0 < n; m0 += KC) {
        // stage in[o][m0 .. m0+KC)[s0 .. s0+NS) (zero-padded)
        for (int e = threadIdx.x; e < KC * NS; e += NT) {
            const int r = e / NS, cidx = e % NS;
            const int m = m0 + r, s = s0 + cidx;
            sB[r][cidx] = (m < n && s < S) ? a.in[obase + size_t(m) * S + s] : 0.0f;
        }
        __syncthreads();
        if (PREC == 1) {
#pragma unroll
            for (int kk = 0; kk < KC / 4; ++kk) {
                const int m = m0 + 4 * kk + lg;
                float bfrag[2];
#pragma unroll
                for (int u = 0; u < 2; ++u) bfrag[u] = sB[4 * kk + lg][u * 16 + li];
#pragma unroll
                for (int t = 0; t < MT; ++t) {
                    const int k = kbase + t * 16 + li;
                    const float afrag = (m < n && k < n) ? a.ct[size_t(m) * n + k] : 0.0f;
#pragma unroll
                    for (int u = 0; u < 2; ++u)
                        acc[t][u] = __builtin_amdgcn_mfma_f32_16x16x4f32(afrag, bfrag[u], acc[t][u], 0, 0, 0);
                }
            }
        } else {
            bf16x8 bhi[2], blo[2];
#pragma unroll
            for (int u = 0; u < 2; ++u) {
                float x[8];
#pragma unroll
                for (int e = 0; e < 8; ++e) x[e] = sB[8 * lg + e][u * 16 + li];
                split8(x, bhi[u], blo[u]);
            }
            const int mb = m0 + 8 * lg;
            const int nv = n - mb < 8 ? (n - mb > 0 ? n - mb : 0) : 8;
#pragma unroll
            for (int t = 0; t < MT; ++t) {
                const int k = kbase + t * 16 + li;
                const int kv = k < n ? nv : 0;
                const bool al = (n & 7) == 0;
                const bf16x8 ahi = load_bf16x8(a.chi + size_t(k < n ? k : 0) * n + (kv ? mb : 0), kv, al);
                const bf16x8 alo = load_bf16x8(a.clo + size_t(k < n ? k : 0) * n + (kv ? mb : 0), kv, al);
#pragma unroll
                for (int u = 0; u < 2; ++u) {
                    acc[t][u] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ahi, bhi[u], acc[t][u], 0, 0, 0);
                    acc[t][u] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ahi, blo[u], acc[t][u], 0, 0, 0);
                    acc[t][u] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(alo, bhi[u], acc[t][u], 0, 0, 0);
                }
            }
        }
        __syncthreads();
    }
    // epilogue: lane holds rows 4*lg + r of each 16x16 tile, column li
#pragma unroll
    for (int t = 0; t < MT; ++t)
#pragma unroll
        for (int u = 0; u < 2; ++u)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int k = kbase + t * 16 + 4 * lg + r;
                const int s = s0 + u * 16 + li;
                if (k < n && s < S) {
                    const size_t idx = obase + size_t(k) * S + s;
                    float v = acc[t][u][r] * a.scale;
                    if (a.poisson) v = poisson_apply(v, idx, a);
                    a.out[idx] = v;
                }
            }
}

// ---------------------------------------------------------------------------
// Row pass: out[r][k] = sum_m in[r][m] C[k][m]   (rows contiguous, x axis)
// ---------------------------------------------------------------------------
template <int PREC, int MT>
__global__ __launch_bounds__(NT) void dct_row_kernel(PassArgs a) {
    __shared__ float sA[NS][KC + 1];
    const int n = a.n, R = a.O;
    const int r0 = blockIdx.x * NS;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int li = lane & 15, lg = lane >> 4;
    const int kbase = wave * (MT * 16);

    f32x4 acc[2][MT];
#pragma unroll
    for (int u = 0; u < 2; ++u)
#pragma unroll
        for (int t = 0; t < MT; ++t) acc[u][t] = f32x4{0.f, 0.f, 0.f, 0.f};

    for (int m0 = 0; m0 < n; m0 += KC) {
        for (int e = threadIdx.x; e < NS * KC; e += NT) {
            const int rr = e / KC, cidx = e % KC;
            const int r = r0 + rr, m = m0 + cidx;
            sA[rr][cidx] = (r < R && m < n) ? a.in[size_t(r) * n + m] : 0.0f;
        }
        __syncthreads();
        if (PREC == 1) {
#pragma unroll
            for (int kk = 0; kk < KC / 4; ++kk) {
                const int m = m0 + 4 * kk + lg;
                float afrag[2];
#pragma unroll
                for (int u = 0; u < 2; ++u) afrag[u] = sA[u * 16 + li][4 * kk + lg];
#pragma unroll
                for (int t = 0; t < MT; ++t) {
                    const int k = kbase + t * 16 + li;
                    const float bfrag = (m < n && k < n) ? a.ct[size_t(m) * n + k] : 0.0f;
#pragma unroll
                    for (int u = 0; u < 2; ++u)
                        acc[u][t] = __builtin_amdgcn_mfma_f32_16x16x4f32(afrag[u], bfrag, acc[u][t], 0, 0, 0);
                }
            }
        } else {
            bf16x8 ahi[2], alo[2];
#pragma unroll
            for (int u = 0; u < 2; ++u) {
                float x[8];
#pragma unroll
                for (int e = 0; e < 8; ++e) x[e] = sA[u * 16 + li][8 * lg + e];
                split8(x, ahi[u], alo[u]);
            }
            const int mb = m0 + 8 * lg;
            const int nv = n - mb < 8 ? (n - mb > 0 ? n - mb : 0) : 8;
#pragma unroll
            for (int t = 0; t < MT; ++t) {
                const int k = kbase + t * 16 + li;
                const int kv = k < n ? nv : 0;
                const bool al = (n & 7) == 0;
                const bf16x8 bhi = load_bf16x8(a.chi + size_t(k < n ? k : 0) * n + (kv ? mb : 0), kv, al);
                const bf16x8 blo = load_bf16x8(a.clo + size_t(k < n ? k : 0) * n + (kv ? mb : 0), kv, al);
#pragma unroll
                for (int u = 0; u < 2; ++u) {
                    acc[u][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ahi[u], bhi, acc[u][t], 0, 0, 0);
                    acc[u][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(alo[u], bhi, acc[u][t], 0, 0, 0);
                    acc[u][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ahi[u], blo, acc[u][t], 0, 0, 0);
                }
            }
        }
        __syncthreads();
    }
    // lane holds rows 4*lg + rr (data rows) of each tile, column li (output k)
#pragma unroll
    for (int u = 0; u < 2; ++u)
#pragma unroll
        for (int t = 0; t < MT; ++t)
#pragma unroll
            for (int rr = 0; rr < 4; ++rr) {
                const int r = r0 + u * 16 + 4 * lg + rr;
                const int k = kbase + t * 16 + li;
                if (r < R && k < n) a.out[size_t(r) * n + k] = acc[u][t][rr] * a.scale;
            }
}

template <int PREC, int MT>
hipError_t launch_pass(hipStream_t s, const PassArgs& a, int axis) {
    if (axis == 0) {
        const int blocks = (a.O + NS - 1) / NS;
        hipLaunchKernelGGL((dct_row_kernel<PREC, MT>), dim3(blocks), dim3(NT), 0, s, a);
    } else {
        const int blocks = a.O * ((a.S + NS - 1) / NS);
        hipLaunchKernelGGL((dct_col_kernel<PREC, MT>), dim3(blocks), dim3(NT), 0, s, a);
    }
    return hipGetLastError();
}

template <int PREC>
hipError_t launch_prec(hipStream_t s, const PassArgs& a, int axis) {
    const int mt = (a.n + 63) / 64;   // 16-row tiles per wave
    if (mt <= 1) return launch_pass<PREC, 1>(s, a, axis);
    if (mt <= 2) return launch_pass<PREC, 2>(s, a, axis);
    if (mt <= 4) return launch_pass<PREC, 4>(s, a, axis);
    if (mt <= 8) return launch_pass<PREC, 8>(s, a, axis);
    return hipErrorInvalidValue;   // extents above 512 use the reference-order passes
}

}  // namespace

int dct_mfma_max_extent() { return 512; }

hipError_t launch_dct_mfma_pass(hipStream_t s, const float* in, float* out, const float* table,
                                const uint16_t* table_hi, const uint16_t* table_lo, Geometry g,
                                int axis, int inverse, int last, int precision,
                                const float* eig_x, const float* eig_y, const float* eig_z,
                                int fuse_poisson, int quirk) {
    PassArgs a;
    a.in = in; a.out = out; a.ct = table; a.chi = table_hi; a.clo = table_lo;
    a.g = g; a.cx = eig_x; a.cy = eig_y; a.cz = eig_z;
    a.quirk = quirk;
    a.poisson = (last && !inverse && fuse_poisson) ? 1 : 0;
    if (axis == 0) {
        a.n = g.w; a.S = 1; a.O = g.h * g.d;
    } else if (axis == 1) {
        a.n = g.h; a.S = g.w; a.O = g.d;
    } else {
        a.n = g.d; a.S = g.w * g.h; a.O = 1;
    }
    a.scale = 1.0f;
    if (last) a.scale = inverse ? float(1.0 / (double(g.n()) * 8.0)) : 8.0f;
    return precision == 2 ? launch_prec<2>(s, a, axis) : launch_prec<1>(s, a, axis);
}

}  // namespace amb
