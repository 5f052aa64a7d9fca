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
// One kernel for every axis: a tiled product D[i][j] = sum_m X[i][m] Y[m][j]
// with 128 x 128 output tiles per block (2 x 2 waves of 64 x 64), 32-deep
// K chunks staged in LDS with K contiguous for both operands (one 16-byte LDS
// read per lane and split for the BF16 operands).
//   column pass (y, z): X = C[k][m] (table), Y = in[o][m][s], D = out[o][k][s]
//   row pass (x):       X = in[r][m],       Y[m][k] = C[k][m], D = out[r][k]
// The data tile is read once per K chunk with 16-byte loads; tiles that share
// a data panel are adjacent and run on one XCD (xcd_contiguous), sharing its L2.  Operand k-slots are
// assigned identically for X and Y, so the products do not depend on the
// hardware's k order within a slot group.
#include "amb_internal.hpp"

namespace amb {
namespace {

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef short i16x8 __attribute__((ext_vector_type(8)));


// Poisson eigenvalue divide on the output element (x, y, z) of the forward z
// pass, i = z and j = y W + x (the same arithmetic as poisson_kernel in
// dct.hip, solve_poisson_in_cosine_domain.cu:42-53; with the index quirk the
// coordinates are recovered from the linear index through float arithmetic as
// the reference does).
__device__ __forceinline__ float poisson_apply(float v, int i, int j, const Geometry& g, int quirk,
                                               const float* cx, const float* cy,
                                               const float* cz) {
    int x, y, z;
    if (!quirk) {
        z = i; y = j / g.w; x = j - y * g.w;
    } else {
        const unsigned wh = unsigned(g.w) * unsigned(g.h);
        const unsigned idx = unsigned(i) * wh + unsigned(j);
        const float zf = floorf(float(unsigned(idx) / wh));
        int rest = int(float(int(idx)) - zf * float(wh));
        const float yf = floorf(float(unsigned(rest) / unsigned(g.w)));
        rest = int(float(rest) - yf * float(unsigned(g.w)));
        x = rest; y = int(yf); z = int(zf);
    }
    if (x == 0 && y == 0 && z == 0) return 0.0f;
    const double den =
        g.d == 1 ? 4.0 - 2.0 * double(cx[x]) - 2.0 * double(cy[y])   // the 2-D kernel
                 : 6.0 - 2.0 * double(cx[x]) - 2.0 * double(cy[y]) - 2.0 * double(cz[z]);
    return float(double(v) / den);
}

constexpr int BM = 128, BN = 128, BK = 32;
constexpr int PF32 = BK + 4;   // LDS row pitch (floats) of an FP32 tile
constexpr int PB16 = BK + 8;   // LDS row pitch (bf16) of a split tile

struct GemmArgs {
    const float* data;      // FP32 volume (X in the row pass, Y in the column pass)
    const float* t32;       // FP32 table [k][m]
    const uint16_t* thi;    // BF16 split tables [k][m]
    const uint16_t* tlo;
    float* out;
    int M, N, K;            // output rows, output columns, summed extent
    int O;                  // outer count (column pass: planes above the axis)
    int rowpass;            // 1: X = data [M][K], Y = table [N][K]; 0: X = table, Y = data [K][N]
    float scale;
    int poisson, quirk;
    Geometry g;
    const float* cx;
    const float* cy;
    const float* cz;
};

__device__ __forceinline__ float poisson_at(float v, int i, int j, const GemmArgs& a) {
    return poisson_apply(v, i, j, a.g, a.quirk, a.cx, a.cy, a.cz);
}

// LDS tiles: X [BM][pitch], Y [BN][pitch]; FP32 (PREC 1) or bf16 hi/lo (PREC 2)
template <int PREC>
struct GemmLds;
template <>
struct GemmLds<1> {
    float x[BM][PF32];
    float y[BN][PF32];
};
template <>
struct GemmLds<2> {
    uint16_t xh[BM][PB16], xl[BM][PB16];
    uint16_t yh[BN][PB16], yl[BN][PB16];
};

// Column pass, data operand Y.  FP32: kept as loaded, [k][j] with a pitch of
// BN + 16 floats (the 16 lanes of a fragment read consecutive j; the four
// k rows of a read are 16 banks apart).  BF16: transposed to [j][k] with the
// 4-element k groups XOR-swizzled by bits 4-5 of j, so that the transposing
// 8-byte writes of 16 lanes (rows 4 apart) hit distinct banks while a
// fragment's 8 k values stay contiguous and in order.
constexpr int PYK = BN + 16;
static_assert(BK * PYK == BN * PF32, "column-pass FP32 Y tile reuses the Y array");
__device__ __forceinline__ int yswz(int j) { return ((j >> 4) & 3) << 3; }

// Staging.  A block has NT threads (256 for F32, 512 for BF16x3); each moves
// NQ = 1024 / NT rows of 4 data values and 16 (F32) or 8 (BF16, hi and lo)
// table entries per K chunk.  The global loads of chunk c + 1 are issued into
// registers (DataRegs / TableRegs) before the MFMAs of chunk c and written to
// LDS after them.
template <int NT>
struct DataRegs {
    float v[1024 / NT][4];   // [piece row][4 consecutive elements]
};
struct TableRegs {
    float f[4][4];    // FP32 table: 4 x float4
    i16x8 h[2], l[2]; // BF16 table: up to 2 x 8 hi / lo
};

// Data operand, K contiguous in memory (row pass X = in[r][m]): thread e
// loads rows r0 + (e / 8) + (NT / 8) q (q < NQ), columns m0 + 4 (e % 8) .. + 4.
template <int NT>
__device__ __forceinline__ void load_data_kc(DataRegs<NT>& R, const float* src, int rows, int r0,
                                             int m0, int K) {
    const int e = threadIdx.x, c4 = (e % 8) * 4;
#pragma unroll
    for (int q = 0; q < 1024 / NT; ++q) {
        const int rg = r0 + e / 8 + (NT / 8) * q, m = m0 + c4;
        if (rg < rows && m + 3 < K && (K & 3) == 0) {
            const float4 t = *reinterpret_cast<const float4*>(src + size_t(rg) * K + m);
            R.v[q][0] = t.x; R.v[q][1] = t.y; R.v[q][2] = t.z; R.v[q][3] = t.w;
        } else {
#pragma unroll
            for (int c = 0; c < 4; ++c)
                R.v[q][c] = (rg < rows && m + c < K) ? src[size_t(rg) * K + m + c] : 0.0f;
        }
    }
}

// Data operand, N contiguous in memory (column pass Y = in[m][s]): thread e
// loads rows m0 + NQ (e / 32) .. + NQ at columns j0 + 4 (e % 32) .. + 4.
template <int NT>
__device__ __forceinline__ void load_data_nc(DataRegs<NT>& R, const float* src, int N, int K,
                                             int j0, int m0) {
    constexpr int NQ = 1024 / NT;
    const int e = threadIdx.x, kq = (e / 32) * NQ, j = j0 + (e % 32) * 4;
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        const int m = m0 + kq + q;
        if (m < K && j + 3 < N && (N & 3) == 0) {
            const float4 t = *reinterpret_cast<const float4*>(src + size_t(m) * N + j);
            R.v[q][0] = t.x; R.v[q][1] = t.y; R.v[q][2] = t.z; R.v[q][3] = t.w;
        } else {
#pragma unroll
            for (int c = 0; c < 4; ++c) R.v[q][c] = (m < K && j + c < N) ? src[size_t(m) * N + j + c] : 0.0f;
        }
    }
}

// Table [k][m] (K contiguous): FP32 (256 threads) thread e loads rows
// t0 + e / 8 + 32 q, columns m0 + 4 (e % 8); BF16 thread e loads rows
// t0 + e / 4 + (NT / 4) q (q < 512 / NT), columns m0 + 8 (e % 4).
template <int PREC, int NT>
__device__ __forceinline__ void load_table(TableRegs& T, const GemmArgs& a, int t0, int nrows,
                                           int m0) {
    const int K = a.K, e = threadIdx.x;
    if constexpr (PREC == 1) {
        const int c4 = (e % 8) * 4;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int tr = t0 + e / 8 + 32 * q, m = m0 + c4;
            if (tr < nrows && m + 3 < K && (K & 3) == 0) {
                const float4 t = *reinterpret_cast<const float4*>(a.t32 + size_t(tr) * K + m);
                T.f[q][0] = t.x; T.f[q][1] = t.y; T.f[q][2] = t.z; T.f[q][3] = t.w;
            } else {
#pragma unroll
                for (int c = 0; c < 4; ++c)
                    T.f[q][c] = (tr < nrows && m + c < K) ? a.t32[size_t(tr) * K + m + c] : 0.0f;
            }
        }
    } else {
        const int c8 = (e % 4) * 8;
#pragma unroll
        for (int q = 0; q < 512 / NT; ++q) {
            const int tr = t0 + e / 4 + (NT / 4) * q, m = m0 + c8;
            if (tr < nrows && m + 7 < K && (K & 7) == 0) {
                T.h[q] = *reinterpret_cast<const i16x8*>(a.thi + size_t(tr) * K + m);
                T.l[q] = *reinterpret_cast<const i16x8*>(a.tlo + size_t(tr) * K + m);
            } else {
#pragma unroll
                for (int c = 0; c < 8; ++c) {
                    const bool ok = tr < nrows && m + c < K;
                    T.h[q][c] = ok ? short(a.thi[size_t(tr) * K + m + c]) : short(0);
                    T.l[q][c] = ok ? short(a.tlo[size_t(tr) * K + m + c]) : short(0);
                }
            }
        }
    }
}

// 4 floats -> packed bf16 hi and lo (2 x 64 bits): hi = rne(x), lo = rne(x - hi),
// two values per conversion instruction
__device__ __forceinline__ void split4(const float (&x)[4], uint2& hi, uint2& lo) {
    const uint32_t h01 = StoreBF16::rne2(x[0], x[1]), h23 = StoreBF16::rne2(x[2], x[3]);
    hi = make_uint2(h01, h23);
    lo = make_uint2(StoreBF16::rne2(x[0] - __uint_as_float(h01 << 16),
                                    x[1] - __uint_as_float(h01 & 0xffff0000u)),
                    StoreBF16::rne2(x[2] - __uint_as_float(h23 << 16),
                                    x[3] - __uint_as_float(h23 & 0xffff0000u)));
}

// 4 K-consecutive values of row r of operand tile X (isx) or Y, at column k
template <int PREC>
__device__ __forceinline__ void put4(GemmLds<PREC>& L, bool isx, int r, int k, const float (&x)[4]) {
    if constexpr (PREC == 1) {
        float* d = isx ? &L.x[r][k] : &L.y[r][k];
        *reinterpret_cast<float4*>(d) = make_float4(x[0], x[1], x[2], x[3]);
    } else {
        uint2 h, l;
        split4(x, h, l);
        *reinterpret_cast<uint2*>(isx ? &L.xh[r][k] : &L.yh[r][k]) = h;
        *reinterpret_cast<uint2*>(isx ? &L.xl[r][k] : &L.yl[r][k]) = l;
    }
}

template <int PREC, int NT>
__device__ __forceinline__ void store_data_kc(GemmLds<PREC>& L, const DataRegs<NT>& R) {
    const int e = threadIdx.x, c4 = (e % 8) * 4;
#pragma unroll
    for (int q = 0; q < 1024 / NT; ++q) put4<PREC>(L, true, e / 8 + (NT / 8) * q, c4, R.v[q]);
}

// 2 floats -> packed bf16 hi and lo (32 bits each)
__device__ __forceinline__ void split2(float x0, float x1, uint32_t& hi, uint32_t& lo) {
    hi = StoreBF16::rne2(x0, x1);
    lo = StoreBF16::rne2(x0 - __uint_as_float(hi << 16), x1 - __uint_as_float(hi & 0xffff0000u));
}

// column pass: FP32 rows k as loaded; BF16 transposed (NQ x 4 piece) and swizzled
template <int PREC, int NT>
__device__ __forceinline__ void store_data_nc(GemmLds<PREC>& L, const DataRegs<NT>& R) {
    constexpr int NQ = 1024 / NT;
    const int e = threadIdx.x, k4 = (e / 32) * NQ, j = (e % 32) * 4;
    if constexpr (PREC == 1) {
        static_assert(NT == 256, "the F32 column pass stages 4 x 4 pieces");
        float* y = &L.y[0][0];
#pragma unroll
        for (int q = 0; q < 4; ++q)
            *reinterpret_cast<float4*>(y + (k4 + q) * PYK + j) =
                make_float4(R.v[q][0], R.v[q][1], R.v[q][2], R.v[q][3]);
    } else if constexpr (NQ == 4) {
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const float x[4] = {R.v[0][c], R.v[1][c], R.v[2][c], R.v[3][c]};
            put4<PREC>(L, false, j + c, k4 ^ yswz(j + c), x);
        }
    } else {   // 2 k values per row (the swizzle moves whole 8-element groups)
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            uint32_t h, l;
            split2(R.v[0][c], R.v[1][c], h, l);
            *reinterpret_cast<uint32_t*>(&L.yh[j + c][k4 ^ yswz(j + c)]) = h;
            *reinterpret_cast<uint32_t*>(&L.yl[j + c][k4 ^ yswz(j + c)]) = l;
        }
    }
}

template <int PREC, int NT>
__device__ __forceinline__ void store_table(GemmLds<PREC>& L, bool isx, const TableRegs& T) {
    const int e = threadIdx.x;
    if constexpr (PREC == 1) {
        const int c4 = (e % 8) * 4;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            float* d = isx ? &L.x[e / 8 + 32 * q][c4] : &L.y[e / 8 + 32 * q][c4];
            *reinterpret_cast<float4*>(d) = make_float4(T.f[q][0], T.f[q][1], T.f[q][2], T.f[q][3]);
        }
    } else {
        const int c8 = (e % 4) * 8;
#pragma unroll
        for (int q = 0; q < 512 / NT; ++q) {
            const int r = e / 4 + (NT / 4) * q;
            *reinterpret_cast<i16x8*>(isx ? &L.xh[r][c8] : &L.yh[r][c8]) = T.h[q];
            *reinterpret_cast<i16x8*>(isx ? &L.xl[r][c8] : &L.yl[r][c8]) = T.l[q];
        }
    }
}

// Threads per block: F32 4 waves of 64 x 64 outputs; BF16x3 8 waves of 32 x 64
// (half the accumulators per wave, so twice the waves per SIMD to hide the
// loads' latency).
template <int PREC>
constexpr int gemm_threads() { return PREC == 2 ? 512 : 256; }

template <int PREC, int ROW>
__global__ __launch_bounds__(gemm_threads<PREC>()) void dct_gemm_kernel(GemmArgs a) {
    constexpr int NT = gemm_threads<PREC>();
    constexpr int WR = BM / (NT / 128);   // output rows per wave (64 or 32)
    constexpr int TT = WR / 16;           // 16-row MFMA tiles per wave
    __shared__ __align__(16) GemmLds<PREC> L;
    const int M = a.M, N = a.N, K = a.K;
    const int tm = (M + BM - 1) / BM, tn = (N + BN - 1) / BN;
    // tile order: the tiles sharing a data panel are adjacent, and each XCD
    // takes a contiguous range of tiles, so a panel is fetched from HBM once
    // into one XCD's L2 instead of once per XCD its tiles land on
    int bid = int(xcd_contiguous(blockIdx.x, gridDim.x));
    int it, jt, o;
    if (ROW) {                  // data panel = rows it; j tiles fastest
        jt = bid % tn; bid /= tn; it = bid % tm; o = bid / tm;
    } else {                    // data panel = columns jt; i tiles fastest
        it = bid % tm; bid /= tm; jt = bid % tn; o = bid / tn;
    }
    const int i0 = it * BM, j0 = jt * BN;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int li = lane & 15, lg = lane >> 4;
    const int wi = (wave >> 1) * WR, wj = (wave & 1) * 64;
    const size_t obase = size_t(o) * size_t(M) * size_t(N);

    f32x4 acc[TT][4];
#pragma unroll
    for (int t = 0; t < TT; ++t)
#pragma unroll
        for (int u = 0; u < 4; ++u) acc[t][u] = f32x4{0.f, 0.f, 0.f, 0.f};

    const float* dsrc = ROW ? a.data : a.data + size_t(o) * size_t(K) * size_t(N);
    // FP32: two register sets, chunk c + 2 is loaded while chunk c is
    // multiplied; BF16x3: one set (chunk c + 1), which keeps four waves per
    // SIMD (a second set costs the column pass its second block; in the row
    // pass alone it measured no gain at 512 x 512 x 256)
    constexpr int DEPTH = PREC == 1 ? 2 : 1;
    DataRegs<NT> R0, R1;
    TableRegs T0, T1;
    auto load = [&](DataRegs<NT>& R, TableRegs& T, int m0) {
        if (ROW) {
            load_data_kc<NT>(R, dsrc, M, i0, m0, K);
            load_table<PREC, NT>(T, a, j0, N, m0);
        } else {
            load_table<PREC, NT>(T, a, i0, M, m0);
            load_data_nc<NT>(R, dsrc, N, K, j0, m0);
        }
    };
    auto chunk = [&](DataRegs<NT>& R, TableRegs& T, int m0) {
        if (ROW) {
            store_data_kc<PREC, NT>(L, R);
            store_table<PREC, NT>(L, false, T);
        } else {
            store_table<PREC, NT>(L, true, T);
            store_data_nc<PREC, NT>(L, R);
        }
        __syncthreads();
        if (m0 + DEPTH * BK < K) load(R, T, m0 + DEPTH * BK);
        if constexpr (PREC == 1) {
#pragma unroll
            for (int kk = 0; kk < BK / 4; ++kk) {
                const int kc = 4 * kk + lg;
                float xf[4], yf[4];
#pragma unroll
                for (int t = 0; t < 4; ++t) xf[t] = L.x[wi + t * 16 + li][kc];
#pragma unroll
                for (int u = 0; u < 4; ++u)
                    yf[u] = ROW ? L.y[wj + u * 16 + li][kc] : (&L.y[0][0])[kc * PYK + wj + u * 16 + li];
#pragma unroll
                for (int t = 0; t < 4; ++t)
#pragma unroll
                    for (int u = 0; u < 4; ++u)
                        acc[t][u] = __builtin_amdgcn_mfma_f32_16x16x4f32(xf[t], yf[u], acc[t][u], 0, 0, 0);
            }
        } else {
            bf16x8 xh[TT], xl[TT], yh[4], yl[4];
#pragma unroll
            for (int t = 0; t < TT; ++t) {
                xh[t] = __builtin_bit_cast(bf16x8, *reinterpret_cast<const i16x8*>(&L.xh[wi + t * 16 + li][8 * lg]));
                xl[t] = __builtin_bit_cast(bf16x8, *reinterpret_cast<const i16x8*>(&L.xl[wi + t * 16 + li][8 * lg]));
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int r = wj + u * 16 + li;
                const int kq = ROW ? 8 * lg : (8 * lg) ^ yswz(r);
                yh[u] = __builtin_bit_cast(bf16x8, *reinterpret_cast<const i16x8*>(&L.yh[r][kq]));
                yl[u] = __builtin_bit_cast(bf16x8, *reinterpret_cast<const i16x8*>(&L.yl[r][kq]));
            }
            // each product over all 16 tiles before the next: consecutive MFMAs
            // are independent (same per-accumulator order hi*hi, hi*lo, lo*hi)
#pragma unroll
            for (int t = 0; t < TT; ++t)
#pragma unroll
                for (int u = 0; u < 4; ++u)
                    acc[t][u] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(xh[t], yh[u], acc[t][u], 0, 0, 0);
#pragma unroll
            for (int t = 0; t < TT; ++t)
#pragma unroll
                for (int u = 0; u < 4; ++u)
                    acc[t][u] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(xh[t], yl[u], acc[t][u], 0, 0, 0);
#pragma unroll
            for (int t = 0; t < TT; ++t)
#pragma unroll
                for (int u = 0; u < 4; ++u)
                    acc[t][u] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(xl[t], yh[u], acc[t][u], 0, 0, 0);
        }
        __syncthreads();
    };
    load(R0, T0, 0);
    if constexpr (DEPTH == 2) {
        if (BK < K) load(R1, T1, BK);
        for (int m0 = 0; m0 < K; m0 += 2 * BK) {
            chunk(R0, T0, m0);
            if (m0 + BK < K) chunk(R1, T1, m0 + BK);
        }
    } else {
        for (int m0 = 0; m0 < K; m0 += BK) chunk(R0, T0, m0);
    }
    // lane holds D rows 4*lg + r, column li of each 16 x 16 tile
#pragma unroll
    for (int t = 0; t < TT; ++t)
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int i = i0 + wi + t * 16 + 4 * lg + r;
                const int j = j0 + wj + u * 16 + li;
                if (i < M && j < N) {
                    const size_t idx = obase + size_t(i) * N + j;
                    float v = acc[t][u][r] * a.scale;
                    if (a.poisson) v = poisson_at(v, i, j, a);
                    a.out[idx] = v;
                }
            }
}

}  // namespace

int dct_mfma_max_extent() { return 4096; }

hipError_t launch_dct_mfma_pass(hipStream_t s, const float* in, float* out, const float* table,
                                const uint16_t* table_hi, const uint16_t* table_lo, Geometry g,
                                int axis, int inverse, int last, int precision,
                                const float* eig_x, const float* eig_y, const float* eig_z,
                                int fuse_poisson, int quirk) {
    GemmArgs a;
    a.data = in; a.out = out; a.t32 = table; a.thi = table_hi; a.tlo = table_lo;
    a.g = g; a.cx = eig_x; a.cy = eig_y; a.cz = eig_z;
    a.quirk = quirk;
    a.poisson = (last && !inverse && fuse_poisson) ? 1 : 0;
    if (axis == 0) {           // rows r of length W
        a.rowpass = 1; a.M = g.h * g.d; a.N = g.w; a.K = g.w; a.O = 1;
    } else if (axis == 1) {    // per plane: out[z][k][x] = sum_m C[k][m] in[z][m][x]
        a.rowpass = 0; a.M = g.h; a.N = g.w; a.K = g.h; a.O = g.d;
    } else {                   // out[k][s] = sum_m C[k][m] in[m][s], s over the plane
        a.rowpass = 0; a.M = g.d; a.N = g.w * g.h; a.K = g.d; a.O = 1;
    }
    a.scale = 1.0f;
    if (last) a.scale = inverse ? float(1.0 / (double(g.n()) * 8.0)) : 8.0f;
    const long tm = (a.M + BM - 1) / BM, tn = (a.N + BN - 1) / BN;
    const long blocks = tm * tn * a.O;
    if (blocks > 0x7fffffffL) return hipErrorInvalidValue;
    const dim3 grid{static_cast<unsigned>(blocks), 1, 1};
    const dim3 block{unsigned(precision == 2 ? gemm_threads<2>() : gemm_threads<1>()), 1, 1};
    if (precision == 2) {
        if (a.rowpass) hipLaunchKernelGGL((dct_gemm_kernel<2, 1>), grid, block, 0, s, a);
        else hipLaunchKernelGGL((dct_gemm_kernel<2, 0>), grid, block, 0, s, a);
    } else {
        if (a.rowpass) hipLaunchKernelGGL((dct_gemm_kernel<1, 1>), grid, block, 0, s, a);
        else hipLaunchKernelGGL((dct_gemm_kernel<1, 0>), grid, block, 0, s, a);
    }
    return hipGetLastError();
}

}  // namespace amb
