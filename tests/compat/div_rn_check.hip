// Device check of csrc/div_rn.hpp's division by a constant against the
// device's IEEE division operator (hipcc divides FP32 with correct rounding by
// default), in the fused sweep's use: b = alpha within [2^-60, 2^36], y = RN(1/b)
// (the host's double-then-float reciprocal, equal to the device's 1.0f / b),
// a = |p| or |q| >= 0.  For every significand of b at exponents -60, -20, -1,
// 0, 1, 9, 23 and 35:
//   8 numerators hashed over significands and exponents in [-90, 63]: the
//   quotient equals a / b bit for bit;
//   a = 0: the quotient is +0;
//   2 numerators below 2^-90 (subnormals included): the quotient and a / b
//   are on the same side of 1 (the sweep only compares such a quotient with 1).
// Prints the mismatch counts; exit status 0 when all are zero.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>

#include "../../ambient_image_processor_amd/csrc/div_rn.hpp"

namespace {

constexpr int N_EXP = 8;
__constant__ int kExp[N_EXP] = {-60, -20, -1, 0, 1, 9, 23, 35};
constexpr uint32_t N_SIG = 1u << 23;

__device__ uint32_t mix(uint32_t x) {   // 32-bit integer hash (xorshift-multiply)
    x ^= x >> 16; x *= 0x7feb352du;
    x ^= x >> 15; x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

__global__ void check(unsigned long long* bad) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= N_SIG * N_EXP) return;
    const uint32_t m = t % N_SIG;
    const int e = kExp[t / N_SIG];
    const float b = __uint_as_float((uint32_t(127 + e) << 23) | m);
    const float y = 1.0f / b;
    unsigned long long nq = 0, nz = 0, nt = 0;
    for (uint32_t k = 0; k < 8; ++k) {
        const uint32_t h = mix(t * 8u + k), h2 = mix(h ^ 0x9e3779b9u);
        const int ea = -90 + int(h2 % 154u);   // exponents -90 .. 63
        const float a = __uint_as_float((uint32_t(127 + ea) << 23) | (h & 0x7fffffu));
        if (__float_as_uint(amb::div_rn(a, b, y)) != __float_as_uint(a / b)) ++nq;
    }
    if (__float_as_uint(amb::div_rn(0.0f, b, y)) != 0u) ++nz;
    for (uint32_t k = 0; k < 2; ++k) {
        const uint32_t h = mix(t * 2u + k + 0x51ed27u);
        const float a = __uint_as_float(h % 0x12800000u);   // [0, 2^-90)
        if ((amb::div_rn(a, b, y) > 1.0f) != (a / b > 1.0f)) ++nt;
    }
    if (nq) atomicAdd(&bad[0], nq);
    if (nz) atomicAdd(&bad[1], nz);
    if (nt) atomicAdd(&bad[2], nt);
}

}  // namespace

int main() {
    unsigned long long* bad = nullptr;
    if (hipMalloc(&bad, 3 * sizeof(unsigned long long)) != hipSuccess) return 2;
    if (hipMemset(bad, 0, 3 * sizeof(unsigned long long)) != hipSuccess) return 2;
    const uint32_t n = N_SIG * N_EXP, threads = 256;
    hipLaunchKernelGGL(check, dim3((n + threads - 1) / threads), dim3(threads), 0, 0, bad);
    if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess) return 2;
    unsigned long long h[3] = {0, 0, 0};
    if (hipMemcpy(h, bad, sizeof(h), hipMemcpyDeviceToHost) != hipSuccess) return 2;
    (void)hipFree(bad);
    std::printf("divisors %u quotients %llu mismatches %llu; zero mismatches %llu; "
                "tiny-numerator comparisons with 1 mismatches %llu\n",
                n, 8ull * n, h[0], h[1], h[2]);
    return h[0] == 0 && h[1] == 0 && h[2] == 0 ? 0 : 1;
}
