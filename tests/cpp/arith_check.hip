// Exhaustive check of the short correctly rounded forms in lvn_device.hpp against the
// compiler's IEEE sqrtf and division, on the GPU:
//   sqrt_ge1(x)  vs sqrtf(x)  for every float in [1, 4) (both exponent parities), every
//                             97th float of each binade [4, 2^128), and +inf;
//   recip_1_2(d) vs 1.f / d   for every float in [1, 2).
// lvn_arith_check(out) writes {sqrt mismatches, sqrt values tried, recip mismatches,
// recip values tried} and returns 0, or a HIP error code.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../leven_amd/csrc/lvn_device.hpp"

namespace {

constexpr uint32_t kOne = 0x3f800000u;        // 1.0f
constexpr uint32_t kFour = 0x40800000u;       // 4.0f
constexpr uint32_t kMantissa = 1u << 23;
constexpr uint32_t kSparse = 97;
constexpr uint32_t kDenseSqrt = 2 * kMantissa;                           // [1, 4)
constexpr uint32_t kSparseBinades = 254 - 129 + 1;                       // exponents of [4, 2^128)
constexpr uint32_t kSparsePerBinade = (kMantissa + kSparse - 1) / kSparse;
constexpr uint32_t kSqrtCount = kDenseSqrt + kSparseBinades * kSparsePerBinade + 1;  // + inf

__device__ uint32_t sqrt_input(uint32_t i) {
    if (i < kDenseSqrt) return kOne + i;
    i -= kDenseSqrt;
    if (i < kSparseBinades * kSparsePerBinade) {
        const uint32_t binade = i / kSparsePerBinade, k = i % kSparsePerBinade;
        return kFour + binade * kMantissa + k * kSparse;
    }
    return 0x7f800000u;  // +inf
}

__global__ __launch_bounds__(256) void k_check(unsigned long long* counts) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    unsigned bad_sqrt = 0, bad_recip = 0;
    if (i < kSqrtCount) {
        const float x = __uint_as_float(sqrt_input(i));
        bad_sqrt = __float_as_uint(lvn::sqrt_ge1(x)) != __float_as_uint(sqrtf(x));
    }
    if (i < kMantissa) {
        const float d = __uint_as_float(kOne + i);
        bad_recip = __float_as_uint(lvn::recip_1_2(d)) != __float_as_uint(1.f / d);
    }
    const unsigned long long bs = __popcll(__ballot(bad_sqrt)), br = __popcll(__ballot(bad_recip));
    if ((threadIdx.x & 63) == 0 && (bs | br)) {
        atomicAdd(&counts[0], bs);
        atomicAdd(&counts[2], br);
    }
}

}  // namespace

extern "C" int lvn_arith_check(unsigned long long out[4]) {
    unsigned long long* d = nullptr;
    hipError_t e = hipMalloc(&d, 4 * sizeof(unsigned long long));
    if (e != hipSuccess) return (int)e;
    e = hipMemset(d, 0, 4 * sizeof(unsigned long long));
    if (e == hipSuccess) {
        const uint32_t blocks = (kSqrtCount + 255) / 256;  // kSqrtCount > kMantissa: covers both
        hipLaunchKernelGGL(k_check, dim3(blocks), dim3(256), 0, 0, d);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpy(out, d, 4 * sizeof(unsigned long long), hipMemcpyDeviceToHost);
    (void)hipFree(d);
    out[1] = kSqrtCount;
    out[3] = kMantissa;
    return (int)e;
}
