// datagen.hip -- seeded, counter-based generators for the synthetic benchmark columns.
//
// rand(seed, stream, i) = mix(key(seed, stream) + i), key(seed, stream) = mix(mix(seed) ^ stream)
// where mix is the SplitMix64 finaliser (Steele, Lea, Flood: "Fast splittable pseudorandom
// number generators", OOPSLA 2014), including its golden-ratio increment.  Row i of a column
// depends only on (seed, stream, i), so the device and ahead_amd/datagen.py produce identical
// columns and flip sets.
#include "host_common.hpp"

namespace {

__host__ __device__ __forceinline__ uint64_t mix(uint64_t x) {
    uint64_t z = x + 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

__host__ __device__ __forceinline__ uint64_t stream_key(uint64_t seed, uint64_t stream) {
    return mix(mix(seed) ^ stream);
}

template <typename T>
__global__ __launch_bounds__(256) void k_gen_uniform(T *out, uint64_t n, uint64_t key, uint64_t lo,
                                                     uint64_t range, uint64_t a) {
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < n; i += stride) {
        const uint64_t r = mix(key + i);
        const uint64_t v = lo + (((r >> 32) * range) >> 32);
        out[i] = static_cast<T>(a ? v * a : v);
    }
}

template <typename T>
__global__ __launch_bounds__(256) void k_flip(T *codes, uint64_t n, uint64_t key_hit, uint64_t key_bit,
                                              uint64_t threshold) {
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < n; i += stride) {
        if (mix(key_hit + i) < threshold) {
            const uint64_t bit = ((mix(key_bit + i) >> 32) * (8 * sizeof(T))) >> 32;
            codes[i] = static_cast<T>(codes[i] ^ (static_cast<T>(1) << bit));
        }
    }
}

uint32_t grid_for(uint64_t n) {
    const uint64_t blocks = (n + 255) / 256;
    return static_cast<uint32_t>(blocks < 4096 ? (blocks ? blocks : 1) : 4096);
}

}  // namespace

using namespace ahead_host;

extern "C" int ahead_hip_gen_uniform(void *out, uint32_t width, uint64_t n, uint64_t seed,
                                     uint64_t stream, uint64_t lo, uint64_t hi, uint64_t a,
                                     void *hstream) {
    if (out == nullptr && n) return fail(AHEAD_ERR_INVALID, "gen_uniform: NULL output");
    if (hi < lo || hi - lo >= (1ull << 32)) return fail(AHEAD_ERR_INVALID, "gen_uniform: range must be < 2^32");
    if (n == 0) return AHEAD_OK;
    const uint64_t key = stream_key(seed, stream), range = hi - lo + 1;
    hipStream_t s = static_cast<hipStream_t>(hstream);
    const dim3 g(grid_for(n)), b(256);
    switch (width) {
        case 1: hipLaunchKernelGGL(k_gen_uniform<uint8_t>, g, b, 0, s, static_cast<uint8_t *>(out), n, key, lo, range, a); break;
        case 2: hipLaunchKernelGGL(k_gen_uniform<uint16_t>, g, b, 0, s, static_cast<uint16_t *>(out), n, key, lo, range, a); break;
        case 4: hipLaunchKernelGGL(k_gen_uniform<uint32_t>, g, b, 0, s, static_cast<uint32_t *>(out), n, key, lo, range, a); break;
        case 8: hipLaunchKernelGGL(k_gen_uniform<uint64_t>, g, b, 0, s, static_cast<uint64_t *>(out), n, key, lo, range, a); break;
        default: return fail(AHEAD_ERR_INVALID, "gen_uniform: width must be 1, 2, 4 or 8");
    }
    AHEAD_HIP_TRY(hipGetLastError());
    return AHEAD_OK;
}

extern "C" int ahead_hip_inject_flips(void *codes, uint32_t width, uint64_t n, uint64_t seed,
                                      uint64_t stream, uint64_t threshold, void *hstream) {
    if (codes == nullptr && n) return fail(AHEAD_ERR_INVALID, "inject_flips: NULL codes");
    if (n == 0 || threshold == 0) return AHEAD_OK;
    const uint64_t kh = stream_key(seed, stream), kb = stream_key(seed, stream + 1);
    hipStream_t s = static_cast<hipStream_t>(hstream);
    const dim3 g(grid_for(n)), b(256);
    switch (width) {
        case 2: hipLaunchKernelGGL(k_flip<uint16_t>, g, b, 0, s, static_cast<uint16_t *>(codes), n, kh, kb, threshold); break;
        case 4: hipLaunchKernelGGL(k_flip<uint32_t>, g, b, 0, s, static_cast<uint32_t *>(codes), n, kh, kb, threshold); break;
        case 8: hipLaunchKernelGGL(k_flip<uint64_t>, g, b, 0, s, static_cast<uint64_t *>(codes), n, kh, kb, threshold); break;
        default: return fail(AHEAD_ERR_INVALID, "inject_flips: width must be 2, 4 or 8");
    }
    AHEAD_HIP_TRY(hipGetLastError());
    return AHEAD_OK;
}
