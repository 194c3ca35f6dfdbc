// Float min/max through integer atomics (device code only): floats are mapped
// to unsigned ints whose order is the float order (negatives reversed), so
// atomicMin/atomicMax on the mapped bits give exact float minima and maxima
// whatever the order of arrival.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace pico {

__device__ __forceinline__ uint32_t map_f(float v) {
    const uint32_t u = __float_as_uint(v);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float unmap_f(uint32_t m) {
    return __uint_as_float((m & 0x80000000u) ? (m & 0x7fffffffu) : ~m);
}
// identities: the mapped +inf (start of a minimum) and -inf (start of a maximum)
constexpr uint32_t kMappedPosInf = 0xff800000u;
constexpr uint32_t kMappedNegInf = 0x007fffffu;

__device__ __forceinline__ void atomic_min_f(uint32_t* a, float v) { atomicMin(a, map_f(v)); }
__device__ __forceinline__ void atomic_max_f(uint32_t* a, float v) { atomicMax(a, map_f(v)); }

}  // namespace pico
