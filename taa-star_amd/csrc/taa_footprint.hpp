// The Catmull-Rom history footprint on one axis (taa.comp:441-514): base texel k = int(floor(uv *
// n - 0.5)) (saturating, as V_CVT_I32_F32 converts), taps k - 1 .. k + 2, each clamped to [0, n - 1]
// unless all four lie inside. Host and device, so that the index rules are checked on the CPU
// (tests/test_footprint.py).
#pragma once

#include <hip/hip_runtime.h>

namespace taastar {

// the base pre-clamped to [-2, n + 1]: every tap clamped to the edge still reads the texel the
// unclamped tap would, and k + 2 cannot overflow for any input (n below 2^28, as the host limits
// every image to 4 GiB)
__host__ __device__ __forceinline__ int cr_base(int k_raw, int n) {
    const int k = k_raw > -2 ? k_raw : -2;
    return k < n + 1 ? k : n + 1;
}

// all four taps of a pre-clamped base inside the image
__host__ __device__ __forceinline__ bool cr_inside(int k, int n) { return k >= 1 && k + 2 <= n - 1; }

// the same test for a base that was not pre-clamped (any int, INT_MIN and INT_MAX included)
__host__ __device__ __forceinline__ bool cr_inside_raw(int k, int n) { return k >= 1 && k <= n - 3; }

}  // namespace taastar
