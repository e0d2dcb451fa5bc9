// Workgroup -> tile maps of the launches, as pure functions of the workgroup index. Host and
// device, so that the maps are checked on the CPU (tests/test_tiles.py). Workgroups are dealt to
// the eight XCDs round robin (XCD = index % 8); each map gives XCD k a contiguous range of tiles so
// that neighbouring tiles share that XCD's L2.
#pragma once

#include <hip/hip_runtime.h>

namespace taastar {

// one launch over n tiles with n workgroups: XCD k takes the k-th of eight contiguous ranges (the
// first n % 8 ranges one tile longer), its workgroups in order
__host__ __device__ __forceinline__ int xcd_tile_of(int b, int n) {
    const int q = n >> 3, r = n & 7;
    const int xcd = b & 7, slot = b >> 3;
    return xcd * q + (xcd < r ? xcd : r) + slot;
}

// a batched launch over total tiles with 8 * per workgroups (per = ceil(total / 8), the grid the
// host sizes): XCD k takes tiles [k * per, (k + 1) * per); -1 for a workgroup of the rounding
__host__ __device__ __forceinline__ int batch_tile_of(int b, int total) {
    const int per = (total + 7) >> 3;
    const int slot = b >> 3, t = (b & 7) * per + slot;
    return slot < per && t < total ? t : -1;
}

}  // namespace taastar
