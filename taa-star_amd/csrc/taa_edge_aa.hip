// Edge-AA fallback for the texels the segmentation mask puts in category 1 (the prepare + FXAA
// passes of source/taa.hpp:1061-1107), with the library's own filter: every texel gets alpha =
// luma (the prepare pass); a category-1 texel with enough local luma contrast is blended with the
// 4-neighbour across its edge (include/taastar/taa_abi.h states the filter). FP32 arithmetic
// without contraction (this file is compiled with -ffp-contract=off), in the oracle's order.
//
// One texel per lane, 64x4 texels per workgroup, so a wave reads and writes 64 consecutive texels
// of a row. Most texels are outside category 1 and cost a mask read and a copy; the 3x3
// neighbourhood of the others is read from global memory (L1 / L2 hits: the rows around a wave's
// row are read by the neighbouring waves of the same workgroup).
#include <hip/hip_runtime.h>

#include "taa_device.hpp"
#include "taa_internal.hpp"

namespace taastar {
namespace {

struct EdgeAaArgs {
    const void* src;
    const uint32_t* mask;
    void* dst;
    int W, H, pitch, seg_pitch;
    float subpix, thr, thr_min;
};

__device__ __forceinline__ float aa_luma(f3 c) { return c.x * 0.299f + c.y * 0.587f + c.z * 0.114f; }

template <bool S16>
__global__ __launch_bounds__(256) void k_edge_aa(EdgeAaArgs a) {
    const int x = (int)blockIdx.x * 64 + (int)(threadIdx.x & 63u);
    const int y = (int)blockIdx.y * 4 + (int)(threadIdx.x >> 6);
    if (x >= a.W || y >= a.H) return;
    const uint32_t i = (uint32_t)y * (uint32_t)a.pitch + (uint32_t)x;
    const f3 cC = xyz(ld_rgba<S16>(a.src, i));
    const float lC = aa_luma(cC);
    float4 out = make_float4(cC.x, cC.y, cC.z, lC);
    const uint32_t m = gload<uint32_t>(a.mask, ((uint32_t)y * (uint32_t)a.seg_pitch + (uint32_t)x) * 4u);
    if ((m & 3u) == 1u) {
        auto at = [&](int dx, int dy) {
            const uint32_t j = (uint32_t)clampi(y + dy, 0, a.H - 1) * (uint32_t)a.pitch + (uint32_t)clampi(x + dx, 0, a.W - 1);
            return xyz(ld_rgba<S16>(a.src, j));
        };
        const float lN = aa_luma(at(0, -1)), lS = aa_luma(at(0, 1)), lW = aa_luma(at(-1, 0)), lE = aa_luma(at(1, 0));
        const float lNW = aa_luma(at(-1, -1)), lNE = aa_luma(at(1, -1)), lSW = aa_luma(at(-1, 1)), lSE = aa_luma(at(1, 1));
        const float lmin = fminf(lC, fminf(fminf(lN, lS), fminf(lW, lE)));
        const float lmax = fmaxf(lC, fmaxf(fmaxf(lN, lS), fmaxf(lW, lE)));
        const float range = lmax - lmin;
        if (!(range < fmaxf(a.thr_min, lmax * a.thr))) {
            const float gx = (lNE + 2.0f * lE + lSE) - (lNW + 2.0f * lW + lSW);
            const float gy = (lSW + 2.0f * lS + lSE) - (lNW + 2.0f * lN + lNE);
            int dx = 0, dy = 0;
            if (fabsf(gy) >= fabsf(gx)) dy = fabsf(lN - lC) >= fabsf(lS - lC) ? -1 : 1;
            else dx = fabsf(lW - lC) >= fabsf(lE - lC) ? -1 : 1;
            const float avg = (((lN + lS) + (lW + lE)) * 2.0f + ((lNW + lNE) + (lSW + lSE))) / 12.0f;
            const float s = gclamp(fabsf(avg - lC) / range, 0.0f, 1.0f);
            const float blend = a.subpix * 0.5f * s;
            const f3 cX = at(dx, dy);
            out = make_float4(cC.x * (1.0f - blend) + cX.x * blend, cC.y * (1.0f - blend) + cX.y * blend,
                              cC.z * (1.0f - blend) + cX.z * blend, lC);
        }
    }
    st_rgba<S16>(a.dst, i, out);
}

}  // namespace

int launch_edge_aa(const void* src, const uint32_t* mask, void* dst, int W, int H, int pitch, int seg_pitch, int fmt,
                   float subpix, float thr, float thr_min, void* stream) {
    EdgeAaArgs a{src, mask, dst, W, H, pitch, seg_pitch, subpix, thr, thr_min};
    const dim3 grid((unsigned)((W + 63) / 64), (unsigned)((H + 3) / 4));
    hipStream_t st = static_cast<hipStream_t>(stream);
    if (fmt == TAA_FMT_RGBA16F) hipLaunchKernelGGL(k_edge_aa<true>, grid, dim3(256), 0, st, a);
    else hipLaunchKernelGGL(k_edge_aa<false>, grid, dim3(256), 0, st, a);
    return (int)hipGetLastError();
}

}  // namespace taastar
