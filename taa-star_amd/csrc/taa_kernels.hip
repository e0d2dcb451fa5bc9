// gfx950 kernels of the temporal resolve (shaders/taa.comp) and FidelityFX CAS
// (shaders/sharpen_cas.comp, shaders/ffx_cas.h:424-537).
//
// Work decomposition (one 256-thread workgroup per output tile TW x TH):
//   1. stage the current colour (tone-mapped, converted to YCoCg once per texel) and depth of the
//      tile plus an apron into LDS: 1 px for the 3x3 neighbourhood, 2 px when CAS is fused;
//   2. resolve every texel of the resolve region (the tile, plus a 1-px ring when CAS is fused)
//      -- neighbourhood box from LDS, closest-depth tap from LDS, velocity and the 4x4
//      Catmull-Rom history footprint gathered from global memory (L2 / MALL resident);
//   3. with CAS fused, keep the resolved screen colour of the region in LDS and sharpen the tile
//      from it, so the unsharpened image never goes to HBM.
// Tiles are mapped to workgroups so that each XCD (blockIdx % 8) works on one contiguous band of
// tiles, which keeps apron and history re-reads inside that XCD's L2.
//
// This file is compiled with -ffp-contract=off: the FP32 arithmetic path evaluates every
// expression of taa.comp in source order with correctly rounded division and square root, so it
// reproduces the scalar FP32 transcription bit for bit.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "taa_device.hpp"
#include "taa_fast_math.hpp"
#include "taa_internal.hpp"

namespace taastar {
namespace {

// ------------------------------------------------------------------------------------------
// tile geometry and LDS

template <int TW_, int TH_, int PAD_, bool CAS_, bool UP_ = false, bool SEG_ = false, bool VEL16_ = false>
struct Geo {
    static constexpr int TW = TW_, TH = TH_, PAD = PAD_;
    static constexpr bool CAS = CAS_, UP = UP_;  // UP: upsampling path (input size != output size)
    static constexpr bool SEG = SEG_;            // SEG: the segmentation mask's inputs are staged too
    // VEL16: an RG16F velocity image is staged as well (its 32-bit texels); measured with RGBA16F
    // colour at 4K: 30.6 -> 31.4 Gpix/s. The RG32F equivalent ran slower with RGBA32F colour
    // (24.3 -> 17.4 at 1080p: the larger region costs the schedule its overlapped gathers).
    static constexpr bool VEL16 = VEL16_;
    static constexpr int NT = 256;
    static constexpr int A = CAS ? 2 : 1;  // input apron
    static constexpr int SW = TW + 2 * A, SH = TH + 2 * A, SP = SW + PAD;
    static constexpr int RA = CAS ? 1 : 0;  // resolve apron (CAS needs the resolved cross)
    static constexpr int RW = TW + 2 * RA, RH = TH + 2 * RA, RP = RW + PAD;
    // upsampling path: the input texels under the resolved region (+3 on every side) are staged
    // in a UCW x UCH rectangle, which holds them for every scale >= 1 (at most RW + 6 columns)
    static constexpr int UCW = UP ? RW + 8 : 1, UCH = UP ? RH + 8 : 1;
    struct Smem {
        float c0[UP ? 1 : SH * SP], c1[UP ? 1 : SH * SP], c2[UP ? 1 : SH * SP], d[UP ? 1 : SH * SP];
        // the resolved region for the fused CAS; the FP16 arithmetic on FP16 storage keeps each
        // texel's stored fp16 bits (8 bytes) in the space of r0 and r1 instead (packed())
        alignas(8) float r0[CAS ? RH * RP : 1];
        float r1[CAS ? RH * RP : 1], r2[CAS ? RH * RP : 1];
        __device__ uint2* packed() { return reinterpret_cast<uint2*>(r0); }
        float ux[SW], uy[SH];  // uv of the tile's (clamped) columns and rows
        // segmentation inputs of the staged region: unit normal, raw YCoCg luma, material id
        float nx[SEG ? SH * SP : 1], ny[SEG ? SH * SP : 1], nz[SEG ? SH * SP : 1], lum[SEG ? SH * SP : 1];
        uint32_t mat[SEG ? SH * SP : 1];
        uint32_t vel[VEL16 ? SH * SP : 1];  // RG16F velocity of the staged region
        // upsampling path: converted colour, depth and velocity (as ld_vel reads it) of the staged
        // rectangle. Staging only x and y of an RG velocity in the production form frees enough LDS
        // for a fifth 30x14 workgroup per CU, which ran slower (1080p -> 4K FP16: 0.221 against 0.205 ms)
        float4 uvel[UCW * UCH];
        float uc0[UCW * UCH], uc1[UCW * UCH], uc2[UCW * UCH], ud[UCW * UCH];
        // upsampling path, per resolved column (row): the bilinear axis of the depth and velocity
        // samples at uv + (d - 1) input texels, d = 0..2, and the current-sample search at the
        // corner offsets 0 and c (position, hit)
        BilinearAxis uax[UP ? 3 * RW : 1], uay[UP ? 3 * RH : 1];
        int ucx[UP ? RW : 1], ucy[UP ? RH : 1];
    };
};

// The input-sized images of the upsampling path as staged for one tile: input texel (x, y)
// (clamped coordinates) is entry (y - y0) * n + (x - x0) when 0 <= x - x0 < w and 0 <= y - y0 < h;
// w = 0 when nothing is staged.
struct UpIn {
    const float* c0;
    const float* c1;
    const float* c2;
    const float* d;
    const float4* v;
    int x0, y0, w, h, n;
    // the per-column and per-row tables (column j of the resolve region at [d * rw + j])
    const BilinearAxis* ax;
    const BilinearAxis* ay;
    const int* cx;  // getCurrentUpsampledColor's input column per output column, -1: none found
    const int* cy;  // ... input row per output row, -1: none
    int rw, rh;
};

// the staged segmentation planes (row pitch SP, same indexing as the colour planes)
struct SegPlanes {
    const float* nx;
    const float* ny;
    const float* nz;
    const float* lum;
    const uint32_t* mat;
};

struct PixelResult {
    f3 screen;       // un-tonemapped rgb (alpha 1)
    float4 history;  // antiAliased, dynamic mask
};

// sobel (taa.comp:582-587) and length(), arguments in the reference's order: row y-1, then the
// left and right texels of row y, then row y+1
__device__ __forceinline__ float sobel_length(const float (&t)[8]) {
    const float gx = t[0] - t[5] + 2.0f * t[1] - 2.0f * t[6] + t[2] - t[7];
    const float gy = t[0] - t[2] + 2.0f * t[3] - 2.0f * t[4] + t[5] - t[7];
    return sqrtf(gx * gx + gy * gy);
}

// calc_segmentation_value (taa.comp:589-702) at texel (x, y), one size. `total` receives the
// weighted indicator sum when the evaluation gets that far (else 0). Normals, luma, material ids
// and depth of the 3x3 neighbourhood come from the staged planes (their apron is clamped to the
// edge like CLAMP_TO_TEX); the previous frame's material id and mask from global memory.
template <int SP>
__device__ uint32_t segmentation_value(const FrameDesc& fd, const SegDesc& sg, const SegPlanes& sp,
                                       const taa_uniforms_t& U, const taa_params_t& P, int x, int y, float hu,
                                       float hv, const float* __restrict__ sd, int l, float& total) {
    total = 0.0f;
    const uint32_t fl = P.mRayTraceAugmentFlags;
    const int W = fd.W, H = fd.H;
    if (fl & TAA_RTFLAG_FXD) {
        const int b = 100;
        if (x < b || y < b || x >= W - b || y >= H - b) return 1u;
    }
    if (fl & TAA_RTFLAG_ALL) return 2u;
    const bool useHistoryCount = (fl & TAA_RTFLAG_CNT) != 0u;
    const uint32_t newCountValue = useHistoryCount ? (uint32_t)P.mRayTraceHistoryCount << 16 : 0u;
    if ((fl & TAA_RTFLAG_OUT) && (hu < 0.0f || hv < 0.0f || hu >= 1.0f || hv >= 1.0f)) return 1u;
    const uint32_t matId = sp.mat[l];
    if (fl & TAA_RTFLAG_DIS) {
        const int tx = to_int(hu * (float)W), ty = to_int(hv * (float)H);  // uv_to_tc(historyUv)
        if (tx >= 0 && ty >= 0 && tx < W && ty < H) {
            const uint32_t prevMatId = gload<uint32_t>(sg.prev_material, (uint32_t)(ty * sg.pitch_l + tx) * 4u);
            if (prevMatId != matId && (prevMatId & 0x80000000u) != 0u) return 2u | newCountValue;
        }
    }
    const int ox[8] = {-1, 0, 1, -1, 1, -1, 0, 1}, oy[8] = {-1, -1, -1, 0, 0, 1, 1, 1};
    float nrmValue = 0.0f, dptValue = 0.0f, matValue = 0.0f, lumValue = 0.0f;
    if (fl & TAA_RTFLAG_NRM) {
        auto normal = [&](int o) { return f3{sp.nx[l + o], sp.ny[l + o], sp.nz[l + o]}; };  // sample_normal
        const f3 nC = normal(0);
        auto dn = [&](int dx, int dy) { const f3 n = normal(dy * SP + dx); return nC.x * n.x + nC.y * n.y + nC.z * n.z; };
        const float dL = dn(-1, 0), dR = dn(1, 0), dT = dn(0, -1), dB = dn(0, 1);
        nrmValue = 1.0f - fmaxf(0.0f, fminf(fminf(fminf(dL, dR), dT), dB));
    }
    if (fl & TAA_RTFLAG_DPT) {
        float t[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) t[k] = linearize_depth(sd[l + oy[k] * SP + ox[k]], U.mCamNearPlane, U.mCamFarPlane);
        dptValue = sobel_length(t);
    }
    if (fl & TAA_RTFLAG_MID) {
        if (matId != sp.mat[l - 1] || matId != sp.mat[l + 1] || matId != sp.mat[l - SP] || matId != sp.mat[l + SP])
            matValue = 1.0f;
    }
    if (fl & TAA_RTFLAG_LUM) {
        float t[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) t[k] = sp.lum[l + oy[k] * SP + ox[k]];
        lumValue = sobel_length(t);
    }
    total = nrmValue * P.mRayTraceAugment_WNrm + dptValue * P.mRayTraceAugment_WDpt + matValue * P.mRayTraceAugment_WMId +
            lumValue * P.mRayTraceAugment_WLum;
    if (total >= P.mRayTraceAugment_Thresh) return 2u | newCountValue;
    if (useHistoryCount) {
        const uint32_t oldCnt = (gload<uint32_t>(sg.prev_segmask, (uint32_t)(y * sg.pitch + x) * 4u) & 0xffff0000u) >> 16;
        if (oldCnt > 0u) return 2u | ((oldCnt - 1u) << 16);
    }
    return 0u;
}

// calc_segmentation_value (taa.comp:589-702) on the upsampling path, at the output texel (x, y) as
// the reference passes it (taa.comp:775). The renderer's images have the input size and are read
// from global memory: the neighbours clamped to the input image (CLAMP_TO_TEX, taa.comp:42), the
// centre's material id and normal at (x, y) itself -- outside the input image for most output
// texels, where Vulkan leaves the load undefined unless the device bounds-checks images: zero
// here, the value robustImageAccess2 defines. The border of TAA_RTFLAG_FXD is measured against the
// input size (textureSize_loRes, taa.comp:595), the previous material read at historyUv in input
// texels (taa.comp:622-624), the previous mask at (x, y) in the output-sized mask.
template <bool S16>
__device__ uint32_t segmentation_value_up(const FrameDesc& fd, const SegDesc& sg, const taa_uniforms_t& U,
                                          const taa_params_t& P, int x, int y, float hu, float hv, float& total) {
    total = 0.0f;
    const uint32_t fl = P.mRayTraceAugmentFlags;
    const int Wl = fd.Wl, Hl = fd.Hl, pl = fd.pitch_l, sp = sg.pitch_l;
    if (fl & TAA_RTFLAG_FXD) {
        const int b = 100;
        if (x < b || y < b || x >= Wl - b || y >= Hl - b) return 1u;
    }
    if (fl & TAA_RTFLAG_ALL) return 2u;
    const bool useHistoryCount = (fl & TAA_RTFLAG_CNT) != 0u;
    const uint32_t newCountValue = useHistoryCount ? (uint32_t)P.mRayTraceHistoryCount << 16 : 0u;
    if ((fl & TAA_RTFLAG_OUT) && (hu < 0.0f || hv < 0.0f || hu >= 1.0f || hv >= 1.0f)) return 1u;
    const bool inside = x < Wl && y < Hl;  // x, y >= 0
    // input texel (xx + dx, yy + dy) clamped to the input image: its offset in texels at row pitch q
    auto cl = [&](int xx, int yy, int q) { return (uint32_t)(clampi(yy, 0, Hl - 1) * q + clampi(xx, 0, Wl - 1)); };
    const uint32_t matId = inside ? gload<uint32_t>(sg.material, (uint32_t)(y * sp + x) * 4u) : 0u;
    if (fl & TAA_RTFLAG_DIS) {
        const int tx = to_int(hu * (float)Wl), ty = to_int(hv * (float)Hl);  // uv_to_tc(historyUv, textureSize_loRes)
        if (tx >= 0 && ty >= 0 && tx < Wl && ty < Hl) {
            const uint32_t prevMatId = gload<uint32_t>(sg.prev_material, (uint32_t)(ty * sp + tx) * 4u);
            if (prevMatId != matId && (prevMatId & 0x80000000u) != 0u) return 2u | newCountValue;
        }
    }
    const int ox[8] = {-1, 0, 1, -1, 1, -1, 0, 1}, oy[8] = {-1, -1, -1, 0, 0, 1, 1, 1};
    float nrmValue = 0.0f, dptValue = 0.0f, matValue = 0.0f, lumValue = 0.0f;
    if (fl & TAA_RTFLAG_NRM) {
        auto normal_of = [&](float4 q) {  // sample_normal (taa.comp:574-579)
            float sz, cz, sw, cw;
            seg_sincos(q.z, sz, cz);
            seg_sincos(q.w, sw, cw);
            return f3{cz * cw, sz * cw, sw};
        };
        const f3 nC = normal_of(inside ? gload<float4>(sg.uv_normal, (uint32_t)(y * sp + x) * 16u) : make_float4(0.f, 0.f, 0.f, 0.f));
        auto dn = [&](int dx, int dy) {
            const f3 n = normal_of(gload<float4>(sg.uv_normal, cl(x + dx, y + dy, sp) * 16u));
            return nC.x * n.x + nC.y * n.y + nC.z * n.z;
        };
        const float dL = dn(-1, 0), dR = dn(1, 0), dT = dn(0, -1), dB = dn(0, 1);
        nrmValue = 1.0f - fmaxf(0.0f, fminf(fminf(fminf(dL, dR), dT), dB));
    }
    if (fl & TAA_RTFLAG_DPT) {
        float t[8];
#pragma unroll
        for (int k = 0; k < 8; ++k)
            t[k] = linearize_depth(ld_f32(fd.depth, cl(x + ox[k], y + oy[k], pl)), U.mCamNearPlane, U.mCamFarPlane);
        dptValue = sobel_length(t);
    }
    if (fl & TAA_RTFLAG_MID) {
        auto mat = [&](int dx, int dy) { return gload<uint32_t>(sg.material, cl(x + dx, y + dy, sp) * 4u); };
        if (matId != mat(-1, 0) || matId != mat(1, 0) || matId != mat(0, -1) || matId != mat(0, 1)) matValue = 1.0f;
    }
    if (fl & TAA_RTFLAG_LUM) {
        float t[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) t[k] = rgb_to_ycocg(xyz(ld_rgba<S16>(fd.color, cl(x + ox[k], y + oy[k], pl)))).x;
        lumValue = sobel_length(t);
    }
    total = nrmValue * P.mRayTraceAugment_WNrm + dptValue * P.mRayTraceAugment_WDpt + matValue * P.mRayTraceAugment_WMId +
            lumValue * P.mRayTraceAugment_WLum;
    if (total >= P.mRayTraceAugment_Thresh) return 2u | newCountValue;
    if (useHistoryCount) {
        const uint32_t oldCnt = (gload<uint32_t>(sg.prev_segmask, (uint32_t)(y * sg.pitch + x) * 4u) & 0xffff0000u) >> 16;
        if (oldCnt > 0u) return 2u | ((oldCnt - 1u) << 16);
    }
    return 0u;
}

// One output texel of taa.comp main (taa.comp:707-960), FP32 arithmetic. `l` indexes the texel
// in the staged LDS planes (row pitch SP). SEG: the segmentation mask's images are in `sg`.
// FA: the FP16 path's arithmetic (TAA_ARITH_FP16) where the bits of a selection do not depend on
// it -- fp32 moments with fused multiply-adds and the hardware square root, the Catmull-Rom and
// bilinear history filters with fp16 texels in mixed FMAs, one reciprocal for the clip, Lottes and
// the near-clamp ratio, the blend as one FMA per channel; every texel index, tap choice and
// rejection decision stays that of the FP32 path (used by the upsampling resolve's FP16 form)
template <bool S16, int SP, bool UP, bool SEG = false, bool STG = false, bool VEL16 = false, bool FA = false, int UCW = 1>
__device__ __forceinline__ PixelResult resolve_pixel_f32(const FrameDesc& fd, const taa_uniforms_t& U, const taa_params_t& P,
                                         int x, int y, int l, const float* __restrict__ s0,
                                         const float* __restrict__ s1, const float* __restrict__ s2,
                                         const float* __restrict__ sd, float uvx, float uvy, bool write_aux,
                                         const SegDesc* sg = nullptr, const SegPlanes* spl = nullptr,
                                         const UpIn* up = nullptr, int rx = 0, int ry = 0,
                                         const uint32_t* __restrict__ svel = nullptr) {
    // uvx, uvy: tc_to_uv(x, y) = ((x, y) + 0.5) / size (taa.comp:131), tabulated per tile
    const int W = fd.W, H = fd.H;
    const uint32_t gi = (uint32_t)(y * fd.pitch + x);
    PixelResult R;
    // input (render) resolution: iuv_lores = int(uv * input size) (taa.comp:724); == (x, y) at one size
    const int Wl = fd.Wl, Hl = fd.Hl, pl = fd.pitch_l;
    int xl = x, yl = y;
    if constexpr (UP) {
        xl = to_int(uvx * (float)Wl);
        yl = to_int(uvy * (float)Hl);
    }
    auto lo = [&](int xx, int yy) { return (uint32_t)(clampi(yy, 0, Hl - 1) * pl + clampi(xx, 0, Wl - 1)); };
    // upsampling path: input texel (xx, yy) (clamped coordinates) from the staged rectangle (STG: the
    // tile staged every input texel its resolve reads; row pitch UCW), else from global memory --
    // the same values either way. The staged rectangle holds at position (sx, sy) the texel
    // (clamp(x0 + sx), clamp(y0 + sy)), so a position relative to it needs no clamp.
    // (the row offset is below UCW * UCH: a 24-bit multiply, full rate where a 32-bit one is quarter rate)
    auto staged = [&](int xx, int yy) { return __mul24(yy - up->y0, UCW) + (xx - up->x0); };
    auto color_at = [&](int xx, int yy) -> f3 {
        if constexpr (STG) {
            const int k = staged(xx, yy);
            return f3{up->c0[k], up->c1[k], up->c2[k]};
        } else {
            return stage_convert(P, xyz(ld_rgba<S16>(fd.color, (uint32_t)(yy * pl + xx))));
        }
    };
    auto depth_at = [&](int xx, int yy) -> float {
        if constexpr (STG) return up->d[staged(xx, yy)];
        else return ld_f32(fd.depth, (uint32_t)(yy * pl + xx));
    };
    // the taps of the bilinear depth and velocity samples (non-staged tiles): the clamped texels
    // STG: the output texel's own input texel (xl, yl) in the staged rectangle
    const int sbase = STG ? staged(xl, yl) : 0;
    // velocity texel (x + dx, y + dy) clamped to the image (global index gidx), dx, dy in -1..1:
    // from the staged RG16F region when the tile staged it (svel), else from global memory
    auto vel_tap = [&](int dx, int dy, uint32_t gidx) -> float4 {
        if constexpr (VEL16) {
            if (svel) {
                const uint32_t v = svel[l + dy * SP + dx];
                return make_float4(h2f(v & 0xffffu), h2f(v >> 16), 0.f, 1.f);  // as ld_vel reads RG16F
            }
        }
        return ld_vel(fd.velocity, fd.vel_fmt, gidx);
    };
    // the bilinear taps of a depth or velocity sample at uv + (dx, dy) input texels, dx, dy in -1..1
    // (SAMPLE_TEX at uv + toU * offset), from the tile's tables
    auto up_taps = [&](int dx, int dy) {
        return bilinear_taps(up->ax[(dx + 1) * up->rw + rx], up->ay[(dy + 1) * up->rh + ry]);
    };
    // bilinear depth and velocity samples at uv + (dx, dy) input texels. STG: a staged tile's tables
    // hold the first tap's position p = (x - x0) + (y - y0) * UCW before the edge clamp, the other
    // taps at p + 1, p + UCW, p + UCW + 1 -- the rectangle holds the clamped texel at every
    // unclamped position, so these are the four texels of the clamped form
    auto up_depth = [&](int dx, int dy) -> float {
        const BilinearTaps t = up_taps(dx, dy);
        if constexpr (STG) {
            const float* q = up->d + (t.ya + t.xa);
            return bilinear1_quad(q[0], q[1], q[UCW], q[UCW + 1], t);
        } else {
            return bilinear1_at(t, [&](int xx, int yy) { return ld_f32(fd.depth, (uint32_t)(yy * pl + xx)); });
        }
    };
    // the nine depth samples of the closest-depth search in scan order. STG, when the three rows'
    // first taps are consecutive rows of the rectangle (everywhere but at a top edge, and where
    // the float position lands within rounding of a texel boundary): the horizontal lerps of the
    // four rows are shared by the samples above and below them -- the same operations on the same
    // values, twelve lerps instead of eighteen
    auto up_depth9 = [&](float (&ds)[9]) {
        if constexpr (STG) {
            const BilinearAxis y0 = up->ay[ry], y1 = up->ay[up->rh + ry], y2 = up->ay[2 * up->rh + ry];
            if (y1.a == y0.a + UCW && y2.a == y1.a + UCW) {
#pragma unroll
                for (int qx = 0; qx < 3; ++qx) {
                    const BilinearAxis x = up->ax[qx * up->rw + rx];
                    float h[4];
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const float* q = up->d + (y0.a + r * UCW + x.a);
                        h[r] = q[0] * x.u + q[1] * x.t;
                    }
                    ds[qx] = h[0] * y0.u + h[1] * y0.t;
                    ds[3 + qx] = h[1] * y1.u + h[2] * y1.t;
                    ds[6 + qx] = h[2] * y2.u + h[3] * y2.t;
                }
                return;
            }
        }
#pragma unroll
        for (int k = 0; k < 9; ++k) ds[k] = up_depth(k % 3 - 1, k / 3 - 1);
    };
    auto up_vel = [&](int dx, int dy) -> float4 {
        const BilinearTaps t = up_taps(dx, dy);
        if constexpr (STG) {
            const float4* q = up->v + (t.ya + t.xa);
            return bilinear4_quad(q[0], q[1], q[UCW], q[UCW + 1], t);
        } else {
            return bilinear4_at(t, [&](int xx, int yy) { return ld_vel(fd.velocity, fd.vel_fmt, (uint32_t)(yy * pl + xx)); });
        }
    };
    const float toUx = 1.0f / (float)Wl, toUy = 1.0f / (float)Hl;  // 1 / textureSize(depth, velocity)

    if (P.mPassThrough || U.mBypassHistoryUpdate) {  // taa.comp:726-737
        float4 h = ld_rgba<S16>(fd.history, gi);
        f3 s = P.mPassThrough ? xyz(ld_rgba<S16>(fd.color, lo(xl, yl))) : xyz(h);
        R.screen = s;
        R.history = make_float4(h.x, h.y, h.z, 1.0f);
        if (write_aux && fd.debug) st_rgba16(fd.debug, gi, make_float4(0.f, 0.f, 0.f, 0.f));
        if (write_aux && fd.inspect) {
            int32_t w[TAA_INSPECT_WORDS];
#pragma unroll
            for (int k = 0; k < TAA_INSPECT_WORDS; ++k) w[k] = (k == TAA_INSPECT_CLOSEST_TAP) ? -1 : 0;
            st_inspect(fd.inspect, gi, w);
        }
        return R;
    }

    // ---- neighbourhood (taa.comp:199-213, 259-319)
    const float jux = U.mJitterNdc[0] * 0.5f * P.mUnjitterFactor, juy = U.mJitterNdc[1] * 0.5f * P.mUnjitterFactor;
    f3 c[9];  // centre, then (-1,-1) (0,-1) (1,-1) (-1,0) (1,0) (-1,1) (0,1) (1,1)
    const int ox[9] = {0, -1, 0, 1, -1, 1, -1, 0, 1};
    const int oy[9] = {0, -1, -1, -1, 0, 0, 1, 1, 1};
    if (!P.mUnjitterNeighbourhood) {
#pragma unroll
        for (int k = 0; k < 9; ++k) {
            if constexpr (STG) {
                const int li = sbase + oy[k] * UCW + ox[k];
                c[k] = f3{up->c0[li], up->c1[li], up->c2[li]};
            } else if constexpr (UP) {
                c[k] = color_at(clampi(xl + ox[k], 0, Wl - 1), clampi(yl + oy[k], 0, Hl - 1));
            } else {
                int li = l + oy[k] * SP + ox[k];
                c[k] = {s0[li], s1[li], s2[li]};
            }
        }
    } else {
        const float ix = 1.0f / (float)Wl, iy = 1.0f / (float)Hl;
        for (int k = 0; k < 9; ++k) {
            float u = jux + ((float)(xl + ox[k]) + 0.5f) * ix, v = juy + ((float)(yl + oy[k]) + 0.5f) * iy;
            c[k] = stage_convert(P, xyz(bilinear_rgba<S16>(fd.color, Wl, Hl, pl, u, v)));
        }
    }
    f3 colMin, colMax, clipTowards;
    if (FA && P.mVarianceClipping) {
        f3 m1 = c[0], m2 = c[0] * c[0];
#pragma unroll
        for (int k = 1; k < 9; ++k) {
            m1 = m1 + c[k];
            m2 = {__builtin_fmaf(c[k].x, c[k].x, m2.x), __builtin_fmaf(c[k].y, c[k].y, m2.y), __builtin_fmaf(c[k].z, c[k].z, m2.z)};
        }
        constexpr float inv9 = 1.0f / 9.0f;
        const f3 mean = m1 * inv9;
        const f3 sigma = {sqrt_fast(fmaxf(__builtin_fmaf(-mean.x, mean.x, m2.x * inv9), 0.0f)),
                          sqrt_fast(fmaxf(__builtin_fmaf(-mean.y, mean.y, m2.y * inv9), 0.0f)),
                          sqrt_fast(fmaxf(__builtin_fmaf(-mean.z, mean.z, m2.z * inv9), 0.0f))};
        const float g = P.mVarClipGamma;
        colMin = {__builtin_fmaf(-g, sigma.x, mean.x), __builtin_fmaf(-g, sigma.y, mean.y), __builtin_fmaf(-g, sigma.z, mean.z)};
        colMax = {__builtin_fmaf(g, sigma.x, mean.x), __builtin_fmaf(g, sigma.y, mean.y), __builtin_fmaf(g, sigma.z, mean.z)};
        clipTowards = mean;
    } else if (P.mVarianceClipping) {

        f3 m1 = c[0] + c[1] + c[2] + c[3] + c[4] + c[5] + c[6] + c[7] + c[8];
        f3 m2 = c[0] * c[0] + c[1] * c[1] + c[2] * c[2] + c[3] * c[3] + c[4] * c[4] + c[5] * c[5] + c[6] * c[6] +
                c[7] * c[7] + c[8] * c[8];
        f3 mean = {div9(m1.x), div9(m1.y), div9(m1.z)};  // m1 / 9
        f3 var = f3{div9(m2.x), div9(m2.y), div9(m2.z)} - mean * mean;
        f3 sigma = {sqrtf(fmaxf(0.0f, var.x)), sqrtf(fmaxf(0.0f, var.y)), sqrtf(fmaxf(0.0f, var.z))};
        colMin = mean - P.mVarClipGamma * sigma;
        colMax = mean + P.mVarClipGamma * sigma;
        clipTowards = mean;
    } else if (P.mShapedNeighbourhood) {
        f3 mn33 = vmin(vmin(vmin(vmin(vmin(vmin(vmin(vmin(c[0], c[1]), c[2]), c[3]), c[4]), c[5]), c[6]), c[7]), c[8]);
        f3 mx33 = vmax(vmax(vmax(vmax(vmax(vmax(vmax(vmax(c[0], c[1]), c[2]), c[3]), c[4]), c[5]), c[6]), c[7]), c[8]);
        f3 mn5 = vmin(vmin(vmin(vmin(c[0], c[2]), c[4]), c[5]), c[7]);
        f3 mx5 = vmax(vmax(vmax(vmax(c[0], c[2]), c[4]), c[5]), c[7]);
        colMin = (mn33 + mn5) * 0.5f;
        colMax = (mx33 + mx5) * 0.5f;
        clipTowards = c[0];
    } else {
        colMin = vmin(vmin(vmin(vmin(vmin(vmin(vmin(vmin(c[0], c[1]), c[2]), c[3]), c[4]), c[5]), c[6]), c[7]), c[8]);
        colMax = vmax(vmax(vmax(vmax(vmax(vmax(vmax(vmax(c[0], c[1]), c[2]), c[3]), c[4]), c[5]), c[6]), c[7]), c[8]);
        clipTowards = c[0];
    }
    if (P.mUseYCoCg && P.mShrinkChromaAxis) {
        const f3 s = {0.5f * 1.0f, 0.5f * 0.5f, 0.5f * 0.5f};
        f3 halfSize = s * (colMax - colMin);
        f3 center = (colMin + colMax) * 0.5f;
        colMin = center - halfSize;
        colMax = center + halfSize;
        bool out = clipTowards.x < colMin.x || clipTowards.y < colMin.y || clipTowards.z < colMin.z ||
                   clipTowards.x > colMax.x || clipTowards.y > colMax.y || clipTowards.z > colMax.z;
        if (out) clipTowards = center;
    }
    // getCurrentColor (taa.comp:216-220), or getCurrentUpsampledColor (taa.comp:222-257, 750-755)
    f3 currentColor = c[0];
    float beta = 1.0f;
    bool upsampled = false;
    if constexpr (UP) {
        if (U.mUpsampling) {
            const int ix = up->cx[rx], iy = up->cy[ry];
            if (ix >= 0 && iy >= 0) {
                beta = 1.0f;
                currentColor = color_at(ix, iy);
            } else {
                beta = 0.0f;
                currentColor = f3{0.0f, 0.0f, 0.0f};
            }
            upsampled = true;
        }
    }
    if (!upsampled) {
        if (P.mUnjitterCurrentSample) {
            float u = jux + ((float)xl + 0.5f) * (1.0f / (float)Wl), v = juy + ((float)yl + 0.5f) * (1.0f / (float)Hl);
            currentColor = stage_convert(P, xyz(bilinear_rgba<S16>(fd.color, Wl, Hl, pl, u, v)));
        } else if (P.mUnjitterNeighbourhood) {
            // c[0] is the unjittered neighbourhood centre; getCurrentColor samples the texel itself
            if constexpr (STG) currentColor = f3{up->c0[sbase], up->c1[sbase], up->c2[sbase]};
            else if constexpr (UP) currentColor = color_at(clampi(xl, 0, Wl - 1), clampi(yl, 0, Hl - 1));
            else currentColor = {s0[l], s1[l], s2[l]};
        }
    }
    float depth;  // taa.comp:756
    if constexpr (STG) depth = up->d[sbase];
    else if constexpr (UP) depth = depth_at(clampi(xl, 0, Wl - 1), clampi(yl, 0, Hl - 1));
    else depth = sd[l];

    // ---- getHistoryPosition (taa.comp:391-438)
    int flags = 0, closest = -1;
    float hu, hv, expectedHistoryDepth;
    {
        // SAMPLE_TEX(uCurrentVelocity, uv): a texel fetch at one size, bilinear when upsampling
        float4 vs;
        if constexpr (UP) vs = up_vel(0, 0);
        else vs = vel_tap(0, 0, gi);
        bool canUseVelocity = !(P.mUseVelocityVectors == 0 || (P.mUseVelocityVectors == 1 && vs.w < 0.5f));
        if (canUseVelocity) {
            flags |= TAA_FLAG_USED_VELOCITY;
            if (P.mVelocitySampleMode == 1) {
                // taa.comp:398-412: the longest vector is tracked but the last tap is assigned
                float4 last;
                if constexpr (UP) last = up_vel(1, 1);
                else last = vel_tap(1, 1, (uint32_t)(min(y + 1, H - 1) * fd.pitch + min(x + 1, W - 1)));
                vs.x = last.x;
                vs.y = last.y;
            } else if (P.mVelocitySampleMode == 2) {
                // findClosestUvAndZ_3x3 (taa.comp:371-389): scan order row-major, strict <
                int best = 0;
                if constexpr (UP) {  // depth taps at the output uv, interpolated
                    float ds[9];
                    up_depth9(ds);
                    float dC = ds[0];
#pragma unroll
                    for (int k = 1; k < 9; ++k) {
                        if (ds[k] < dC) { dC = ds[k]; best = k; }
                    }
                    vs = up_vel(best % 3 - 1, best / 3 - 1);
                } else {
                    float dC = sd[l - SP - 1];
#pragma unroll
                    for (int k = 1; k < 9; ++k) {
                        float d = sd[l + (k / 3 - 1) * SP + (k % 3 - 1)];
                        if (d < dC) { dC = d; best = k; }
                    }
                    int vx = clampi(x + best % 3 - 1, 0, W - 1), vy = clampi(y + best / 3 - 1, 0, H - 1);
                    vs = vel_tap(best % 3 - 1, best / 3 - 1, (uint32_t)(vy * fd.pitch + vx));
                }
                closest = best;
            }
            hu = uvx - vs.x;
            hv = uvy - vs.y;
            expectedHistoryDepth = depth - vs.z;
        } else {
            const float* Mi = U.mInverseViewProjMatrix;
            const float* Mh = U.mHistoryViewProjMatrix;
            float cs[4] = {uvx * 2.0f - 1.0f, uvy * 2.0f - 1.0f, depth, 1.0f};
            float w[4], h[4];
#pragma unroll
            for (int r = 0; r < 4; ++r) w[r] = Mi[r] * cs[0] + Mi[4 + r] * cs[1] + Mi[8 + r] * cs[2] + Mi[12 + r] * cs[3];
#pragma unroll
            for (int r = 0; r < 4; ++r) h[r] = Mh[r] * w[0] + Mh[4 + r] * w[1] + Mh[8 + r] * w[2] + Mh[12 + r] * w[3];
            hu = (h[0] / h[3]) * 0.5f + 0.5f;
            hv = (h[1] / h[3]) * 0.5f + 0.5f;
            expectedHistoryDepth = h[2] / h[3];
        }
    }
    const float vux = uvx - hu, vuy = uvy - hv;
    const float pixelSpeed = FA ? sqrt_fast(vux * vux + vuy * vuy) : sqrtf(vux * vux + vuy * vuy);

    // the previous depth under the history position (taa.comp:818), fetched here, ahead of the
    // history gather, so that the two loads overlap; the test itself keeps its place below
    int dtx = 0, dty = 0;
    float historyDepth = 0.0f;
    if (P.mDepthCulling) {
        dtx = to_int(hu * (float)Wl);  // uHistoryDepth has the input size (taa.comp:818)
        dty = to_int(hv * (float)Hl);
        historyDepth = ld_f32(fd.hdepth, lo(dtx, dty));
    }

    // ---- sample_history_rgba (taa.comp:545-549)
    float4 historyRaw;
    if (FA && P.mInterpolationMode != 1) {
        // the base texel floor(uv * size - 0.5) of the FP32 path (taa.comp:457-458); the weights from its fraction
        const float tX = hu * (float)W - 0.5f, tY = hv * (float)H - 0.5f;
        const float fbx = floorf(tX), fby = floorf(tY);
        const int kx = to_int(fbx), ky = to_int(fby);
        const Hist h = P.mInterpolationMode == 0 ? history_bilinear_fast<S16>(fd.history, W, H, fd.pitch, kx, ky, tX - fbx, tY - fby)
                       : P.mDynamicAntiGhosting ? history_cr_fast<S16, true>(fd.history, W, H, fd.pitch, kx, ky, tX - fbx, tY - fby)
                                                : history_cr_fast<S16, false>(fd.history, W, H, fd.pitch, kx, ky, tX - fbx, tY - fby);
        historyRaw = make_float4(h.c.x, h.c.y, h.c.z, h.a);
    } else if (P.mInterpolationMode == 0) historyRaw = bilinear_rgba<S16>(fd.history, W, H, fd.pitch, hu, hv);
    else if (P.mInterpolationMode == 1) historyRaw = history_bspline<S16>(fd, hu, hv);
    else if (P.mDynamicAntiGhosting) historyRaw = history_catmullrom<S16, true>(fd, hu, hv);
    else historyRaw = history_catmullrom<S16, false>(fd, hu, hv);
    f3 historyColor = xyz(historyRaw);
    if (P.mUseYCoCg) historyColor = rgb_to_ycocg(historyColor);
    float alpha = P.mAlpha;
    bool rejected = false;

    // ---- segmentation mask (taa.comp:773-784, 959)
    uint32_t segMaskValue = 0u;
    float segTotal = 0.0f;
    if constexpr (SEG) {
        if (P.mRayTraceAugment) {
            if constexpr (UP) segMaskValue = segmentation_value_up<S16>(fd, *sg, U, P, x, y, hu, hv, segTotal);
            else segMaskValue = segmentation_value<SP>(fd, *sg, *spl, U, P, x, y, hu, hv, sd, l, segTotal);
            if ((segMaskValue & 3u) != 0u) {
                alpha = P.mRejectionAlpha;
                rejected = true;
            }
            if (write_aux) gstore<uint32_t>(sg->segmask, (uint32_t)(y * sg->pitch + x) * 4u, segMaskValue);
        }
    }

    // ---- rejection (taa.comp:787-823)
    if (P.mRejectOutside) {
        if (hu < 0.0f || hv < 0.0f || hu >= 1.0f || hv >= 1.0f) {
            alpha = P.mRejectionAlpha;
            rejected = true;
            flags |= TAA_FLAG_OUTSIDE;
        }
    }
    float writeDynamicMask = 0.0f;
    if (P.mDynamicAntiGhosting) {
        const float eps = 1e-5f;
        auto mov = [&](int dx, int dy) {
            float4 v;
            if constexpr (UP) v = up_vel(dx, dy);
            else v = vel_tap(dx, dy, (uint32_t)(clampi(y + dy, 0, H - 1) * fd.pitch + clampi(x + dx, 0, W - 1)));
            return (fabsf(v.x) > eps || fabsf(v.y) > eps) && (fabsf(v.w) >= 0.5f);
        };
        bool movL = mov(-1, 0), movR = mov(1, 0), movT = mov(0, -1), movB = mov(0, 1), movC = mov(0, 0);
        if (!(movL || movR || movT || movB || movC) && historyRaw.w > 0.0f) {
            rejected = true;
            flags |= TAA_FLAG_GHOST_REJECT;
        }
        writeDynamicMask = movC ? 1.0f : 0.0f;
    }
    if (P.mDepthCulling) {
        bool rej;
        if (U.pad1 > 0.0f) {
            float n = U.mCamNearPlane, f = U.mCamFarPlane;
            rej = FA ? linear_depth_reject(historyDepth, expectedHistoryDepth, n, f, U.pad1)
                     : fabsf(linearize_depth(historyDepth, n, f) - linearize_depth(expectedHistoryDepth, n, f)) > U.pad1;
        } else {
            float depthEpsilon = 0.1f * (1.0f - historyDepth);
            rej = fabsf(historyDepth - expectedHistoryDepth) > depthEpsilon;
        }
        if (rej) {
            rejected = true;
            flags |= TAA_FLAG_DEPTH_REJECT;
        }
    }

    // ---- rectification (taa.comp:826-845)
    const f3 origHistoryColor = historyColor;
    if (P.mColorClampingOrClipping == 1) historyColor = vmin(vmax(historyColor, colMin), colMax);
    else if (FA && P.mColorClampingOrClipping == 2) historyColor = clip_aabb_fast(colMin, colMax, historyColor);
    else if (P.mColorClampingOrClipping == 2) historyColor = clipAabb(colMin, colMax, historyColor);
    else if (P.mColorClampingOrClipping == 3) historyColor = clipAabbSlow(colMin, colMax, clipTowards, historyColor);
    const f3 rdiff = historyColor - origHistoryColor;
    const bool rectified = fabsf(rdiff.x) > 0.001f || fabsf(rdiff.y) > 0.001f || fabsf(rdiff.z) > 0.001f;

    // ---- blending (taa.comp:850-900)
    if (rejected) {
        alpha = P.mRejectionAlpha;
        beta = 1.0f;
    } else {
        if (P.mVelBasedAlpha)
            alpha = fmaxf(alpha, gmix(alpha, P.mVelBasedAlphaMax, gclamp(pixelSpeed * P.mVelBasedAlphaFactor, 0.0f, 1.0f)));
        if (P.mLumaWeightingLottes) {
            float lc = luminance(P, currentColor), lh = luminance(P, historyColor);
            float diff = FA ? fabsf(lc - lh) * rcp_fast(fmaxf(fmaxf(lc, lh), 0.2f)) : fabsf(lc - lh) / fmaxf(fmaxf(lc, lh), 0.2f);
            float w = 1.0f - diff;
            float ww = w * w;
            alpha = gmix(P.mMaxAlpha, P.mMinAlpha, ww);
        }
        if (P.mReduceBlendNearClamp) {
            float mnl = luminance(P, colMin), mxl = luminance(P, colMax), hl = luminance(P, origHistoryColor);
            float distToClamp = FA ? 2.0f * fabsf(fminf(hl - mnl, mxl - hl)) * rcp_fast(mxl - mnl)
                                   : 2.0f * fabsf(fminf(hl - mnl, mxl - hl)) / (mxl - mnl);
            if (mxl - mnl < 0.001f) distToClamp = 1.0f;
            alpha *= gclamp(4.0f * distToClamp, 0.0f, 1.0f);
        }
    }
    if (U.mResetHistory) { alpha = 1.0f; beta = 1.0f; }
    const float ab = alpha * beta;
    f3 aa = {gmix(historyColor.x, currentColor.x, ab), gmix(historyColor.y, currentColor.y, ab), gmix(historyColor.z, currentColor.z, ab)};
    if (FA)
        aa = {__builtin_fmaf(ab, currentColor.x - historyColor.x, historyColor.x),
              __builtin_fmaf(ab, currentColor.y - historyColor.y, historyColor.y),
              __builtin_fmaf(ab, currentColor.z - historyColor.z, historyColor.z)};
    if (P.mUseYCoCg) aa = ycocg_to_rgb(aa);
    if (P.mAddNoise) {  // taa.comp:551-556 (GLSL sin(): outside the parity contract)
        float sx = uvx + U.mSinTime[0] + 0.6959174f, sy = uvy + U.mSinTime[0] + 0.6959174f;
        float s = sinf(sx * 12.9898f + sy * 78.233f);
        float n0 = s * 43758.5453f, n1 = s * 28001.8384f, n2 = s * 50849.4141f;
        aa.x = aa.x + ((n0 - floorf(n0)) * 2.0f - 1.0f) * P.mNoiseFactor;
        aa.y = aa.y + ((n1 - floorf(n1)) * 2.0f - 1.0f) * P.mNoiseFactor;
        aa.z = aa.z + ((n2 - floorf(n2)) * 2.0f - 1.0f) * P.mNoiseFactor;
    }
    R.history = make_float4(aa.x, aa.y, aa.z, writeDynamicMask);
    R.screen = un_tonemap_rgb(P, aa);

    if (write_aux && fd.debug) {  // taa.comp:912-942
        float4 g = make_float4(0.f, 0.f, 0.f, 0.f);
        switch (P.mDebugMode) {
            case 0: { f3 t = colMax - colMin; g = make_float4(t.x, t.y, t.z, 0.f); break; }
            case 1: { f3 t = colMax - colMin; float p = t.x * t.y * t.z; g = make_float4(p, p, p, 0.f); break; }
            case 2: g = make_float4(rejected ? 1.f : 0.f, sqrtf(rdiff.x * rdiff.x + rdiff.y * rdiff.y + rdiff.z * rdiff.z), 0.f, 0.f); break;
            case 3: g = make_float4(alpha, alpha, alpha, 0.f); break;
            case 4: g = ld_vel(fd.velocity, fd.vel_fmt, lo(x, y)); break;
            case 5: g = make_float4(pixelSpeed, 0.f, 0.f, 0.f); break;
            case 6: g = make_float4(R.screen.x, R.screen.y, R.screen.z, 1.0f); break;
            case 7: g = R.history; break;
            case 8:  // the mask category (taa.comp:932-938); category 3 does not occur
                if ((segMaskValue & 3u) == 0u) g = make_float4(0.f, 0.f, 1.f, 0.f);
                else if ((segMaskValue & 3u) == 1u) g = make_float4(1.f, 0.f, 0.f, 0.f);
                else if ((segMaskValue & 3u) == 2u) g = make_float4(1.f, 1.f, 0.f, 0.f);
                break;
            default: break;
        }
        const float ds = P.mDebugScale;
        g = make_float4(g.x * ds * P.mDebugMask[0], g.y * ds * P.mDebugMask[1], g.z * ds * P.mDebugMask[2], g.w * ds * P.mDebugMask[3]);
        if (P.mDebugCenter) g = make_float4(g.x * 0.5f + 0.5f, g.y * 0.5f + 0.5f, g.z * 0.5f + 0.5f, g.w * 0.5f + 0.5f);
        st_rgba16(fd.debug, gi, g);
    }
    if (write_aux && fd.inspect) {
        if (rejected) flags |= TAA_FLAG_REJECTED;
        if (rectified) flags |= TAA_FLAG_RECTIFIED;
        int32_t w[TAA_INSPECT_WORDS];
        w[TAA_INSPECT_HIST_TC_X] = to_int(floorf(hu * (float)W - 0.5f));
        w[TAA_INSPECT_HIST_TC_Y] = to_int(floorf(hv * (float)H - 0.5f));
        w[TAA_INSPECT_CLOSEST_TAP] = closest;
        w[TAA_INSPECT_FLAGS] = flags;
        w[TAA_INSPECT_COLMIN + 0] = __float_as_int(colMin.x);
        w[TAA_INSPECT_COLMIN + 1] = __float_as_int(colMin.y);
        w[TAA_INSPECT_COLMIN + 2] = __float_as_int(colMin.z);
        w[TAA_INSPECT_COLMAX + 0] = __float_as_int(colMax.x);
        w[TAA_INSPECT_COLMAX + 1] = __float_as_int(colMax.y);
        w[TAA_INSPECT_COLMAX + 2] = __float_as_int(colMax.z);
        w[TAA_INSPECT_BLEND] = __float_as_int(ab);
        w[TAA_INSPECT_DEPTH_TC_X] = dtx;
        w[TAA_INSPECT_DEPTH_TC_Y] = dty;
        w[TAA_INSPECT_HIST_UV_X] = __float_as_int(hu);
        w[TAA_INSPECT_HIST_UV_Y] = __float_as_int(hv);
        w[TAA_INSPECT_SEG_TOTAL] = __float_as_int(segTotal);
        st_inspect(fd.inspect, gi, w);
    }
    return R;
}

// ------------------------------------------------------------------------------------------
// one tile

// The production feature set of the exact kernel (variance clip box, clip toward its centre, YCoCg,
// Catmull-Rom history, closest-depth velocity from an RG image of the colour's precision, depth
// culling, Lottes luma weighting, outside rejection; no debug or inspect image) with its switches
// as constants in local copies of the parameters, so that the compiler drops every other branch.
// The arithmetic is the general kernel's, operation for operation. exact_prod_form() decides on the
// host which launches take it.
struct ProdFix {
    FrameDesc fd;
    taa_uniforms_t U;
    taa_params_t P;
};
__device__ __forceinline__ void prod_fix(ProdFix& f, const FrameDesc& fd, const taa_uniforms_t& U, bool s16, bool seg) {
    f.fd = fd;
    f.fd.debug = nullptr;
    f.fd.inspect = nullptr;
    f.fd.vel_fmt = s16 ? TAA_VEL_RG16F : TAA_VEL_RG32F;
    f.U = U;
    f.U.mBypassHistoryUpdate = 0;
    f.P = U.param[fd.param_idx];
    taa_params_t& p = f.P;
    p.mPassThrough = 0;
    p.mUnjitterNeighbourhood = 0;
    p.mUnjitterCurrentSample = 0;
    p.mVarianceClipping = 1;
    p.mShapedNeighbourhood = 0;
    p.mUseYCoCg = 1;
    p.mShrinkChromaAxis = 0;
    p.mUseVelocityVectors = 2;
    p.mVelocitySampleMode = 2;
    p.mDepthCulling = 1;
    p.mInterpolationMode = 2;
    p.mDynamicAntiGhosting = 0;
    p.mRayTraceAugment = seg ? 1 : 0;  // the masked resolve keeps the mask's own parameters
    p.mColorClampingOrClipping = 2;
    p.mVelBasedAlpha = 0;
    p.mLumaWeightingLottes = 1;
    p.mReduceBlendNearClamp = 0;
    p.mRejectOutside = 1;
    p.mAddNoise = 0;
    p.mToneMapLumaKaris = 0;
}

template <class G, bool S16, bool SEG = G::SEG, bool FA = false, bool PROD = false>
__device__ __forceinline__ void resolve_tile(const FrameDesc& fd_in, const taa_uniforms_t& U_in, int tile, typename G::Smem& sm,
                                             const SegDesc* sg = nullptr) {
    ProdFix pf;
    if constexpr (PROD) prod_fix(pf, fd_in, U_in, S16, SEG);
    const FrameDesc& fd = PROD ? pf.fd : fd_in;
    const taa_uniforms_t& U = PROD ? pf.U : U_in;
    const taa_params_t& P = PROD ? pf.P : U.param[fd.param_idx];
    const int W = fd.W, H = fd.H, pitch = fd.pitch;
    const int ty = tile / fd.tiles_x, tx = tile - ty * fd.tiles_x;
    const int x0 = fd.x_begin + tx * G::TW, y0 = ty * G::TH;
    const int gx0 = x0 - G::A, gy0 = y0 - G::A;  // image coordinates of LDS texel (0,0)
    const int tid = threadIdx.x;

    // 1. stage colour (converted once per texel) and depth, clamped to the edge (one size only:
    // the upsampling path reads its inputs where the output texel maps to)
    const bool vel_staged = G::VEL16 && fd.vel_fmt == TAA_VEL_RG16F;
    for (int i = tid; i < (G::UP ? 0 : G::SW * G::SH); i += G::NT) {
        const int sy = i / G::SW, sx = i - sy * G::SW;
        const int gx = clampi(gx0 + sx, 0, W - 1), gy = clampi(gy0 + sy, 0, H - 1);
        const uint32_t gi = (uint32_t)(gy * pitch + gx);
        const f3 raw = xyz(ld_rgba<S16>(fd.color, gi));
        f3 t = stage_convert(P, raw);
        const int li = sy * G::SP + sx;
        sm.c0[li] = t.x;
        sm.c1[li] = t.y;
        sm.c2[li] = t.z;
        sm.d[li] = ld_f32(fd.depth, gi);
        if constexpr (G::VEL16)
            if (vel_staged) sm.vel[li] = gload<uint32_t>(fd.velocity, gi * 4u);
        if constexpr (G::SEG) {  // sample_normal, sample_luminance and the material id (taa.comp:570-579)
            const uint32_t si = (uint32_t)(gy * sg->pitch_l + gx);
            const float4 q = gload<float4>(sg->uv_normal, si * 16u);
            float sz, cz, sw, cw;
            seg_sincos(q.z, sz, cz);
            seg_sincos(q.w, sw, cw);
            sm.nx[li] = cz * cw;
            sm.ny[li] = sz * cw;
            sm.nz[li] = sw;
            sm.lum[li] = rgb_to_ycocg(raw).x;
            sm.mat[li] = gload<uint32_t>(sg->material, si * 4u);
        }
    }
    // upsampling path: stage the input texels under the resolved region. xl = int(uv * input size)
    // grows with the output column (uv as tabulated below), so the region's first and last
    // columns bound it. With an input no larger than the output, every fetch of a resolved texel
    // lies within 2 input texels of its (xl, yl): the neighbourhood at +-1, the bilinear taps of
    // depth and velocity at uv + (-1..1) input texels (floor(uv * size - 0.5) + {-1, 0, 1, 2} around
    // xl, the float error of uv * size being far below half a texel), the upsampled sample inside
    // the output texel's footprint (+-1); the rectangle keeps 3. An input larger than the output
    // stages nothing and every fetch reads global memory.
    UpIn ui{sm.uc0, sm.uc1, sm.uc2, sm.ud, sm.uvel, 0, 0, 0, 0, G::UCW,
            sm.uax, sm.uay, sm.ucx, sm.ucy, G::RW, G::RH};
    if constexpr (G::UP) {
        const int Wl = fd.Wl, Hl = fd.Hl, pl = fd.pitch_l;
        // the staged rectangle: the input texels under the resolve region, +3 on every side
        const int cx_lo = clampi(x0 - G::RA, 0, W - 1), cx_hi = clampi(x0 + G::TW + G::RA - 1, 0, W - 1);
        const int cy_lo = clampi(y0 - G::RA, 0, H - 1), cy_hi = clampi(y0 + G::TH + G::RA - 1, 0, H - 1);
        const int xl_lo = to_int(((float)cx_lo + 0.5f) / (float)W * (float)Wl);
        const int xl_hi = to_int(((float)cx_hi + 0.5f) / (float)W * (float)Wl);
        const int yl_lo = to_int(((float)cy_lo + 0.5f) / (float)H * (float)Hl);
        const int yl_hi = to_int(((float)cy_hi + 0.5f) / (float)H * (float)Hl);
        const int w = xl_hi - xl_lo + 7, h = yl_hi - yl_lo + 7;
        // staged only for inputs no larger than the output (scale >= 1), where the bounds above hold
        const bool stage = Wl <= W && Hl <= H && w <= G::UCW && h <= G::UCH;
        if (stage) {
            ui.x0 = xl_lo - 3;
            ui.y0 = yl_lo - 3;
            ui.w = w;
            ui.h = h;
        }
        // per-column and per-row tables of the resolve region (clamped texel, uv as tabulated
        // below): the bilinear axes of SAMPLE_TEX at uv + (d - 1) / input size, and the input
        // column (row) getCurrentUpsampledColor fetches from (upsample_pick). A staged tile keeps the
        // bilinear taps as positions in the staged rectangle (x - x0, (y - y0) * UCW), which hold
        // the same texels.
        {
            const float toUx = 1.0f / (float)Wl, toUy = 1.0f / (float)Hl;
            const float jux = U.mJitterNdc[0] * 0.5f * P.mUnjitterFactor, juy = U.mJitterNdc[1] * 0.5f * P.mUnjitterFactor;
            const float lx = (float)Wl, ly = (float)Hl, hxs = (float)W, hys = (float)H;
            const float tjx = jux * lx * -1.0f, tjy = juy * ly * -1.0f;
            for (int i = tid; i < 4 * (G::RW + G::RH); i += G::NT) {
                const bool col = i < 4 * G::RW;
                const int n = col ? G::RW : G::RH, e = col ? i : i - 4 * G::RW;
                const int q = e / n, j = e - q * n;  // q = 0..2: bilinear offset, 3: the sample search
                const int c = col ? clampi(x0 - G::RA + j, 0, W - 1) : clampi(y0 - G::RA + j, 0, H - 1);
                const float uv = col ? ((float)c + 0.5f) / (float)W : ((float)c + 0.5f) / (float)H;
                if (q < 3) {
                    int o0;
                    BilinearAxis ba = col ? bilinear_axis(Wl, uv + toUx * (float)(q - 1), &o0)
                                          : bilinear_axis(Hl, uv + toUy * (float)(q - 1), &o0);
                    if (stage) {  // the first tap's position before the edge clamp; the second follows it
                        ba.a = col ? o0 - ui.x0 : (o0 - ui.y0) * G::UCW;
                        ba.b = col ? ba.a + 1 : ba.a + G::UCW;
                    }
                    (col ? sm.uax : sm.uay)[q * n + j] = ba;
                } else {
                    (col ? sm.ucx : sm.ucy)[j] = col ? upsample_pick(lx, hxs, tjx, c, Wl, true)
                                                     : upsample_pick(ly, hys, tjy, c, Hl, false);
                }
            }
        }
        if (stage) {
            for (int i = tid; i < w * h; i += G::NT) {
                const int sy = i / w, sx = i - sy * w;
                const int gx = clampi(ui.x0 + sx, 0, Wl - 1), gy = clampi(ui.y0 + sy, 0, Hl - 1);
                const uint32_t gi = (uint32_t)(gy * pl + gx);
                const f3 t = stage_convert(P, xyz(ld_rgba<S16>(fd.color, gi)));
                const int k = sy * G::UCW + sx;
                sm.uc0[k] = t.x;
                sm.uc1[k] = t.y;
                sm.uc2[k] = t.z;
                sm.ud[k] = ld_f32(fd.depth, gi);
                sm.uvel[k] = ld_vel(fd.velocity, fd.vel_fmt, gi);
            }
        }
    }
    if (tid < G::SW) sm.ux[tid] = ((float)clampi(gx0 + tid, 0, W - 1) + 0.5f) / (float)W;
    else if (tid < G::SW + G::SH) sm.uy[tid - G::SW] = ((float)clampi(gy0 + tid - G::SW, 0, H - 1) + 0.5f) / (float)H;
    __syncthreads();

    // 2. resolve the region
    for (int i = tid; i < G::RW * G::RH; i += G::NT) {
        const int ry = i / G::RW, rx = i - ry * G::RW;
        const int px = x0 - G::RA + rx, py = y0 - G::RA + ry;
        const int cx = clampi(px, 0, W - 1), cy = clampi(py, 0, H - 1);
        const bool owned = rx >= G::RA && rx < G::RA + G::TW && ry >= G::RA && ry < G::RA + G::TH && px < fd.x_end && py < H;
        const int l = (cy - gy0) * G::SP + (cx - gx0);
        const SegPlanes spl{sm.nx, sm.ny, sm.nz, sm.lum, sm.mat};
        const float uvx = sm.ux[cx - gx0], uvy = sm.uy[cy - gy0];
        PixelResult r;
        if constexpr (G::UP) {
            // every input fetch of a texel lies within 2 input texels of its (xl, yl) = int(uv * input
            // size) (bilinear taps at uv +- one texel, the neighbourhood, the upsampled sample), and
            // xl, yl grow with the output column and row, so a staged tile's rectangle holds them
            // all: the staged variant reads only LDS
            if (ui.w > 0)
                r = resolve_pixel_f32<S16, G::SP, true, SEG, true, false, FA, G::UCW>(fd, U, P, cx, cy, l, sm.c0, sm.c1, sm.c2, sm.d,
                                                                             uvx, uvy, owned, sg, &spl, &ui, rx, ry);
            else
                r = resolve_pixel_f32<S16, G::SP, true, SEG, false, false, FA>(fd, U, P, cx, cy, l, sm.c0, sm.c1, sm.c2, sm.d, uvx, uvy,
                                                                    owned, sg, &spl, &ui, rx, ry);
        } else {
            r = resolve_pixel_f32<S16, G::SP, false, SEG, false, G::VEL16, FA>(fd, U, P, cx, cy, l, sm.c0, sm.c1, sm.c2, sm.d,
                                                                                uvx, uvy, owned, sg, &spl, nullptr, 0, 0,
                                                                                vel_staged ? sm.vel : nullptr);
        }
        if (owned) {
            const uint32_t gi = (uint32_t)(cy * pitch + cx);
            st_rgba<S16>(fd.history_out, gi, r.history);
            if constexpr (!G::CAS) st_rgba<S16>(fd.screen, gi, make_float4(r.screen.x, r.screen.y, r.screen.z, 1.0f));
        }
        if constexpr (G::CAS) {
            // the CAS pass reads the stored screen image: round to the storage format first
            const int li = ry * G::RP + rx;
            if constexpr (FA && S16) {  // its fp16 bits, for the FP16 path's packed-half CAS
                sm.packed()[li] = make_uint2(pack_f2h(r.screen.x, r.screen.y), f2h(r.screen.z));
            } else {
                f3 s = r.screen;
                if constexpr (S16) s = {round_f16(s.x), round_f16(s.y), round_f16(s.z)};
                sm.r0[li] = s.x;
                sm.r1[li] = s.y;
                sm.r2[li] = s.z;
            }
        }
    }

    // 3. fused CAS from the resolved region
    if constexpr (G::CAS) {
        __syncthreads();
        const float peak = __uint_as_float(fd.cas_peak_bits);
        for (int i = tid; i < G::TW * G::TH; i += G::NT) {
            const int oy = i / G::TW, ox = i - oy * G::TW;
            const int px = x0 + ox, py = y0 + oy;
            if (px >= fd.x_end || py >= H) continue;
            const int li = (oy + 1) * G::RP + (ox + 1);
            if constexpr (FA && S16) {
                const uint2* q = sm.packed();
                gstore<uint2>(fd.screen, (uint32_t)(py * pitch + px) * 8u,
                              cas_filter_h2(q[li - G::RP], q[li - 1], q[li], q[li + 1], q[li + G::RP], peak));
                continue;
            }
            auto at = [&](int k) { return f3{sm.r0[k], sm.r1[k], sm.r2[k]}; };
            f3 o = FA ? cas_filter_fma(at(li - G::RP), at(li - 1), at(li), at(li + 1), at(li + G::RP), peak)
                      : cas_filter(at(li - G::RP), at(li - 1), at(li), at(li + 1), at(li + G::RP), peak);
            st_rgba<S16>(fd.screen, (uint32_t)(py * pitch + px), make_float4(o.x, o.y, o.z, 1.0f));
        }
    }
}

struct ResolveArgs {
    FrameDesc fd;
    taa_uniforms_t U;
};

// PROD: the production feature set with its switches fixed (prod_fix)
template <int TW, int TH, int PAD, bool CAS, bool S16, bool PROD>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(S16 ? 4 : 3))) void k_resolve(const ResolveArgs a) {
    using G = Geo<TW, TH, PAD, CAS, false, false, S16>;
    __shared__ typename G::Smem sm;
    const int n = a.fd.tiles_x * a.fd.tiles_y;
    resolve_tile<G, S16, false, false, PROD>(a.fd, a.U, xcd_tile(blockIdx.x, n), sm);
}

// the resolve with the segmentation mask of mRayTraceAugment (FA: the FP16 path's arithmetic where
// no selection depends on it; PROD: the production feature set with the mask, prod_params(.., true))
template <int TW, int TH, int PAD, bool CAS, bool S16, bool FA, bool PROD = false>
__global__ __launch_bounds__(256) void k_resolve_seg(const ResolveArgs a, const SegDesc sg) {
    using G = Geo<TW, TH, PAD, CAS, false, true>;
    __shared__ typename G::Smem sm;
    const int n = a.fd.tiles_x * a.fd.tiles_y;
    resolve_tile<G, S16, true, FA, PROD>(a.fd, a.U, xcd_tile(blockIdx.x, n), sm, &sg);
}

// the upsampling resolve with the segmentation mask (its inputs read from global memory)
template <int TW, int TH, int PAD, bool CAS, bool S16, bool FA>
__global__ __launch_bounds__(256) void k_resolve_up_seg(const ResolveArgs a, const SegDesc sg) {
    using G = Geo<TW, TH, PAD, CAS, true>;
    __shared__ typename G::Smem sm;
    const int n = a.fd.tiles_x * a.fd.tiles_y;
    resolve_tile<G, S16, true, FA>(a.fd, a.U, xcd_tile(blockIdx.x, n), sm, &sg);
}

// Upsampling path (input size != output size, or mUpsampling): exact FP32 arithmetic, or with FA
// the FP16 path's (TAA_ARITH_FP16); the input texels under the tile are staged in LDS (converted
// colour, depth, velocity), the rest of the resolve as above, fused CAS from LDS.
// PROD: the production feature set with its switches fixed (prod_fix). Occupancy hint: the FP16
// arithmetic's general form fits five waves per SIMD in registers; the production form's 30x14
// tiles hold four workgroups per CU in LDS, so four
template <int TW, int TH, int PAD, bool CAS, bool S16, bool FA, bool PROD>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(FA && S16 ? (PROD ? 4 : 5) : 1))) void k_resolve_up(const ResolveArgs a) {
    using G = Geo<TW, TH, PAD, CAS, true>;
    __shared__ typename G::Smem sm;
    const int n = a.fd.tiles_x * a.fd.tiles_y;
    resolve_tile<G, S16, false, FA, PROD>(a.fd, a.U, xcd_tile(blockIdx.x, n), sm);
}

// Batched resolve: one workgroup per tile of the flattened tile space of every frame in the batch
// (batch_tile: XCD k takes the k-th contiguous eighth of the tiles).
// PROD: every frame of the batch has the production feature set (exact_prod_form); its RG16F
// velocity is then staged like k_resolve's.
template <int TW, int TH, int PAD, bool CAS, bool S16, bool PROD>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(S16 ? 4 : 3))) void k_resolve_batch(const BatchEntry* __restrict__ entries, int n_entries, int total_tiles) {
    using G = Geo<TW, TH, PAD, CAS, false, false, PROD && S16>;
    __shared__ typename G::Smem sm;
    const int t = batch_tile(total_tiles);
    if (t < 0) return;
    const BatchEntry& e = batch_entry(entries, n_entries, t);
    resolve_tile<G, S16, false, false, PROD>(e.fd, e.U, t - e.tile_begin, sm);
}

// Batched upsampling resolve: the frames of a batch whose input size differs from the output size
// (or that set mUpsampling), walked like k_resolve_batch, each tile as in k_resolve_up (its tile
// shapes and occupancy hint).
// PROD: every frame of the batch has the production feature set (prod_params)
template <int TW, int TH, int PAD, bool CAS, bool S16, bool FA, bool PROD>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(FA && S16 ? (PROD ? 4 : 5) : 1))) void k_resolve_up_batch(const BatchEntry* __restrict__ entries, int n_entries, int total_tiles) {
    using G = Geo<TW, TH, PAD, CAS, true>;
    __shared__ typename G::Smem sm;
    const int t = batch_tile(total_tiles);
    if (t < 0) return;
    const BatchEntry& e = batch_entry(entries, n_entries, t);
    resolve_tile<G, S16, false, FA, PROD>(e.fd, e.U, t - e.tile_begin, sm);
}

// sharpen_cas.comp on its own (one thread per pixel, clamped loads)
template <bool S16>
__global__ __launch_bounds__(256) void k_cas(const void* __restrict__ src, void* __restrict__ dst, int W, int H, int pitch, uint32_t peak_bits) {
    const int x = blockIdx.x * 16 + (threadIdx.x & 15), y = blockIdx.y * 16 + (threadIdx.x >> 4);
    if (x >= W || y >= H) return;
    auto ld = [&](int xx, int yy) {
        return xyz(ld_rgba<S16>(src, (uint32_t)(clampi(yy, 0, H - 1) * pitch + clampi(xx, 0, W - 1))));
    };
    f3 o = cas_filter(ld(x, y - 1), ld(x - 1, y), ld(x, y), ld(x + 1, y), ld(x, y + 1), __uint_as_float(peak_bits));
    st_rgba<S16>(dst, (uint32_t)(y * pitch + x), make_float4(o.x, o.y, o.z, 1.0f));
}

// ------------------------------------------------------------------------------------------
// dispatch tables

template <bool CAS, bool S16, bool PROD>
hipError_t launch_resolve_t(const FrameDesc& fd, const taa_uniforms_t& U, int shape, hipStream_t st) {
    ResolveArgs a;
    a.fd = fd;
    a.U = U;
    const int n = fd.tiles_x * fd.tiles_y;
    constexpr const TileShapeDef* S = CAS ? kShapesCas : kShapesPlain;
    switch (shape) {
        case 0: hipLaunchKernelGGL((k_resolve<S[0].tw, S[0].th, S[0].pad, CAS, S16, PROD>), dim3(n), dim3(256), 0, st, a); break;
        case 1: hipLaunchKernelGGL((k_resolve<S[1].tw, S[1].th, S[1].pad, CAS, S16, PROD>), dim3(n), dim3(256), 0, st, a); break;
        case 2: hipLaunchKernelGGL((k_resolve<S[2].tw, S[2].th, S[2].pad, CAS, S16, PROD>), dim3(n), dim3(256), 0, st, a); break;
        default: hipLaunchKernelGGL((k_resolve<S[3].tw, S[3].th, S[3].pad, CAS, S16, PROD>), dim3(n), dim3(256), 0, st, a); break;
    }
    return hipGetLastError();
}

template <bool CAS, bool S16, bool PROD>
hipError_t launch_batch_t(const BatchEntry* e, int n, int total, int shape, hipStream_t st) {
    const int grid = batch_grid(total);
    constexpr const TileShapeDef* S = CAS ? kShapesCas : kShapesPlain;
    switch (shape) {
        case 0: hipLaunchKernelGGL((k_resolve_batch<S[0].tw, S[0].th, S[0].pad, CAS, S16, PROD>), dim3(grid), dim3(256), 0, st, e, n, total); break;
        case 1: hipLaunchKernelGGL((k_resolve_batch<S[1].tw, S[1].th, S[1].pad, CAS, S16, PROD>), dim3(grid), dim3(256), 0, st, e, n, total); break;
        case 2: hipLaunchKernelGGL((k_resolve_batch<S[2].tw, S[2].th, S[2].pad, CAS, S16, PROD>), dim3(grid), dim3(256), 0, st, e, n, total); break;
        default: hipLaunchKernelGGL((k_resolve_batch<S[3].tw, S[3].th, S[3].pad, CAS, S16, PROD>), dim3(grid), dim3(256), 0, st, e, n, total); break;
    }
    return hipGetLastError();
}

}  // namespace

bool prod_params(const FrameDesc& fd, const taa_uniforms_t& U, bool seg) {
    const taa_params_t& p = U.param[fd.param_idx];
    const bool s16 = fd.color_fmt == TAA_FMT_RGBA16F;
    return !fd.debug && !fd.inspect && !U.mBypassHistoryUpdate &&
           fd.vel_fmt == (s16 ? TAA_VEL_RG16F : TAA_VEL_RG32F) && !p.mPassThrough && !p.mUnjitterNeighbourhood &&
           !p.mUnjitterCurrentSample && p.mVarianceClipping && !p.mShapedNeighbourhood && p.mUseYCoCg &&
           !p.mShrinkChromaAxis && p.mUseVelocityVectors == 2 && p.mVelocitySampleMode == 2 && p.mDepthCulling &&
           p.mInterpolationMode == 2 && !p.mDynamicAntiGhosting && (p.mRayTraceAugment != 0) == seg &&
           p.mColorClampingOrClipping == 2 &&
           !p.mVelBasedAlpha && p.mLumaWeightingLottes && !p.mReduceBlendNearClamp && p.mRejectOutside && !p.mAddNoise &&
           !p.mToneMapLumaKaris;
}

bool exact_prod_form(const FrameDesc& fd, const taa_uniforms_t& U) { return !fd.up && !U.mUpsampling && prod_params(fd, U); }

int launch_resolve(const FrameDesc& fd, const taa_uniforms_t& U, int shape, int arith, void* stream) {
    (void)arith;
    hipStream_t st = static_cast<hipStream_t>(stream);
    const bool s16 = fd.color_fmt == TAA_FMT_RGBA16F;
    hipError_t e;
    if (exact_prod_form(fd, U)) {
        if (fd.cas_on) e = s16 ? launch_resolve_t<true, true, true>(fd, U, shape, st) : launch_resolve_t<true, false, true>(fd, U, shape, st);
        else e = s16 ? launch_resolve_t<false, true, true>(fd, U, shape, st) : launch_resolve_t<false, false, true>(fd, U, shape, st);
    } else {
        if (fd.cas_on) e = s16 ? launch_resolve_t<true, true, false>(fd, U, shape, st) : launch_resolve_t<true, false, false>(fd, U, shape, st);
        else e = s16 ? launch_resolve_t<false, true, false>(fd, U, shape, st) : launch_resolve_t<false, false, false>(fd, U, shape, st);
    }
    return (int)e;
}

// the launch's tiling follows the form's tile shape: tiles_x / tiles_y of `a` are set here
template <bool FA, bool PROD>
hipError_t launch_resolve_up_t(ResolveArgs a, hipStream_t st) {
    FrameDesc& fd = a.fd;
    const bool s16 = fd.color_fmt == TAA_FMT_RGBA16F;
    constexpr TileShapeDef C = kShapesCas[PROD ? kUpShapeCasProd : kUpShapeCas], Pl = kShapesPlain[kUpShapePlain];
    const TileShapeDef d = fd.cas_on ? C : Pl;
    fd.tiles_x = tiles_for(fd.x_end - fd.x_begin, d.tw);
    fd.tiles_y = tiles_for(fd.H, d.th);
    const int n = fd.tiles_x * fd.tiles_y;
    if (fd.cas_on) {
        if (s16) hipLaunchKernelGGL((k_resolve_up<C.tw, C.th, C.pad, true, true, FA, PROD>), dim3(n), dim3(256), 0, st, a);
        else hipLaunchKernelGGL((k_resolve_up<C.tw, C.th, C.pad, true, false, FA, PROD>), dim3(n), dim3(256), 0, st, a);
    } else {
        if (s16) hipLaunchKernelGGL((k_resolve_up<Pl.tw, Pl.th, Pl.pad, false, true, FA, PROD>), dim3(n), dim3(256), 0, st, a);
        else hipLaunchKernelGGL((k_resolve_up<Pl.tw, Pl.th, Pl.pad, false, false, FA, PROD>), dim3(n), dim3(256), 0, st, a);
    }
    return hipGetLastError();
}

int launch_resolve_up(const FrameDesc& fd, const taa_uniforms_t& U, int arith, void* stream) {
    hipStream_t st = static_cast<hipStream_t>(stream);
    ResolveArgs a;
    a.fd = fd;
    a.U = U;
    const bool prod = prod_params(fd, U);
    hipError_t e;
    if (arith == TAA_ARITH_FP16) e = prod ? launch_resolve_up_t<true, true>(a, st) : launch_resolve_up_t<true, false>(a, st);
    else e = prod ? launch_resolve_up_t<false, true>(a, st) : launch_resolve_up_t<false, false>(a, st);
    return (int)e;
}

template <bool UP, bool FA>
hipError_t launch_resolve_seg_t(ResolveArgs a, const SegDesc& sg, hipStream_t st) {
    FrameDesc& fd = a.fd;
    const bool s16 = fd.color_fmt == TAA_FMT_RGBA16F;
    constexpr TileShapeDef C = kShapesCas[UP ? kUpShapeCas : kSegShapeCas], Pl = kShapesPlain[UP ? kUpShapePlain : kSegShapePlain];
    const TileShapeDef d = fd.cas_on ? C : Pl;
    fd.tiles_x = tiles_for(fd.x_end - fd.x_begin, d.tw);
    fd.tiles_y = tiles_for(fd.H, d.th);
    const int n = fd.tiles_x * fd.tiles_y;
    if constexpr (UP) {
        if (fd.cas_on) {
            if (s16) hipLaunchKernelGGL((k_resolve_up_seg<C.tw, C.th, C.pad, true, true, FA>), dim3(n), dim3(256), 0, st, a, sg);
            else hipLaunchKernelGGL((k_resolve_up_seg<C.tw, C.th, C.pad, true, false, FA>), dim3(n), dim3(256), 0, st, a, sg);
        } else {
            if (s16) hipLaunchKernelGGL((k_resolve_up_seg<Pl.tw, Pl.th, Pl.pad, false, true, FA>), dim3(n), dim3(256), 0, st, a, sg);
            else hipLaunchKernelGGL((k_resolve_up_seg<Pl.tw, Pl.th, Pl.pad, false, false, FA>), dim3(n), dim3(256), 0, st, a, sg);
        }
    } else {
        if (fd.cas_on && prod_params(fd, a.U, true)) {  // the production feature set with the mask
            if (s16) hipLaunchKernelGGL((k_resolve_seg<C.tw, C.th, C.pad, true, true, FA, true>), dim3(n), dim3(256), 0, st, a, sg);
            else hipLaunchKernelGGL((k_resolve_seg<C.tw, C.th, C.pad, true, false, FA, true>), dim3(n), dim3(256), 0, st, a, sg);
        } else if (fd.cas_on) {
            if (s16) hipLaunchKernelGGL((k_resolve_seg<C.tw, C.th, C.pad, true, true, FA>), dim3(n), dim3(256), 0, st, a, sg);
            else hipLaunchKernelGGL((k_resolve_seg<C.tw, C.th, C.pad, true, false, FA>), dim3(n), dim3(256), 0, st, a, sg);
        } else {
            if (s16) hipLaunchKernelGGL((k_resolve_seg<Pl.tw, Pl.th, Pl.pad, false, true, FA>), dim3(n), dim3(256), 0, st, a, sg);
            else hipLaunchKernelGGL((k_resolve_seg<Pl.tw, Pl.th, Pl.pad, false, false, FA>), dim3(n), dim3(256), 0, st, a, sg);
        }
    }
    return hipGetLastError();
}

int launch_resolve_seg(const FrameDesc& fd, const taa_uniforms_t& U, const SegDesc& sg, int arith, void* stream) {
    hipStream_t st = static_cast<hipStream_t>(stream);
    ResolveArgs a;
    a.fd = fd;
    a.U = U;
    const bool fa = arith == TAA_ARITH_FP16;
    hipError_t e;
    if (fd.up) e = fa ? launch_resolve_seg_t<true, true>(a, sg, st) : launch_resolve_seg_t<true, false>(a, sg, st);
    else e = fa ? launch_resolve_seg_t<false, true>(a, sg, st) : launch_resolve_seg_t<false, false>(a, sg, st);
    return (int)e;
}

int launch_cas(const void* src, void* dst, int W, int H, int pitch, int fmt, uint32_t peak_bits, void* stream) {
    hipStream_t st = static_cast<hipStream_t>(stream);
    dim3 grid((W + 15) / 16, (H + 15) / 16);
    if (fmt == TAA_FMT_RGBA16F) hipLaunchKernelGGL(k_cas<true>, grid, dim3(256), 0, st, src, dst, W, H, pitch, peak_bits);
    else hipLaunchKernelGGL(k_cas<false>, grid, dim3(256), 0, st, src, dst, W, H, pitch, peak_bits);
    return (int)hipGetLastError();
}

template <bool FA, bool PROD>
hipError_t launch_batch_up_t(const BatchEntry* d_entries, int n_entries, int total_tiles, bool cas, bool s16,
                             hipStream_t st) {
    constexpr TileShapeDef C = kShapesCas[PROD ? kUpShapeCasProd : kUpShapeCas], Pl = kShapesPlain[kUpShapePlain];
    const int grid = batch_grid(total_tiles);
    if (cas) {
        if (s16) hipLaunchKernelGGL((k_resolve_up_batch<C.tw, C.th, C.pad, true, true, FA, PROD>), dim3(grid), dim3(256), 0, st, d_entries, n_entries, total_tiles);
        else hipLaunchKernelGGL((k_resolve_up_batch<C.tw, C.th, C.pad, true, false, FA, PROD>), dim3(grid), dim3(256), 0, st, d_entries, n_entries, total_tiles);
    } else {
        if (s16) hipLaunchKernelGGL((k_resolve_up_batch<Pl.tw, Pl.th, Pl.pad, false, true, FA, PROD>), dim3(grid), dim3(256), 0, st, d_entries, n_entries, total_tiles);
        else hipLaunchKernelGGL((k_resolve_up_batch<Pl.tw, Pl.th, Pl.pad, false, false, FA, PROD>), dim3(grid), dim3(256), 0, st, d_entries, n_entries, total_tiles);
    }
    return hipGetLastError();
}

int launch_batch_up(const BatchEntry* d_entries, int n_entries, int total_tiles, bool cas, bool s16, int arith, bool prod,
                    void* stream) {
    hipStream_t st = static_cast<hipStream_t>(stream);
    hipError_t e;
    if (arith == TAA_ARITH_FP16)
        e = prod ? launch_batch_up_t<true, true>(d_entries, n_entries, total_tiles, cas, s16, st)
                 : launch_batch_up_t<true, false>(d_entries, n_entries, total_tiles, cas, s16, st);
    else
        e = prod ? launch_batch_up_t<false, true>(d_entries, n_entries, total_tiles, cas, s16, st)
                 : launch_batch_up_t<false, false>(d_entries, n_entries, total_tiles, cas, s16, st);
    return (int)e;
}

int launch_batch(const BatchEntry* d_entries, int n_entries, int total_tiles, int shape, bool cas, bool s16, int arith,
                 bool prod, void* stream) {
    (void)arith;
    hipStream_t st = static_cast<hipStream_t>(stream);
    hipError_t e;
    if (prod) {
        if (cas) e = s16 ? launch_batch_t<true, true, true>(d_entries, n_entries, total_tiles, shape, st)
                         : launch_batch_t<true, false, true>(d_entries, n_entries, total_tiles, shape, st);
        else e = s16 ? launch_batch_t<false, true, true>(d_entries, n_entries, total_tiles, shape, st)
                     : launch_batch_t<false, false, true>(d_entries, n_entries, total_tiles, shape, st);
    } else {
        if (cas) e = s16 ? launch_batch_t<true, true, false>(d_entries, n_entries, total_tiles, shape, st)
                         : launch_batch_t<true, false, false>(d_entries, n_entries, total_tiles, shape, st);
        else e = s16 ? launch_batch_t<false, true, false>(d_entries, n_entries, total_tiles, shape, st)
                     : launch_batch_t<false, false, false>(d_entries, n_entries, total_tiles, shape, st);
    }
    return (int)e;
}


}  // namespace taastar
