// Device helpers shared by the gfx950 resolve kernels: storage formats, FP32 operations in
// taa.comp order, the CAS filter and the tile-to-XCD mapping.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "taa_footprint.hpp"
#include "taa_tiles.hpp"
#include "taa_internal.hpp"

namespace taastar {
namespace {

// ------------------------------------------------------------------------------------------
// storage

__device__ __forceinline__ float h2f(uint32_t h) { return (float)__builtin_bit_cast(_Float16, (uint16_t)h); }
// fp32 -> fp16 storage conversion: round to nearest even, saturating to +-65504 (NaN stays NaN)
// so that an overflow never becomes an Inf that the history filter would spread as NaN.
__device__ __forceinline__ float sat_f16_range(float f) { return f == f ? __builtin_amdgcn_fmed3f(f, -65504.0f, 65504.0f) : f; }
__device__ __forceinline__ uint32_t f2h(float f) { return (uint32_t)__builtin_bit_cast(uint16_t, (_Float16)sat_f16_range(f)); }
__device__ __forceinline__ float round_f16(float f) { return (float)(_Float16)sat_f16_range(f); }
// two channels into one dword (low half = a), same rounding and saturation as f2h
typedef float f32x2_t __attribute__((ext_vector_type(2)));
typedef _Float16 f16x2_t __attribute__((ext_vector_type(2)));
__device__ __forceinline__ uint32_t pack_f2h(float a, float b) {
    const f32x2_t v = {sat_f16_range(a), sat_f16_range(b)};
    return __builtin_bit_cast(uint32_t, __builtin_convertvector(v, f16x2_t));
}

// Global memory through a uniform base pointer and a 32-bit byte offset: every image is smaller
// than 4 GiB (the host refuses larger ones), so texel offsets are computed in 32 bits and the
// accesses use the scalar-base form of global loads and stores. `imm` is added to the pointer
// (not to the 32-bit offset) so that it folds into the instruction's immediate offset.
#if defined(__HIP_DEVICE_COMPILE__)
#define TAA_GLOBAL __attribute__((address_space(1)))
#define TAA_CONSTANT __attribute__((address_space(4)))
#else
#define TAA_GLOBAL  // host pass: these functions are parsed but never compiled for the CPU
#define TAA_CONSTANT
#endif
template <class T>
__device__ __forceinline__ T gload(const void* base, uint32_t off, int imm = 0) {
    return *(const TAA_GLOBAL T*)((const TAA_GLOBAL char*)base + off + imm);
}
template <class T>
__device__ __forceinline__ void gstore(void* base, uint32_t off, T v, int imm = 0) {
    *(TAA_GLOBAL T*)((TAA_GLOBAL char*)base + off + imm) = v;
}

// four dwords at an 8-byte-aligned address (two RGBA16F texels): one global_load_dwordx4
typedef uint32_t u32x4_a8 __attribute__((ext_vector_type(4), aligned(8)));
__device__ __forceinline__ float4 rgba16_bits(uint32_t lo, uint32_t hi) {
    return make_float4(h2f(lo & 0xffffu), h2f(lo >> 16), h2f(hi & 0xffffu), h2f(hi >> 16));
}

// texel i of an RGBA image (element index, not bytes)
template <bool S16>
__device__ __forceinline__ float4 ld_rgba(const void* p, uint32_t i) {
    if constexpr (S16) {
        uint2 v = gload<uint2>(p, i * 8u);
        return make_float4(h2f(v.x & 0xffffu), h2f(v.x >> 16), h2f(v.y & 0xffffu), h2f(v.y >> 16));
    } else {
        return gload<float4>(p, i * 16u);
    }
}
template <bool S16>
__device__ __forceinline__ void st_rgba(void* p, uint32_t i, float4 v) {
    if constexpr (S16) {
        gstore<uint2>(p, i * 8u, make_uint2(pack_f2h(v.x, v.y), pack_f2h(v.z, v.w)));
    } else {
        gstore<float4>(p, i * 16u, v);
    }
}
__device__ __forceinline__ void st_rgba16(void* p, uint32_t i, float4 v) { st_rgba<true>(p, i, v); }
__device__ __forceinline__ float ld_f32(const void* p, uint32_t i) { return gload<float>(p, i * 4u); }

// Velocity texel; RG formats read back as (r, g, 0, 1).
__device__ __forceinline__ float4 ld_vel(const void* p, int fmt, uint32_t i) {
    switch (fmt) {
        case TAA_VEL_RG32F: { float2 v = gload<float2>(p, i * 8u); return make_float4(v.x, v.y, 0.f, 1.f); }
        case TAA_VEL_RG16F: { uint32_t v = gload<uint32_t>(p, i * 4u); return make_float4(h2f(v & 0xffffu), h2f(v >> 16), 0.f, 1.f); }
        case TAA_VEL_RGBA32F: return gload<float4>(p, i * 16u);
        default: {
            uint2 v = gload<uint2>(p, i * 8u);
            return make_float4(h2f(v.x & 0xffffu), h2f(v.x >> 16), h2f(v.y & 0xffffu), h2f(v.y >> 16));
        }
    }
}

// the 16 inspect words of texel i, as four 16-byte stores
__device__ __forceinline__ void st_inspect(void* p, uint32_t i, const int32_t (&w)[TAA_INSPECT_WORDS]) {
    const uint32_t o = i * (uint32_t)(TAA_INSPECT_WORDS * 4);
    gstore<int4>(p, o, make_int4(w[0], w[1], w[2], w[3]));
    gstore<int4>(p, o, make_int4(w[4], w[5], w[6], w[7]), 16);
    gstore<int4>(p, o, make_int4(w[8], w[9], w[10], w[11]), 32);
    gstore<int4>(p, o, make_int4(w[12], w[13], w[14], w[15]), 48);
}

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return min(max(v, lo), hi); }

// x / 9 correctly rounded, as the IEEE division gives it, in three operations: q0 = x * RN(1/9),
// the FMA residual r = x - 9 q0 and the FMA correction q0 + r RN(1/9) (Markstein's correction
// with a correctly rounded reciprocal); a zero or infinite q0 is the quotient itself. Identical to
// x / 9.0f for every single-precision x (checked exhaustively on the CPU by the tests).
__device__ __forceinline__ float div9(float x) {
    const float y = 1.0f / 9.0f;
    const float q0 = x * y;
    const float r = __builtin_fmaf(-q0, 9.0f, x);
    return (q0 == 0.0f || __builtin_isinf(q0)) ? q0 : __builtin_fmaf(r, y, q0);
}

// int(x): NaN -> 0, saturating to the int range, truncation toward zero -- exactly what
// V_CVT_I32_F32 does, so emit that instruction instead of letting the compiler treat
// out-of-range conversions as undefined.
__device__ __forceinline__ int to_int(float v) {
    int r;
    asm("v_cvt_i32_f32 %0, %1" : "=v"(r) : "v"(v));
    return r;
}

// ------------------------------------------------------------------------------------------
// FP32 arithmetic in taa.comp order

struct f3 { float x, y, z; };
__device__ __forceinline__ f3 operator+(f3 a, f3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ f3 operator-(f3 a, f3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ f3 operator*(f3 a, f3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
__device__ __forceinline__ f3 operator/(f3 a, f3 b) { return {a.x / b.x, a.y / b.y, a.z / b.z}; }
__device__ __forceinline__ f3 operator*(float s, f3 a) { return {s * a.x, s * a.y, s * a.z}; }
__device__ __forceinline__ f3 operator*(f3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ f3 operator/(f3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }
__device__ __forceinline__ f3 vmin(f3 a, f3 b) { return {fminf(a.x, b.x), fminf(a.y, b.y), fminf(a.z, b.z)}; }
__device__ __forceinline__ f3 vmax(f3 a, f3 b) { return {fmaxf(a.x, b.x), fmaxf(a.y, b.y), fmaxf(a.z, b.z)}; }
__device__ __forceinline__ float gmix(float x, float y, float a) { return x * (1.0f - a) + y * a; }
__device__ __forceinline__ float gclamp(float x, float lo, float hi) { return fminf(fmaxf(x, lo), hi); }
__device__ __forceinline__ f3 xyz(float4 v) { return {v.x, v.y, v.z}; }

// taa.comp:158-174
__device__ __forceinline__ f3 rgb_to_ycocg(f3 c) {
    return {.25f * c.x + .5f * c.y + .25f * c.z, .5f * c.x - .5f * c.z, -.25f * c.x + .5f * c.y - .25f * c.z};
}
__device__ __forceinline__ f3 ycocg_to_rgb(f3 c) {
    float tmp = c.x - c.z;
    return {tmp + c.y, c.x + c.z, tmp - c.y};
}
// taa.comp:182-197
__device__ __forceinline__ f3 tonemap_rgb(const taa_params_t& P, f3 hdr) {
    if (P.mToneMapLumaKaris) {
        float luma = fmaxf(fmaxf(hdr.x, hdr.y), hdr.z);
        return hdr / (1.0f + luma);
    }
    return hdr;
}
__device__ __forceinline__ f3 un_tonemap_rgb(const taa_params_t& P, f3 ldr) {
    if (P.mToneMapLumaKaris) {
        float luma = fmaxf(fmaxf(ldr.x, ldr.y), ldr.z);
        return ldr / (1.0f - luma);
    }
    return ldr;
}
__device__ __forceinline__ f3 stage_convert(const taa_params_t& P, f3 c) {
    c = tonemap_rgb(P, c);
    return P.mUseYCoCg ? rgb_to_ycocg(c) : c;
}
__device__ __forceinline__ float luminance(const taa_params_t& P, f3 c) { return P.mUseYCoCg ? c.x : rgb_to_ycocg(c).x; }
__device__ __forceinline__ float linearize_depth(float d, float n, float f) { return n * f / (f + d * (n - f)); }

// sin and cos of the G-buffer's normal angles (taa.comp:574-579): Cody-Waite reduction by pi/2
// with a three-part constant (Cody & Waite, Software Manual for the Elementary Functions, 1980),
// then the Taylor polynomials of sin (degree 9) and cos (degree 10) on [-pi/4, pi/4] in Horner
// form, every step an FMA or a correctly rounded operation; within 1.5 ulp of sin / cos on
// [-200, 200], and the same bits as the host evaluation of the same steps.
__device__ __forceinline__ void seg_sincos(float a, float& s, float& c) {
    const float k = rintf(a * 0x1.45f306p-1f);
    float r = __builtin_fmaf(-k, 0x1.921fb6p+0f, a);
    r = __builtin_fmaf(-k, -0x1.777a5cp-25f, r);
    r = __builtin_fmaf(-k, -0x1.ee59dap-50f, r);
    const float r2 = r * r;
    float ps = __builtin_fmaf(0x1.71de3ap-19f, r2, -0x1.a01a02p-13f);
    ps = __builtin_fmaf(ps, r2, 0x1.111112p-7f);
    ps = __builtin_fmaf(ps, r2, -0x1.555556p-3f);
    const float sr = __builtin_fmaf(r * r2, ps, r);
    float pc = __builtin_fmaf(-0x1.27e4fcp-22f, r2, 0x1.a01a02p-16f);
    pc = __builtin_fmaf(pc, r2, -0x1.6c16c2p-10f);
    pc = __builtin_fmaf(pc, r2, 0x1.555556p-5f);
    pc = __builtin_fmaf(pc, r2, -0.5f);
    const float cr = __builtin_fmaf(r2, pc, 1.0f);
    switch (to_int(k) & 3) {
        case 0: s = sr; c = cr; break;
        case 1: s = cr; c = -sr; break;
        case 2: s = -sr; c = -cr; break;
        default: s = -cr; c = sr; break;
    }
}

// Bilinear (clamp to edge, exact weights) -- the sampler of taa.hpp:274 at a position that is
// not a texel centre -- with the texel fetch `at(x, y)` (clamped coordinates) supplied by the
// caller, so that a staged copy and the image in global memory give the same bits.
// One axis of a bilinear sample: the two texels (clamped) and their weights t (second) and
// u = 1 - t (first). The axes are independent, so per-column and per-row tables of it give the
// same bits as evaluating it per sample.
struct BilinearAxis {
    int a, b;
    float t, u;
};
// `origin` receives the first tap before the edge clamp (pre-clamped to [-1, W] so that origin + 1
// cannot overflow for absurd positions)
__device__ __forceinline__ BilinearAxis bilinear_axis(int W, float u, int* origin = nullptr) {
    const float px = u * (float)W - 0.5f;
    const float fx = floorf(px);
    BilinearAxis r;
    r.t = px - fx;
    const int x0 = clampi(to_int(fx), -1, W);
    if (origin) *origin = x0;
    r.a = clampi(x0, 0, W - 1);
    r.b = clampi(x0 + 1, 0, W - 1);
    r.u = 1.0f - r.t;
    return r;
}
struct BilinearTaps {
    int xa, xb, ya, yb;
    float tx, ty, ux, uy;
};
__device__ __forceinline__ BilinearTaps bilinear_taps(const BilinearAxis& x, const BilinearAxis& y) {
    return BilinearTaps{x.a, x.b, y.a, y.b, x.t, y.t, x.u, y.u};
}
__device__ __forceinline__ BilinearTaps bilinear_taps(int W, int H, float u, float v) {
    return bilinear_taps(bilinear_axis(W, u), bilinear_axis(H, v));
}
// the filter of the four taps a (xa, ya), b (xb, ya), c (xa, yb), d (xb, yb): lerp in x, then in y
__device__ __forceinline__ float4 bilinear4_quad(float4 a, float4 b, float4 c, float4 d, const BilinearTaps& t) {
    const float tx = t.tx, ty = t.ty, ux = t.ux, uy = t.uy;
    float4 top = make_float4(a.x * ux + b.x * tx, a.y * ux + b.y * tx, a.z * ux + b.z * tx, a.w * ux + b.w * tx);
    float4 bot = make_float4(c.x * ux + d.x * tx, c.y * ux + d.y * tx, c.z * ux + d.z * tx, c.w * ux + d.w * tx);
    return make_float4(top.x * uy + bot.x * ty, top.y * uy + bot.y * ty, top.z * uy + bot.z * ty, top.w * uy + bot.w * ty);
}
__device__ __forceinline__ float bilinear1_quad(float a, float b, float c, float d, const BilinearTaps& t) {
    return (a * t.ux + b * t.tx) * t.uy + (c * t.ux + d * t.tx) * t.ty;
}
template <class F>
__device__ __forceinline__ float4 bilinear4_at(const BilinearTaps& t, F&& at) {
    return bilinear4_quad(at(t.xa, t.ya), at(t.xb, t.ya), at(t.xa, t.yb), at(t.xb, t.yb), t);
}
template <class F>
__device__ __forceinline__ float4 bilinear4_at(int W, int H, float u, float v, F&& at) {
    return bilinear4_at(bilinear_taps(W, H, u, v), at);
}
template <class F>
__device__ __forceinline__ float bilinear1_at(const BilinearTaps& t, F&& at) {
    return bilinear1_quad(at(t.xa, t.ya), at(t.xb, t.ya), at(t.xa, t.yb), at(t.xb, t.yb), t);
}
template <class F>
__device__ __forceinline__ float bilinear1_at(int W, int H, float u, float v, F&& at) {
    return bilinear1_at(bilinear_taps(W, H, u, v), at);
}

// over an RGBA image in global memory
template <bool S16>
__device__ float4 bilinear_rgba(const void* img, int W, int H, int pitch, float u, float v) {
    return bilinear4_at(W, H, u, v, [&](int x, int y) { return ld_rgba<S16>(img, (uint32_t)(y * pitch + x)); });
}

// Bilinear samples of the velocity (format fmt) and of a depth image at a non-centre uv, with the
// weights and the order of operations of bilinear_rgba: texture() of the upsampling path, where
// the output texel's uv is not a centre of the input-sized depth and velocity images.
__device__ float4 bilinear_vel(const void* img, int fmt, int W, int H, int pitch, float u, float v) {
    return bilinear4_at(W, H, u, v, [&](int x, int y) { return ld_vel(img, fmt, (uint32_t)(y * pitch + x)); });
}
__device__ float bilinear_f32(const float* img, int W, int H, int pitch, float u, float v) {
    return bilinear1_at(W, H, u, v, [&](int x, int y) { return ld_f32(img, (uint32_t)(y * pitch + x)); });
}

// getCurrentUpsampledColor (taa.comp:222-257), one axis: the jittered input sample position s
// (input texels) for output texel tc and corner offset o, and whether it lands in tc after scaling
// by h / l. l, h: input and output size; tj = jitter * l * -1 (input texels).
__device__ __forceinline__ float upsample_axis(float l, float h, float tj, int tc, float o, bool& hit) {
    const float sc = h / l;
    const float s = floorf(l * ((float)tc + o) / h) + 0.5f + tj;
    hit = to_int(s * sc) == tc;
    return s;
}
constexpr float kUpsampleCorner = 0.999999f;  // the second corner offset of the search (taa.comp:236-239)

// getCurrentUpsampledColor's search (taa.comp:236-257) on one axis of output texel tc: the two
// candidates at the corner offsets 0 and c, the second taken when it lands in the output texel,
// else the first. The reference's four corners are the products of the two axes' candidates and it
// keeps the last that lands on both (order (0,0), (c,0), (0,c), (c,c)): that is this pick on each
// axis whenever each axis has one. Returns the input texel it fetches (floor, clamped to the edge),
// or -1 when no candidate lands -- or, on the x axis (`x`), when the pick is negative, since the
// reference tests only foundTc.x >= 0.
__device__ __forceinline__ int upsample_pick(float l, float h, float tj, int tc, int n, bool x) {
    bool h0, h1;
    const float s0 = upsample_axis(l, h, tj, tc, 0.0f, h0);
    const float s1 = upsample_axis(l, h, tj, tc, kUpsampleCorner, h1);
    const float s = h1 ? s1 : s0;
    if (!(h0 || h1) || (x && !(s >= 0.0f))) return -1;
    return clampi(to_int(floorf(s)), 0, n - 1);
}

// RN(h * a) for the fp16 value h in the low (HI false) or high half of `bits`: one mixed FMA whose
// f16 operand converts exactly and whose addend is -0 (x + -0 = x for every x, the sign of a zero
// product included), so the same bits as converting h and then multiplying, in one instruction
template <bool HI>
__device__ __forceinline__ float mul_f16(uint32_t bits, float a) {
    float r;
    if constexpr (HI) asm("v_fma_mix_f32 %0, %1, %2, %3 op_sel:[1,0,0] op_sel_hi:[1,0,0]" : "=v"(r) : "v"(bits), "v"(a), "s"(-0.0f));
    else asm("v_fma_mix_f32 %0, %1, %2, %3 op_sel_hi:[1,0,0]" : "=v"(r) : "v"(bits), "v"(a), "s"(-0.0f));
    return r;
}

// Catmull-Rom history fetch (taa.comp:441-514), 16-tap texelFetch form with the edge clamp
// at size-1; identical to the 9-tap bilinear form up to rounding. Every tap is
// texel * w_x * w_y summed in the order of taa.comp:480-495. The alpha lane (read only by the
// dynamic anti-ghosting test) is accumulated when ALPHA; an interior footprint reads each row's
// four texels from one base address.
template <bool S16, bool ALPHA>
__device__ float4 history_catmullrom(const FrameDesc& fd, float hu, float hv) {
    const int W = fd.W, H = fd.H;
    float iTx = hu * (float)W, iTy = hv * (float)H;
    float fbx = floorf(iTx - 0.5f), fby = floorf(iTy - 0.5f);
    float tcx = fbx + 0.5f, tcy = fby + 0.5f;
    float fx = iTx - tcx, fy = iTy - tcy;
    float f2x = fx * fx, f2y = fy * fy, f3x = f2x * fx, f3y = f2y * fy;
    float wx[4] = {-0.5f * f3x + f2x - 0.5f * fx, 1.5f * f3x - 2.5f * f2x + 1.0f,
                   -1.5f * f3x + 2.0f * f2x + 0.5f * fx, 0.5f * f3x - 0.5f * f2x};
    float wy[4] = {-0.5f * f3y + f2y - 0.5f * fy, 1.5f * f3y - 2.5f * f2y + 1.0f,
                   -1.5f * f3y + 2.0f * f2y + 0.5f * fy, 0.5f * f3y - 0.5f * f2y};
    int kx = cr_base(to_int(fbx), W), ky = cr_base(to_int(fby), H);
    // cr_inside forms kx + 2 of the pre-clamped base. Written as kx <= W - 3 the same test put the
    // batched RGBA32F launch into a schedule 18% slower (4.88 -> 5.93 ms per 48-frame round).
    const bool interior = cr_inside(kx, W) && cr_inside(ky, H);
    int xs[4], ys[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        xs[i] = interior ? kx - 1 + i : clampi(kx - 1 + i, 0, W - 1);
        ys[i] = (interior ? ky - 1 + i : clampi(ky - 1 + i, 0, H - 1)) * fd.pitch;
    }
    float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
    // tap (i, j) given texel * wx[i] (each channel rounded): texel * wx[i] * wy[j], summed in order
    auto accum = [&](int i, int j, float4 ta) {
        const float b = wy[j];
        const float4 s = make_float4(ta.x * b, ta.y * b, ta.z * b, ALPHA ? ta.w * b : 0.0f);
        if (i == 0 && j == 0) acc = s;
        else acc = make_float4(acc.x + s.x, acc.y + s.y, acc.z + s.z, ALPHA ? acc.w + s.w : 0.0f);
    };
    auto add = [&](int i, int j, float4 t) {
        const float a = wx[i];
        accum(i, j, make_float4(t.x * a, t.y * a, t.z * a, ALPHA ? t.w * a : 0.0f));
    };
    // the same from the texel's two fp16 dwords, the conversion folded into the first product
    auto add16 = [&](int i, int j, uint32_t lo, uint32_t hi) {
        const float a = wx[i];
        accum(i, j, make_float4(mul_f16<false>(lo, a), mul_f16<true>(lo, a), mul_f16<false>(hi, a),
                                ALPHA ? mul_f16<true>(hi, a) : 0.0f));
    };
    if (S16 && interior) {  // each row's four fp16 texels as two 16-byte loads, all eight issued first
        u32x4_a8 rlo[4], rhi[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const uint32_t rb = (uint32_t)(ys[j] + xs[0]) * 8u;
            rlo[j] = gload<u32x4_a8>(fd.history, rb, 0);
            rhi[j] = gload<u32x4_a8>(fd.history, rb, 16);
        }
        // keep the eight loads ahead of the filter: a schedule that moves each load next to its
        // use exposes every load's latency in turn (measured 20% slower)
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const u32x4_a8 lo = rlo[j], hi = rhi[j];
            add16(0, j, lo.x, lo.y);
            add16(1, j, lo.z, lo.w);
            add16(2, j, hi.x, hi.y);
            add16(3, j, hi.z, hi.w);
        }
        return acc;
    }
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
        for (int i = 0; i < 4; ++i) add(i, j, ld_rgba<S16>(fd.history, (uint32_t)(ys[j] + xs[i])));
    return acc;
}

// B-spline history fetch (taa.comp:517-543): four bilinear taps.
template <bool S16>
__device__ float4 history_bspline(const FrameDesc& fd, float hu, float hv) {
    const int W = fd.W, H = fd.H;
    float sx = (float)W, sy = (float)H;
    float ix = 1.0f / sx, iy = 1.0f / sy;
    float iTx = hu * sx, iTy = hv * sy;
    float tcx = floorf(iTx - 0.5f) + 0.5f, tcy = floorf(iTy - 0.5f) + 0.5f;
    float fx = iTx - tcx, fy = iTy - tcy;
    float f2x = fx * fx, f2y = fy * fy, f3x = f2x * fx, f3y = f2y * fy;
    float w0x = f2x - 0.5f * (f3x + fx), w0y = f2y - 0.5f * (f3y + fy);
    float w1x = 1.5f * f3x - 2.5f * f2x + 1.0f, w1y = 1.5f * f3y - 2.5f * f2y + 1.0f;
    float w3x = 0.5f * (f3x - f2x), w3y = 0.5f * (f3y - f2y);
    float w2x = 1.0f - w0x - w1x - w3x, w2y = 1.0f - w0y - w1y - w3y;
    float s0x = w0x + w1x, s0y = w0y + w1y, s1x = w2x + w3x, s1y = w2y + w3y;
    float f0x = w1x / (w0x + w1x), f0y = w1y / (w0y + w1y);
    float f1x = w3x / (w2x + w3x), f1y = w3y / (w2y + w3y);
    float t0x = (tcx - 1.0f + f0x) * ix, t0y = (tcy - 1.0f + f0y) * iy;
    float t1x = (tcx + 1.0f + f1x) * ix, t1y = (tcy + 1.0f + f1y) * iy;
    float4 a = bilinear_rgba<S16>(fd.history, W, H, fd.pitch, t0x, t0y);
    float4 b = bilinear_rgba<S16>(fd.history, W, H, fd.pitch, t1x, t0y);
    float4 c = bilinear_rgba<S16>(fd.history, W, H, fd.pitch, t0x, t1y);
    float4 d = bilinear_rgba<S16>(fd.history, W, H, fd.pitch, t1x, t1y);
    float4 r0 = make_float4(a.x * s0x + b.x * s1x, a.y * s0x + b.y * s1x, a.z * s0x + b.z * s1x, a.w * s0x + b.w * s1x);
    float4 r1 = make_float4(c.x * s0x + d.x * s1x, c.y * s0x + d.y * s1x, c.z * s0x + d.z * s1x, c.w * s0x + d.w * s1x);
    return make_float4(r0.x * s0y + r1.x * s1y, r0.y * s0y + r1.y * s1y, r0.z * s0y + r1.z * s1y, r0.w * s0y + r1.w * s1y);
}

// taa.comp:323-345 with p = (0,0,0,1)
__device__ __forceinline__ f3 clipAabb(f3 mn, f3 mx, f3 q) {
    const float eps = 1e-7f;
    f3 pClip = 0.5f * (mx + mn);
    f3 eClip = 0.5f * (mx - mn) + f3{eps, eps, eps};
    f3 vClip = q - pClip;
    f3 vUnit = vClip / eClip;
    float ma = fmaxf(fabsf(vUnit.x), fmaxf(fabsf(vUnit.y), fabsf(vUnit.z)));
    if (ma > 1.0f) return pClip + vClip / ma;
    return q;
}
// taa.comp:348-369
__device__ __forceinline__ f3 clipAabbSlow(f3 mn, f3 mx, f3 p, f3 q) {
    f3 r = q - p, rmax = mx - p, rmin = mn - p;
    const float eps = 1e-7f;
    if (r.x > rmax.x + eps) r = r * (rmax.x / r.x);
    if (r.y > rmax.y + eps) r = r * (rmax.y / r.y);
    if (r.z > rmax.z + eps) r = r * (rmax.z / r.z);
    if (r.x < rmin.x - eps) r = r * (rmin.x / r.x);
    if (r.y < rmin.y - eps) r = r * (rmin.y / r.y);
    if (r.z < rmin.z - eps) r = r * (rmin.z / r.z);
    return p + r;
}

// ------------------------------------------------------------------------------------------
// CAS (ffx_cas.h:424-537, sharpen only, green weight for all channels)

__device__ __forceinline__ float prx_lo_sqrt(float a) { return __uint_as_float((__float_as_uint(a) >> 1u) + 0x1fbc4639u); }
__device__ __forceinline__ float prx_lo_rcp(float a) { return __uint_as_float(0x7ef07ebbu - __float_as_uint(a)); }
__device__ __forceinline__ float prx_med_rcp(float a) {
    float b = __uint_as_float(0x7ef19fffu - __float_as_uint(a));
    return b * (-b * a + 2.0f);
}
__device__ __forceinline__ float sat(float a) { return fminf(fmaxf(a, 0.0f), 1.0f); }
__device__ __forceinline__ float min3(float x, float y, float z) { return fminf(x, fminf(y, z)); }
__device__ __forceinline__ float max3(float x, float y, float z) { return fmaxf(x, fmaxf(y, z)); }

__device__ __forceinline__ f3 cas_filter(f3 b, f3 d, f3 e, f3 f, f3 h, float peak) {
    float mxG = max3(max3(d.y, e.y, f.y), b.y, h.y);
    float mnG = min3(min3(d.y, e.y, f.y), b.y, h.y);
    float rcpMG = prx_lo_rcp(mxG);
    float ampG = sat(fminf(mnG, 1.0f - mxG) * rcpMG);
    ampG = prx_lo_sqrt(ampG);
    float wG = ampG * peak;
    float rcpWeight = prx_med_rcp(1.0f + 4.0f * wG);
    f3 o;
    o.x = sat((b.x * wG + d.x * wG + f.x * wG + h.x * wG + e.x) * rcpWeight);
    o.y = sat((b.y * wG + d.y * wG + f.y * wG + h.y * wG + e.y) * rcpWeight);
    o.z = sat((b.z * wG + d.z * wG + f.z * wG + h.z * wG + e.z) * rcpWeight);
    return o;
}


// Workgroup -> tile so that XCD k (blockIdx % 8) gets the contiguous k-th eighth of the tiles.
// Entry of a batch descriptor table that covers tile t (the last entry with tile_begin <= t).
// The table is read-only for the kernel's lifetime: reading it through the constant address
// space gives scalar loads, and the image pointers in it are known to be global memory.
__device__ __forceinline__ const BatchEntry& batch_entry(const BatchEntry* entries, int n_entries, int t) {
    const TAA_CONSTANT BatchEntry* ents = (const TAA_CONSTANT BatchEntry*)entries;
    int lo = 0, hi = n_entries - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (ents[mid].tile_begin <= t) lo = mid;
        else hi = mid - 1;
    }
    return *(const BatchEntry*)(ents + lo);
}

__device__ __forceinline__ int xcd_tile(int b, int n) { return xcd_tile_of(b, n); }

// The tile of a batched launch's workgroup: the grid is batch_grid(total_tiles) = 8 * per
// workgroups, one tile each, and XCD k (blockIdx % 8) takes the k-th contiguous eighth of the
// flattened tile space. -1: a workgroup of the rounding, with no tile.
__device__ __forceinline__ int batch_tile(int total_tiles) { return batch_tile_of((int)blockIdx.x, total_tiles); }

}  // namespace
}  // namespace taastar
