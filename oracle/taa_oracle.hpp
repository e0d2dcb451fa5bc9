// TEST INFRASTRUCTURE -- the oracle. Only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg may load it, and only as the checker / CPU baseline, never as the product.
//
// Scalar C++ restatement of the temporal resolve of cg-tuwien/TAA-STAR `shaders/taa.comp` and of
// the FidelityFX CAS filter it is followed by (`shaders/sharpen_cas.comp`, `shaders/ffx_cas.h`,
// `shaders/ffx_a.h`), templated on the arithmetic type:
//   T = double : the FP64 oracle,
//   T = float  : the plain FP32 CPU transcription (compile with -ffp-contract=off).
// Every GLSL expression is evaluated in the order it is written, component by component,
// with no contraction; GLSL built-ins follow the GLSL 4.60 definitions
// (mix(x,y,a) = x*(1-a) + y*a, clamp(x,lo,hi) = min(max(x,lo),hi)).
//
// Where taa.comp relies on Vulkan's sampler or image units the oracle defines the behaviour
// (the GPU path uses the same definitions):
//   * texture() at an exact texel centre (no unjitter offset) = that texel, clamped to the edge;
//     texture() at any other position = bilinear with exact (unquantised) weights, clamp to edge.
//   * texelFetch / imageLoad outside the image read the edge texel (clamp).
//   * RG velocity formats read back as (r, g, 0, 1).
//   * int(x) for float x: NaN -> 0, saturate to int range, truncate toward zero.
// Parity status per function is listed in the header of oracle_capi.cpp.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>

#include "taastar/taa_abi.h"
#include "taastar/taa_velocity.h"

namespace oracle {

// ------------------------------------------------------------------------------------------
// small vector types with GLSL component-wise semantics
template <class T> struct v2 { T x, y; };
template <class T> struct v3 { T x, y, z; };
template <class T> struct v4 { T x, y, z, w; };
struct iv2 { int x, y; };

template <class T> inline v3<T> operator+(v3<T> a, v3<T> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class T> inline v3<T> operator-(v3<T> a, v3<T> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class T> inline v3<T> operator*(v3<T> a, v3<T> b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
template <class T> inline v3<T> operator/(v3<T> a, v3<T> b) { return {a.x / b.x, a.y / b.y, a.z / b.z}; }
template <class T> inline v3<T> operator*(T s, v3<T> a) { return {s * a.x, s * a.y, s * a.z}; }
template <class T> inline v3<T> operator*(v3<T> a, T s) { return {a.x * s, a.y * s, a.z * s}; }
template <class T> inline v3<T> operator/(v3<T> a, T s) { return {a.x / s, a.y / s, a.z / s}; }
template <class T> inline v4<T> operator+(v4<T> a, v4<T> b) { return {a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w}; }
template <class T> inline v4<T> operator-(v4<T> a, v4<T> b) { return {a.x - b.x, a.y - b.y, a.z - b.z, a.w - b.w}; }
template <class T> inline v4<T> operator*(v4<T> a, T s) { return {a.x * s, a.y * s, a.z * s, a.w * s}; }
template <class T> inline v4<T> operator/(v4<T> a, T s) { return {a.x / s, a.y / s, a.z / s, a.w / s}; }

// GLSL min/max/clamp (IEEE minNum/maxNum: a NaN operand yields the other operand)
template <class T> inline T gmin(T a, T b) { return std::fmin(a, b); }
template <class T> inline T gmax(T a, T b) { return std::fmax(a, b); }
template <class T> inline T gclamp(T x, T lo, T hi) { return gmin(gmax(x, lo), hi); }
template <class T> inline v3<T> vmin(v3<T> a, v3<T> b) { return {gmin(a.x, b.x), gmin(a.y, b.y), gmin(a.z, b.z)}; }
template <class T> inline v3<T> vmax(v3<T> a, v3<T> b) { return {gmax(a.x, b.x), gmax(a.y, b.y), gmax(a.z, b.z)}; }
template <class T> inline T gmix(T x, T y, T a) { return x * (T(1) - a) + y * a; }
template <class T> inline v3<T> vmix(v3<T> x, v3<T> y, T a) { return {gmix(x.x, y.x, a), gmix(x.y, y.y, a), gmix(x.z, y.z, a)}; }

// int(x) as defined above (GLSL leaves out-of-range conversions undefined)
template <class T> inline int to_int(T v) {
    if (v != v) return 0;
    if (v >= T(2147483647.0)) return 2147483647;
    if (v <= T(-2147483648.0)) return (-2147483647 - 1);
    return (int)v;
}
inline int iclamp(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// ------------------------------------------------------------------------------------------
// one frame's images, decoded to float (storage rounding is applied by the caller)
template <class T> struct Frame {
    int W = 0, H = 0;               // output size: uHistoryFrame, uResultScreen, uResultHistory
    int Wl = 0, Hl = 0;             // input size: uCurrentFrame, uCurrentDepth, uCurrentVelocity,
                                    // uHistoryDepth (the render resolution; == W, H without TAAU)
    const float* color = nullptr;   // RGBA  uCurrentFrame
    const float* depth = nullptr;   // D     uCurrentDepth
    const float* vel = nullptr;     // RGBA  uCurrentVelocity (RG formats expanded to r,g,0,1)
    const T* hist = nullptr;        // RGBA  uHistoryFrame
    const float* hdepth = nullptr;  // D     uHistoryDepth
    v4<T> color_at(int x, int y) const {
        x = iclamp(x, 0, Wl - 1); y = iclamp(y, 0, Hl - 1);
        const float* p = color + 4 * ((size_t)y * Wl + x);
        return {T(p[0]), T(p[1]), T(p[2]), T(p[3])};
    }
    v4<T> vel_at(int x, int y) const {
        x = iclamp(x, 0, Wl - 1); y = iclamp(y, 0, Hl - 1);
        const float* p = vel + 4 * ((size_t)y * Wl + x);
        return {T(p[0]), T(p[1]), T(p[2]), T(p[3])};
    }
    v4<T> hist_at(int x, int y) const {
        x = iclamp(x, 0, W - 1); y = iclamp(y, 0, H - 1);
        const T* p = hist + 4 * ((size_t)y * W + x);
        return {p[0], p[1], p[2], p[3]};
    }
    T depth_at(int x, int y) const {
        x = iclamp(x, 0, Wl - 1); y = iclamp(y, 0, Hl - 1);
        return T(depth[(size_t)y * Wl + x]);
    }
    T hdepth_at(int x, int y) const {
        x = iclamp(x, 0, Wl - 1); y = iclamp(y, 0, Hl - 1);
        return T(hdepth[(size_t)y * Wl + x]);
    }
    // segmentation inputs; material == nullptr: no mask. The renderer's images have the input
    // size (Wl x Hl), the masks the output size (taa.hpp:336-338 creates them at the target size)
    const uint32_t* material = nullptr;      // R32UI  uCurrentMaterial   (input size)
    const uint32_t* prev_material = nullptr; // R32UI  uPreviousMaterial  (input size)
    const float* uv_normal = nullptr;        // RGBA32F uCurrentUvNrm     (input size)
    const uint32_t* prev_segmask = nullptr;  // R32UI  uPreviousSegMask   (output size)
    // an input-sized R32UI image at (x, y) clamped to its edge (CLAMP_TO_TEX, taa.comp:42)
    uint32_t u32_in(const uint32_t* img, int x, int y) const {
        x = iclamp(x, 0, Wl - 1); y = iclamp(y, 0, Hl - 1);
        return img[(size_t)y * Wl + x];
    }
    // ... at (x, y) unclamped: imageLoad outside the image, which Vulkan leaves undefined unless the
    // device bounds-checks image accesses; the value robustImageAccess2 defines, 0, is taken
    uint32_t u32_in_or_zero(const uint32_t* img, int x, int y) const {
        if (x < 0 || y < 0 || x >= Wl || y >= Hl) return 0u;
        return img[(size_t)y * Wl + x];
    }
    // an output-sized R32UI image at (x, y) clamped to its edge
    uint32_t u32_out(const uint32_t* img, int x, int y) const {
        x = iclamp(x, 0, W - 1); y = iclamp(y, 0, H - 1);
        return img[(size_t)y * W + x];
    }
};

// Bilinear filter with clamp-to-edge (the sampler of source/taa.hpp:274) at normalised uv,
// exact weights: p = uv*size - 0.5, lerp in x then in y.
template <class T, class Fetch>
inline v4<T> bilinear(Fetch fetch, int W, int H, v2<T> uv) {
    T px = uv.x * T(W) - T(0.5), py = uv.y * T(H) - T(0.5);
    T fx = std::floor(px), fy = std::floor(py);
    T tx = px - fx, ty = py - fy;
    // pre-clamp to [-1, size]: the taps are clamped to the edge anyway, this only keeps x0 + 1
    // from overflowing for absurd positions
    int x0 = iclamp(to_int(fx), -1, W), y0 = iclamp(to_int(fy), -1, H);
    v4<T> a = fetch(x0, y0), b = fetch(x0 + 1, y0), c = fetch(x0, y0 + 1), d = fetch(x0 + 1, y0 + 1);
    T ux = T(1) - tx, uy = T(1) - ty;
    v4<T> top = a * ux + b * tx;
    v4<T> bot = c * ux + d * tx;
    return top * uy + bot * ty;
}

// ------------------------------------------------------------------------------------------
// helpers of taa.comp

// taa.comp:131
template <class T> inline v2<T> tc_to_uv(iv2 tc, iv2 size) {
    return {(T(tc.x) + T(0.5)) / T(size.x), (T(tc.y) + T(0.5)) / T(size.y)};
}
// taa.comp:132
template <class T> inline iv2 uv_to_tc(v2<T> uv, iv2 size) {
    return {to_int(uv.x * T(size.x)), to_int(uv.y * T(size.y))};
}
// taa.comp:158-164
template <class T> inline v3<T> rgb_to_ycocg(v3<T> c) {
    return {T(.25) * c.x + T(.5) * c.y + T(.25) * c.z,
            T(.5) * c.x - T(.5) * c.z,
            T(-.25) * c.x + T(.5) * c.y - T(.25) * c.z};
}
// taa.comp:167-174
template <class T> inline v3<T> ycocg_to_rgb(v3<T> c) {
    T tmp = c.x - c.z;
    return {tmp + c.y, c.x + c.z, tmp - c.y};
}
// taa.comp:177-179
template <class T> inline v3<T> maybe_rgb_to_ycocg(const taa_params_t& P, v3<T> c) { return P.mUseYCoCg ? rgb_to_ycocg(c) : c; }
template <class T> inline v3<T> maybe_ycocg_to_rgb(const taa_params_t& P, v3<T> c) { return P.mUseYCoCg ? ycocg_to_rgb(c) : c; }
template <class T> inline T luminance(const taa_params_t& P, v3<T> c) { return P.mUseYCoCg ? c.x : rgb_to_ycocg(c).x; }
// taa.comp:182-197
template <class T> inline v3<T> tonemap_rgb(const taa_params_t& P, v3<T> hdr) {
    if (P.mToneMapLumaKaris) {
        T luma = gmax(gmax(hdr.x, hdr.y), hdr.z);
        return hdr / (T(1.0) + luma);
    }
    return hdr;
}
template <class T> inline v3<T> un_tonemap_rgb(const taa_params_t& P, v3<T> ldr) {
    if (P.mToneMapLumaKaris) {
        T luma = gmax(gmax(ldr.x, ldr.y), ldr.z);
        return ldr / (T(1.0) - luma);
    }
    return ldr;
}
// taa.comp:559-561
template <class T> inline T linearize_depth(T d, T zNear, T zFar) {
    return zNear * zFar / (zFar + d * (zNear - zFar));
}

template <class T> inline v3<T> xyz(v4<T> v) { return {v.x, v.y, v.z}; }

// current-colour tap: texture(uCurrentFrame, offset + (iuv + d + 0.5) * invsize).rgb
// (taa.comp:204-212, 219). With no unjitter offset this lands on a texel centre.
template <class T>
inline v3<T> current_tap(const Frame<T>& F, iv2 iuv, int dx, int dy, bool unjitter, v2<T> jitter_uv) {
    if (!unjitter) return xyz(F.color_at(iuv.x + dx, iuv.y + dy));
    T invx = T(1) / T(F.Wl), invy = T(1) / T(F.Hl);
    v2<T> uv = {jitter_uv.x + (T(iuv.x + dx) + T(0.5)) * invx, jitter_uv.y + (T(iuv.y + dy) + T(0.5)) * invy};
    auto fetch = [&](int x, int y) { return F.color_at(x, y); };
    return xyz(bilinear<T>(fetch, F.Wl, F.Hl, uv));
}

// taa.comp:199-213 (getNeighbourhood) + taa.comp:259-319 (getColorAndAabb)
template <class T>
inline void getColorAndAabb(const taa_params_t& P, const Frame<T>& F, iv2 iuv, v2<T> jitter_uv,
                            v3<T>& centerCol, v3<T>& minCol, v3<T>& maxCol, v3<T>& cliptowardsCol) {
    const T N = T(9.0);
    bool uj = P.mUnjitterNeighbourhood != 0;
    auto tap = [&](int dx, int dy) {
        return maybe_rgb_to_ycocg(P, tonemap_rgb(P, current_tap(F, iuv, dx, dy, uj, jitter_uv)));
    };
    v3<T> cC = tap(0, 0);
    v3<T> c1 = tap(-1, -1), c2 = tap(0, -1), c3 = tap(1, -1), c4 = tap(-1, 0);
    v3<T> c5 = tap(1, 0), c6 = tap(-1, 1), c7 = tap(0, 1), c8 = tap(1, 1);
    centerCol = cC;
    if (P.mVarianceClipping) {
        v3<T> m1 = cC + c1 + c2 + c3 + c4 + c5 + c6 + c7 + c8;
        v3<T> m2 = cC * cC + c1 * c1 + c2 * c2 + c3 * c3 + c4 * c4 + c5 * c5 + c6 * c6 + c7 * c7 + c8 * c8;
        v3<T> mean = m1 / N;
        v3<T> var = m2 / N - mean * mean;
        v3<T> sigma = {std::sqrt(gmax(T(0), var.x)), std::sqrt(gmax(T(0), var.y)), std::sqrt(gmax(T(0), var.z))};
        T g = T(P.mVarClipGamma);
        minCol = mean - g * sigma;
        maxCol = mean + g * sigma;
        cliptowardsCol = mean;
    } else if (P.mShapedNeighbourhood) {
        v3<T> mn33 = vmin(vmin(vmin(vmin(vmin(vmin(vmin(vmin(cC, c1), c2), c3), c4), c5), c6), c7), c8);
        v3<T> mx33 = vmax(vmax(vmax(vmax(vmax(vmax(vmax(vmax(cC, c1), c2), c3), c4), c5), c6), c7), c8);
        v3<T> mn5 = vmin(vmin(vmin(vmin(cC, c2), c4), c5), c7);
        v3<T> mx5 = vmax(vmax(vmax(vmax(cC, c2), c4), c5), c7);
        minCol = (mn33 + mn5) * T(0.5);
        maxCol = (mx33 + mx5) * T(0.5);
        cliptowardsCol = cC;
    } else {
        minCol = vmin(vmin(vmin(vmin(vmin(vmin(vmin(vmin(cC, c1), c2), c3), c4), c5), c6), c7), c8);
        maxCol = vmax(vmax(vmax(vmax(vmax(vmax(vmax(vmax(cC, c1), c2), c3), c4), c5), c6), c7), c8);
        cliptowardsCol = cC;
    }
    if (P.mUseYCoCg && P.mShrinkChromaAxis) {  // taa.comp:294-314
        const v3<T> s = {T(0.5) * T(1.0), T(0.5) * T(0.5), T(0.5) * T(0.5)};
        v3<T> halfSize = s * (maxCol - minCol);
        v3<T> center = (minCol + maxCol) * T(0.5);
        minCol = center - halfSize;
        maxCol = center + halfSize;
        bool out = cliptowardsCol.x < minCol.x || cliptowardsCol.y < minCol.y || cliptowardsCol.z < minCol.z ||
                   cliptowardsCol.x > maxCol.x || cliptowardsCol.y > maxCol.y || cliptowardsCol.z > maxCol.z;
        if (out) cliptowardsCol = center;
    }
}

// taa.comp:323-345 (called with p = (0,0,0,1), so the .w lane is constant)
template <class T> inline v3<T> clipAabb(v3<T> mn, v3<T> mx, v3<T> q) {
    const T eps = T(1e-7);
    v3<T> pClip = T(0.5) * (mx + mn);
    v3<T> eClip = T(0.5) * (mx - mn) + v3<T>{eps, eps, eps};
    v3<T> vClip = q - pClip;
    v3<T> vUnit = vClip / eClip;
    v3<T> a = {std::fabs(vUnit.x), std::fabs(vUnit.y), std::fabs(vUnit.z)};
    T ma = gmax(a.x, gmax(a.y, a.z));
    if (ma > T(1.0)) return pClip + vClip / ma;
    return q;
}

// taa.comp:348-369 (p = (towards, 1), q = (hist, 1): the .w lane of r is 0 throughout)
template <class T> inline v3<T> clipAabbSlow(v3<T> mn, v3<T> mx, v3<T> p, v3<T> q) {
    v3<T> r = q - p;
    v3<T> rmax = mx - p;
    v3<T> rmin = mn - p;
    const T eps = T(1e-7);
    if (r.x > rmax.x + eps) r = r * (rmax.x / r.x);
    if (r.y > rmax.y + eps) r = r * (rmax.y / r.y);
    if (r.z > rmax.z + eps) r = r * (rmax.z / r.z);
    if (r.x < rmin.x - eps) r = r * (rmin.x / r.x);
    if (r.y < rmin.y - eps) r = r * (rmin.y / r.y);
    if (r.z < rmin.z - eps) r = r * (rmin.z / r.z);
    return p + r;
}

// taa.comp:371-389. Returns the chosen tap index 0..8 (row-major from (-1,-1)); ties keep the
// first tap in scan order (strict <).
template <class T> inline int findClosestTap_3x3(const Frame<T>& F, iv2 iuv, T& dClosest) {
    static const int off[9][2] = {{-1, -1}, {0, -1}, {1, -1}, {-1, 0}, {0, 0}, {1, 0}, {-1, 1}, {0, 1}, {1, 1}};
    int best = 0;
    dClosest = F.depth_at(iuv.x - 1, iuv.y - 1);
    for (int k = 1; k < 9; ++k) {
        T d = F.depth_at(iuv.x + off[k][0], iuv.y + off[k][1]);
        if (d < dClosest) { best = k; dClosest = d; }
    }
    return best;
}

// taa.comp:441-514, the 9-tap form: bilinear taps at tc-1, tc + w2/(w1+w2), tc+2.
template <class T> inline v4<T> catmullrom_9tap(const Frame<T>& F, v2<T> uv) {
    int W = F.W, H = F.H;
    auto fetch = [&](int x, int y) { return F.hist_at(x, y); };
    T sx = T(W), sy = T(H);
    T ix = T(1.0) / sx, iy = T(1.0) / sy;
    T iTx = uv.x * sx, iTy = uv.y * sy;
    T tcx = std::floor(iTx - T(0.5)) + T(0.5), tcy = std::floor(iTy - T(0.5)) + T(0.5);
    T fx = iTx - tcx, fy = iTy - tcy;
    T f2x = fx * fx, f2y = fy * fy, f3x = f2x * fx, f3y = f2y * fy;
    T w0x = T(-0.5) * f3x + f2x - T(0.5) * fx, w0y = T(-0.5) * f3y + f2y - T(0.5) * fy;
    T w1x = T(1.5) * f3x - T(2.5) * f2x + T(1.0), w1y = T(1.5) * f3y - T(2.5) * f2y + T(1.0);
    T w2x = T(-1.5) * f3x + T(2.0) * f2x + T(0.5) * fx, w2y = T(-1.5) * f3y + T(2.0) * f2y + T(0.5) * fy;
    T w3x = T(0.5) * f3x - T(0.5) * f2x, w3y = T(0.5) * f3y - T(0.5) * f2y;
    T wCx = w1x + w2x, wCy = w1y + w2y;
    v2<T> t0 = {(tcx - T(1)) * ix, (tcy - T(1)) * iy};
    v2<T> tC = {(tcx + w2x / wCx) * ix, (tcy + w2y / wCy) * iy};
    v2<T> t3 = {(tcx + T(2)) * ix, (tcy + T(2)) * iy};
    auto S = [&](T u, T v) { return bilinear<T>(fetch, W, H, v2<T>{u, v}); };
    return S(t0.x, t0.y) * w0x * w0y + S(tC.x, t0.y) * wCx * w0y + S(t3.x, t0.y) * w3x * w0y +
           S(t0.x, tC.y) * w0x * wCy + S(tC.x, tC.y) * wCx * wCy + S(t3.x, tC.y) * w3x * wCy +
           S(t0.x, t3.y) * w0x * w3y + S(tC.x, t3.y) * wCx * w3y + S(t3.x, t3.y) * w3x * w3y;
}

// taa.comp:469-495, the 16-tap texelFetch form, with the clamp to [0, size-1] (the reference's
// `#if 0` block clamps to size, one past the edge). Base texel k = floor(uv*size - 0.5).
template <class T> inline v4<T> catmullrom_16tap(const Frame<T>& F, v2<T> uv, iv2* base = nullptr) {
    T sx = T(F.W), sy = T(F.H);
    T iTx = uv.x * sx, iTy = uv.y * sy;
    T tcx = std::floor(iTx - T(0.5)) + T(0.5), tcy = std::floor(iTy - T(0.5)) + T(0.5);
    T fx = iTx - tcx, fy = iTy - tcy;
    T f2x = fx * fx, f2y = fy * fy, f3x = f2x * fx, f3y = f2y * fy;
    T wx[4] = {T(-0.5) * f3x + f2x - T(0.5) * fx, T(1.5) * f3x - T(2.5) * f2x + T(1.0),
               T(-1.5) * f3x + T(2.0) * f2x + T(0.5) * fx, T(0.5) * f3x - T(0.5) * f2x};
    T wy[4] = {T(-0.5) * f3y + f2y - T(0.5) * fy, T(1.5) * f3y - T(2.5) * f2y + T(1.0),
               T(-1.5) * f3y + T(2.0) * f2y + T(0.5) * fy, T(0.5) * f3y - T(0.5) * f2y};
    int kx = to_int(std::floor(iTx - T(0.5))), ky = to_int(std::floor(iTy - T(0.5)));
    if (base) *base = {kx, ky};
    kx = iclamp(kx, -2, F.W + 1);  // result-preserving (taps are edge-clamped), avoids overflow
    ky = iclamp(ky, -2, F.H + 1);
    v4<T> acc = {T(0), T(0), T(0), T(0)};
    bool first = true;
    for (int j = 0; j < 4; ++j)
        for (int i = 0; i < 4; ++i) {
            v4<T> t = F.hist_at(kx - 1 + i, ky - 1 + j) * wx[i] * wy[j];
            acc = first ? t : acc + t;
            first = false;
        }
    return acc;
}

// taa.comp:517-543
template <class T> inline v4<T> bspline_4tap(const Frame<T>& F, v2<T> uv) {
    int W = F.W, H = F.H;
    auto fetch = [&](int x, int y) { return F.hist_at(x, y); };
    T sx = T(W), sy = T(H);
    T ix = T(1.0) / sx, iy = T(1.0) / sy;
    T iTx = uv.x * sx, iTy = uv.y * sy;
    T tcx = std::floor(iTx - T(0.5)) + T(0.5), tcy = std::floor(iTy - T(0.5)) + T(0.5);
    T fx = iTx - tcx, fy = iTy - tcy;
    T f2x = fx * fx, f2y = fy * fy, f3x = f2x * fx, f3y = f2y * fy;
    T w0x = f2x - T(0.5) * (f3x + fx), w0y = f2y - T(0.5) * (f3y + fy);
    T w1x = T(1.5) * f3x - T(2.5) * f2x + T(1.0), w1y = T(1.5) * f3y - T(2.5) * f2y + T(1.0);
    T w3x = T(0.5) * (f3x - f2x), w3y = T(0.5) * (f3y - f2y);
    T w2x = T(1.0) - w0x - w1x - w3x, w2y = T(1.0) - w0y - w1y - w3y;
    T s0x = w0x + w1x, s0y = w0y + w1y, s1x = w2x + w3x, s1y = w2y + w3y;
    T f0x = w1x / (w0x + w1x), f0y = w1y / (w0y + w1y);
    T f1x = w3x / (w2x + w3x), f1y = w3y / (w2y + w3y);
    v2<T> t0 = {(tcx - T(1) + f0x) * ix, (tcy - T(1) + f0y) * iy};
    v2<T> t1 = {(tcx + T(1) + f1x) * ix, (tcy + T(1) + f1y) * iy};
    auto S = [&](T u, T v) { return bilinear<T>(fetch, W, H, v2<T>{u, v}); };
    return (S(t0.x, t0.y) * s0x + S(t1.x, t0.y) * s1x) * s0y + (S(t0.x, t1.y) * s0x + S(t1.x, t1.y) * s1x) * s1y;
}

// Output of one pixel of taa.comp main.
template <class T> struct PixelOut {
    v4<T> screen;   // uResultScreen
    v4<T> history;  // uResultHistory
    v4<T> debug;    // uDebug
    int32_t inspect[TAA_INSPECT_WORDS];
    uint32_t segmask = 0;     // uSegMask
    bool seg_written = false; // the reference stores uSegMask for this texel
};

template <class T> inline void put_f(int32_t* w, int i, T v) { float f = float(v); std::memcpy(&w[i], &f, 4); }

// getCurrentUpsampledColor (taa.comp:222-257): the input sample whose jittered position falls into
// output texel currentTc after scaling, else beta = 0. The chosen sample is fetched at its texel
// (texture() at the texel centre).
template <class T>
inline v3<T> current_upsampled(const taa_params_t& P, const Frame<T>& F, iv2 currentTc, v2<T> jitter_uv, T& beta) {
    const T lx = T(F.Wl), ly = T(F.Hl), hx = T(F.W), hy = T(F.H);
    const v2<T> scale = {hx / lx, hy / ly};
    const v2<T> texelJitter = {jitter_uv.x * lx * T(-1), jitter_uv.y * ly * T(-1)};
    const T almostOne = T(0.999999f);
    v2<T> foundTc = {T(-1), T(-1)};
    const T offs[4][2] = {{T(0), T(0)}, {almostOne, T(0)}, {T(0), almostOne}, {almostOne, almostOne}};
    for (int k = 0; k < 4; ++k) {
        v2<T> s = {std::floor(lx * (T(currentTc.x) + offs[k][0]) / hx) + T(0.5) + texelJitter.x,
                   std::floor(ly * (T(currentTc.y) + offs[k][1]) / hy) + T(0.5) + texelJitter.y};
        if (to_int(s.x * scale.x) == currentTc.x && to_int(s.y * scale.y) == currentTc.y) foundTc = s;
    }
    if (foundTc.x >= T(0)) {
        beta = T(1);
        return maybe_rgb_to_ycocg(P, tonemap_rgb(P, xyz(F.color_at(to_int(std::floor(foundTc.x)), to_int(std::floor(foundTc.y))))));
    }
    beta = T(0);
    return {T(0), T(0), T(0)};
}

// ---- segmentation mask (taa.comp:566-702) ---------------------------------------------------

// sin and cos of the G-buffer's normal angles (taa.comp:574-579). GLSL leaves their precision to
// the implementation; FP64 takes the library functions, FP32 a fixed evaluation the kernels
// repeat operation for operation: Cody-Waite reduction by pi/2 with a three-part constant
// (Cody & Waite, Software Manual for the Elementary Functions, 1980), then the Taylor
// polynomials of sin (degree 9) and cos (degree 10) on [-pi/4, pi/4] in Horner form, all with FMA.
inline void seg_sincos(double a, double& s, double& c) { s = std::sin(a); c = std::cos(a); }
inline void seg_sincos(float a, float& s, float& c) {
    const float k = std::rint(a * 0x1.45f306p-1f);  // a * 2/pi, nearest even
    float r = std::fma(-k, 0x1.921fb6p+0f, a);
    r = std::fma(-k, -0x1.777a5cp-25f, r);
    r = std::fma(-k, -0x1.ee59dap-50f, r);
    const float r2 = r * r;
    float ps = std::fma(0x1.71de3ap-19f, r2, -0x1.a01a02p-13f);  // 1/9!, -1/7!
    ps = std::fma(ps, r2, 0x1.111112p-7f);                        // 1/5!
    ps = std::fma(ps, r2, -0x1.555556p-3f);                       // -1/3!
    const float sr = std::fma(r * r2, ps, r);
    float pc = std::fma(-0x1.27e4fcp-22f, r2, 0x1.a01a02p-16f);  // -1/10!, 1/8!
    pc = std::fma(pc, r2, -0x1.6c16c2p-10f);                      // -1/6!
    pc = std::fma(pc, r2, 0x1.555556p-5f);                        // 1/4!
    pc = std::fma(pc, r2, -0.5f);                                 // -1/2!
    const float cr = std::fma(r2, pc, 1.0f);
    switch (to_int(k) & 3) {
        case 0: s = sr; c = cr; break;
        case 1: s = cr; c = -sr; break;
        case 2: s = -sr; c = -cr; break;
        default: s = -cr; c = sr; break;
    }
}

// sample_normal (taa.comp:574-579): (cos z cos w, sin z cos w, sin w) of the input-sized G-buffer
// texel (x, y); `clamp`: the texel clamped to the edge (the neighbours, read through CLAMP_TO_TEX),
// else a texel outside the image reads (0, 0, 0, 0) as robustImageAccess2 defines (the centre, read
// at the output texel's own coordinates, taa.comp:641, which upsampling puts outside the input)
template <class T> inline v3<T> seg_normal(const Frame<T>& F, int x, int y, bool clamp = true) {
    float z = 0.0f, w = 0.0f;
    if (clamp) { x = iclamp(x, 0, F.Wl - 1); y = iclamp(y, 0, F.Hl - 1); }
    if (x >= 0 && y >= 0 && x < F.Wl && y < F.Hl) {
        const float* p = F.uv_normal + 4 * ((size_t)y * F.Wl + x);
        z = p[2];
        w = p[3];
    }
    T sz, cz, sw, cw;
    seg_sincos(T(z), sz, cz);
    seg_sincos(T(w), sw, cw);
    return {cz * cw, sz * cw, sw};
}

// sobel (taa.comp:582-587), arguments in the reference's order (row y-1, then x-1 / x+1 of row y,
// then row y+1), written component by component
template <class T> inline T sobel_length(const T (&t)[8]) {
    T gx = t[0] - t[5] + T(2) * t[1] - T(2) * t[6] + t[2] - t[7];
    T gy = t[0] - t[2] + T(2) * t[3] - T(2) * t[4] + t[5] - t[7];
    return std::sqrt(gx * gx + gy * gy);  // length()
}

// calc_segmentation_value (taa.comp:589-702) at output texel iuv with the frame's historyUv;
// `total` receives the weighted indicator sum when the function gets that far (else 0). As in the
// reference, iuv is the output texel also where the input (render) size differs: the border test
// of TAA_RTFLAG_FXD is against the input size (textureSize_loRes, taa.comp:595), the neighbours
// are clamped to the input image (CLAMP_TO_TEX), the previous material is read at historyUv in
// input texels (taa.comp:622-624), the centre's material id and normal at iuv itself (zero outside
// the input image, see u32_in_or_zero), the previous mask at iuv in the output-sized mask.
template <class T>
inline uint32_t segmentation_value(const taa_uniforms_t& U, const taa_params_t& P, const Frame<T>& F, iv2 iuv,
                                   v2<T> historyUv, T& total) {
    total = T(0);
    const uint32_t fl = P.mRayTraceAugmentFlags;
    const int Wl = F.Wl, Hl = F.Hl;
    if (fl & TAA_RTFLAG_FXD) {
        const int b = 100;
        if (iuv.x < b || iuv.y < b || iuv.x >= Wl - b || iuv.y >= Hl - b) return 1u;
    }
    if (fl & TAA_RTFLAG_ALL) return 2u;
    const bool useHistoryCount = (fl & TAA_RTFLAG_CNT) != 0;
    const uint32_t newCountValue = useHistoryCount ? (uint32_t)P.mRayTraceHistoryCount << 16 : 0u;
    if (fl & TAA_RTFLAG_OUT) {
        if (historyUv.x < T(0) || historyUv.y < T(0) || historyUv.x >= T(1) || historyUv.y >= T(1)) return 1u;
    }
    const uint32_t matId = F.u32_in_or_zero(F.material, iuv.x, iuv.y);
    if (fl & TAA_RTFLAG_DIS) {
        iv2 tcPrev = uv_to_tc<T>(historyUv, iv2{Wl, Hl});
        if (tcPrev.x >= 0 && tcPrev.y >= 0 && tcPrev.x < Wl && tcPrev.y < Hl) {
            const uint32_t prevMatId = F.u32_in(F.prev_material, tcPrev.x, tcPrev.y);
            if (prevMatId != matId && (prevMatId & 0x80000000u) != 0) return 2u | newCountValue;
        }
    }
    T nrmValue = T(0), dptValue = T(0), matValue = T(0), lumValue = T(0);
    static const int ox[8] = {-1, 0, 1, -1, 1, -1, 0, 1}, oy[8] = {-1, -1, -1, 0, 0, 1, 1, 1};
    if (fl & TAA_RTFLAG_NRM) {
        v3<T> nC = seg_normal(F, iuv.x, iuv.y, false);
        auto d = [&](int dx, int dy) { v3<T> n = seg_normal(F, iuv.x + dx, iuv.y + dy); return nC.x * n.x + nC.y * n.y + nC.z * n.z; };
        T dL = d(-1, 0), dR = d(1, 0), dT = d(0, -1), dB = d(0, 1);
        T mind = gmax(T(0), gmin(gmin(gmin(dL, dR), dT), dB));
        nrmValue = T(1.0) - mind;
    }
    if (fl & TAA_RTFLAG_DPT) {
        T t[8];
        for (int k = 0; k < 8; ++k)
            t[k] = linearize_depth(F.depth_at(iuv.x + ox[k], iuv.y + oy[k]), T(U.mCamNearPlane), T(U.mCamFarPlane));
        dptValue = sobel_length(t);
    }
    if (fl & TAA_RTFLAG_MID) {
        if (matId != F.u32_in(F.material, iuv.x - 1, iuv.y) || matId != F.u32_in(F.material, iuv.x + 1, iuv.y) ||
            matId != F.u32_in(F.material, iuv.x, iuv.y - 1) || matId != F.u32_in(F.material, iuv.x, iuv.y + 1))
            matValue = T(1.0);
    }
    if (fl & TAA_RTFLAG_LUM) {
        T t[8];
        for (int k = 0; k < 8; ++k) t[k] = rgb_to_ycocg(xyz(F.color_at(iuv.x + ox[k], iuv.y + oy[k]))).x;
        lumValue = sobel_length(t);
    }
    total = nrmValue * T(P.mRayTraceAugment_WNrm) + dptValue * T(P.mRayTraceAugment_WDpt) +
            matValue * T(P.mRayTraceAugment_WMId) + lumValue * T(P.mRayTraceAugment_WLum);
    if (total >= T(P.mRayTraceAugment_Thresh)) return 2u | newCountValue;
    if (useHistoryCount) {
        const uint32_t oldCnt = (F.u32_out(F.prev_segmask, iuv.x, iuv.y) & 0xffff0000u) >> 16;
        if (oldCnt > 0) return 2u | ((oldCnt - 1u) << 16);
    }
    return 0u;
}

// taa.comp:707-960 for output texel iuv. `cr16` selects the 16-tap Catmull-Rom form (default)
// instead of the 9-tap bilinear form. With the input (render) size equal to the output size and
// mUpsampling off, every texture() of the current frame lands on a texel centre and is an exact
// fetch. Otherwise (TAAU) depth and velocity are sampled at the output texel's uv with bilinear
// weights (texture() at non-centre positions), the neighbourhood is the input's around
// iuv_lores, and mUpsampling selects the current sample by getCurrentUpsampledColor.
template <class T>
inline PixelOut<T> resolve_pixel(const taa_uniforms_t& U, const Frame<T>& F, iv2 iuv, bool cr16 = true) {
    PixelOut<T> o;
    std::memset(o.inspect, 0, sizeof(o.inspect));
    o.inspect[TAA_INSPECT_CLOSEST_TAP] = -1;
    iv2 size = {F.W, F.H};
    iv2 lsize = {F.Wl, F.Hl};
    const bool taau = F.Wl != F.W || F.Hl != F.H || U.mUpsampling;
    v2<T> uv = tc_to_uv<T>(iuv, size);
    int paramsIdx = (U.splitScreen && iuv.x > U.splitX) ? 1 : 0;
    const taa_params_t& P = U.param[paramsIdx];
    iv2 iuv_lores = uv_to_tc<T>(uv, lsize);  // == iuv without upsampling
    const v2<T> toUv = {T(1) / T(F.Wl), T(1) / T(F.Hl)};  // 1 / textureSize of depth and velocity
    auto sample_vel = [&](v2<T> p) { return bilinear<T>([&](int x, int y) { return F.vel_at(x, y); }, F.Wl, F.Hl, p); };
    auto sample_depth = [&](v2<T> p) {
        return bilinear<T>([&](int x, int y) { T d = F.depth_at(x, y); return v4<T>{d, d, d, d}; }, F.Wl, F.Hl, p).x;
    };

    if (P.mPassThrough) {  // taa.comp:726-731
        v3<T> c = xyz(F.color_at(iuv_lores.x, iuv_lores.y));
        v3<T> h = xyz(F.hist_at(iuv.x, iuv.y));
        o.screen = {c.x, c.y, c.z, T(1)};
        o.history = {h.x, h.y, h.z, T(1)};
        o.debug = {T(0), T(0), T(0), T(0)};
        return o;
    }
    if (U.mBypassHistoryUpdate) {  // taa.comp:732-737
        v3<T> h = xyz(F.hist_at(iuv.x, iuv.y));
        o.screen = {h.x, h.y, h.z, T(1)};
        o.history = {h.x, h.y, h.z, T(1)};
        o.debug = {T(0), T(0), T(0), T(0)};
        return o;
    }

    // JITTER_UV (taa.comp:47)
    v2<T> jitter_uv = {T(U.mJitterNdc[0]) * T(0.5) * T(P.mUnjitterFactor), T(U.mJitterNdc[1]) * T(0.5) * T(P.mUnjitterFactor)};

    v3<T> currentColor, colMin, colMax, colClipTowards;
    getColorAndAabb(P, F, iuv_lores, jitter_uv, currentColor, colMin, colMax, colClipTowards);
    // getCurrentColor (taa.comp:216-220)
    T beta = T(1.0);
    if (U.mUpsampling) currentColor = current_upsampled(P, F, iuv, jitter_uv, beta);
    else currentColor = maybe_rgb_to_ycocg(P, tonemap_rgb(P, current_tap(F, iuv_lores, 0, 0, P.mUnjitterCurrentSample != 0, jitter_uv)));
    T depth = F.depth_at(iuv_lores.x, iuv_lores.y);  // taa.comp:756

    // getHistoryPosition (taa.comp:391-438)
    v2<T> historyUv;
    T expectedHistoryDepth;
    {
        v4<T> vs = taau ? sample_vel(uv) : F.vel_at(iuv.x, iuv.y);
        bool canUseVelocity = true;
        if (P.mUseVelocityVectors == 0 || (P.mUseVelocityVectors == 1 && vs.w < T(0.5))) canUseVelocity = false;
        if (canUseVelocity) {
            o.inspect[TAA_INSPECT_FLAGS] |= TAA_FLAG_USED_VELOCITY;
            if (P.mVelocitySampleMode == 1) {
                // taa.comp:398-412: the loop tracks the longest vector but assigns the last tap
                v4<T> last = taau ? sample_vel(v2<T>{uv.x + toUv.x, uv.y + toUv.y}) : F.vel_at(iuv.x + 1, iuv.y + 1);
                vs.x = last.x;
                vs.y = last.y;
            } else if (P.mVelocitySampleMode == 2) {
                int k;
                if (taau) {  // taa.comp:371-389 at the output uv: depth taps interpolated
                    static const int off[9][2] = {{-1, -1}, {0, -1}, {1, -1}, {-1, 0}, {0, 0}, {1, 0}, {-1, 1}, {0, 1}, {1, 1}};
                    k = 0;
                    T dC = sample_depth(v2<T>{uv.x + toUv.x * T(-1), uv.y + toUv.y * T(-1)});
                    for (int j = 1; j < 9; ++j) {
                        T d = sample_depth(v2<T>{uv.x + toUv.x * T(off[j][0]), uv.y + toUv.y * T(off[j][1])});
                        if (d < dC) { dC = d; k = j; }
                    }
                    vs = sample_vel(v2<T>{uv.x + toUv.x * T(off[k][0]), uv.y + toUv.y * T(off[k][1])});
                } else {
                    T dC;
                    k = findClosestTap_3x3(F, iuv, dC);
                    vs = F.vel_at(iuv.x + (k % 3) - 1, iuv.y + (k / 3) - 1);
                }
                o.inspect[TAA_INSPECT_CLOSEST_TAP] = k;
            }
            historyUv = {uv.x - vs.x, uv.y - vs.y};
            expectedHistoryDepth = depth - vs.z;
        } else {
            const float* Mi = U.mInverseViewProjMatrix;
            const float* Mh = U.mHistoryViewProjMatrix;
            T c[4] = {uv.x * T(2.0) - T(1.0), uv.y * T(2.0) - T(1.0), depth, T(1)};
            T w[4], h[4];
            for (int r = 0; r < 4; ++r)
                w[r] = T(Mi[0 * 4 + r]) * c[0] + T(Mi[1 * 4 + r]) * c[1] + T(Mi[2 * 4 + r]) * c[2] + T(Mi[3 * 4 + r]) * c[3];
            for (int r = 0; r < 4; ++r)
                h[r] = T(Mh[0 * 4 + r]) * w[0] + T(Mh[1 * 4 + r]) * w[1] + T(Mh[2 * 4 + r]) * w[2] + T(Mh[3 * 4 + r]) * w[3];
            historyUv = {(h[0] / h[3]) * T(0.5) + T(0.5), (h[1] / h[3]) * T(0.5) + T(0.5)};
            expectedHistoryDepth = h[2] / h[3];
        }
    }
    v2<T> velUv = {uv.x - historyUv.x, uv.y - historyUv.y};
    T pixelSpeed = std::sqrt(velUv.x * velUv.x + velUv.y * velUv.y);
    put_f(o.inspect, TAA_INSPECT_HIST_UV_X, historyUv.x);
    put_f(o.inspect, TAA_INSPECT_HIST_UV_Y, historyUv.y);

    // sample_history_rgba (taa.comp:545-549)
    v4<T> historyRaw;
    {
        T bx = std::floor(historyUv.x * T(F.W) - T(0.5)), by = std::floor(historyUv.y * T(F.H) - T(0.5));
        o.inspect[TAA_INSPECT_HIST_TC_X] = to_int(bx);
        o.inspect[TAA_INSPECT_HIST_TC_Y] = to_int(by);
        if (P.mInterpolationMode == 0) {
            auto fetch = [&](int x, int y) { return F.hist_at(x, y); };
            historyRaw = bilinear<T>(fetch, F.W, F.H, historyUv);
        } else if (P.mInterpolationMode == 1) {
            historyRaw = bspline_4tap(F, historyUv);
        } else {
            historyRaw = cr16 ? catmullrom_16tap(F, historyUv) : catmullrom_9tap(F, historyUv);
        }
    }
    v3<T> historyColor = maybe_rgb_to_ycocg(P, xyz(historyRaw));
    T alpha = T(P.mAlpha);
    bool rejected = false;

    // ---- segmentation mask (taa.comp:773-784)
    uint32_t segMaskValue = 0;
    if (P.mRayTraceAugment && F.material) {
        T total;
        segMaskValue = segmentation_value(U, P, F, iuv, historyUv, total);
        put_f(o.inspect, TAA_INSPECT_SEG_TOTAL, total);
        if ((segMaskValue & 3u) != 0) {
            alpha = T(P.mRejectionAlpha);
            rejected = true;
        }
        o.segmask = segMaskValue;
        o.seg_written = true;
    }

    // ---- history rejection (taa.comp:787-823)
    if (P.mRejectOutside) {
        if (historyUv.x < T(0) || historyUv.y < T(0) || historyUv.x >= T(1) || historyUv.y >= T(1)) {
            alpha = T(P.mRejectionAlpha);
            rejected = true;
            o.inspect[TAA_INSPECT_FLAGS] |= TAA_FLAG_OUTSIDE;
        }
    }
    T writeDynamicMask = T(0);
    if (P.mDynamicAntiGhosting) {
        const T eps = T(1e-5);
        auto mov = [&](int dx, int dy) {
            v4<T> v = taau ? sample_vel(v2<T>{uv.x + toUv.x * T(dx), uv.y + toUv.y * T(dy)}) : F.vel_at(iuv.x + dx, iuv.y + dy);
            v = {std::fabs(v.x), std::fabs(v.y), std::fabs(v.z), std::fabs(v.w)};
            return (v.x > eps || v.y > eps) && (v.w >= T(0.5));
        };
        bool movL = mov(-1, 0), movR = mov(1, 0), movT = mov(0, -1), movB = mov(0, 1), movC = mov(0, 0);
        bool movement = movL || movR || movT || movB || movC;
        if (!movement && historyRaw.w > T(0.0)) {
            rejected = true;
            o.inspect[TAA_INSPECT_FLAGS] |= TAA_FLAG_GHOST_REJECT;
        }
        writeDynamicMask = movC ? T(1.0) : T(0.0);
    }
    if (P.mDepthCulling) {
        iv2 dtc = uv_to_tc<T>(historyUv, lsize);  // uHistoryDepth has the input size
        o.inspect[TAA_INSPECT_DEPTH_TC_X] = dtc.x;
        o.inspect[TAA_INSPECT_DEPTH_TC_Y] = dtc.y;
        T historyDepth = F.hdepth_at(dtc.x, dtc.y);
        bool rej;
        if (U.pad1 > 0.0f) {  // linear-depth extension carried in uniforms_for_taa.pad1
            T n = T(U.mCamNearPlane), f = T(U.mCamFarPlane);
            rej = std::fabs(linearize_depth(historyDepth, n, f) - linearize_depth(expectedHistoryDepth, n, f)) > T(U.pad1);
        } else {
            T depthEpsilon = T(0.1) * (T(1.0) - historyDepth);
            rej = std::fabs(historyDepth - expectedHistoryDepth) > depthEpsilon;
        }
        if (rej) {
            rejected = true;
            o.inspect[TAA_INSPECT_FLAGS] |= TAA_FLAG_DEPTH_REJECT;
        }
    }

    // ---- rectification (taa.comp:826-845)
    v3<T> origHistoryColor = historyColor;
    switch (P.mColorClampingOrClipping) {
        case 1: historyColor = vmin(vmax(historyColor, colMin), colMax); break;
        case 2: historyColor = clipAabb(colMin, colMax, historyColor); break;
        case 3: historyColor = clipAabbSlow(colMin, colMax, colClipTowards, historyColor); break;
        default: break;
    }
    v3<T> rectified_diff = historyColor - origHistoryColor;
    bool rectified = std::fabs(rectified_diff.x) > T(0.001) || std::fabs(rectified_diff.y) > T(0.001) ||
                     std::fabs(rectified_diff.z) > T(0.001);

    // ---- blending (taa.comp:850-900)
    if (rejected) {
        alpha = T(P.mRejectionAlpha);
        beta = T(1.0);
    } else {
        if (P.mVelBasedAlpha) {
            alpha = gmax(alpha, gmix(alpha, T(P.mVelBasedAlphaMax), gclamp(pixelSpeed * T(P.mVelBasedAlphaFactor), T(0), T(1))));
        }
        if (P.mLumaWeightingLottes) {
            T lumaCurrent = luminance(P, currentColor);
            T lumaHistory = luminance(P, historyColor);
            T diff = std::fabs(lumaCurrent - lumaHistory) / gmax(gmax(lumaCurrent, lumaHistory), T(0.2));
            T w = T(1.0) - diff;
            T ww = w * w;
            alpha = gmix(T(P.mMaxAlpha), T(P.mMinAlpha), ww);
        }
        if (P.mReduceBlendNearClamp) {
            T colMin_lum = luminance(P, colMin);
            T colMax_lum = luminance(P, colMax);
            T history_lum = luminance(P, origHistoryColor);
            T distToClamp = T(2.0) * std::fabs(gmin(history_lum - colMin_lum, colMax_lum - history_lum)) / (colMax_lum - colMin_lum);
            if (colMax_lum - colMin_lum < T(0.001)) distToClamp = T(1.0);
            alpha *= gclamp(T(4) * distToClamp, T(0), T(1));
        }
    }
    if (U.mResetHistory) { alpha = T(1.0); beta = T(1.0); }
    T ab = alpha * beta;
    v3<T> antiAliased = maybe_ycocg_to_rgb(P, vmix(historyColor, currentColor, ab));

    if (P.mAddNoise) {  // taa.comp:551-556, 903-906 (GLSL sin: outside the parity contract)
        T sx = uv.x + T(U.mSinTime[0]) + T(0.6959174), sy = uv.y + T(U.mSinTime[0]) + T(0.6959174);
        T s = std::sin(sx * T(12.9898) + sy * T(78.233));
        const T k[3] = {T(43758.5453), T(28001.8384), T(50849.4141)};
        T n[3];
        for (int i = 0; i < 3; ++i) {
            T v = s * k[i];
            T fr = v - std::floor(v);
            n[i] = (fr * T(2.0) - T(1.0)) * T(P.mNoiseFactor);
        }
        antiAliased = antiAliased + v3<T>{n[0], n[1], n[2]};
    }

    v3<T> scr = un_tonemap_rgb(P, antiAliased);
    o.history = {antiAliased.x, antiAliased.y, antiAliased.z, writeDynamicMask};
    o.screen = {scr.x, scr.y, scr.z, T(1.0)};

    // ---- debug image (taa.comp:912-942)
    v4<T> g = {T(0), T(0), T(0), T(0)};
    switch (P.mDebugMode) {
        case 0: { v3<T> t = colMax - colMin; g = {t.x, t.y, t.z, T(0)}; break; }
        case 1: { v3<T> t = colMax - colMin; T p = t.x * t.y * t.z; g = {p, p, p, T(0)}; break; }
        case 2: {
            T len = std::sqrt(rectified_diff.x * rectified_diff.x + rectified_diff.y * rectified_diff.y + rectified_diff.z * rectified_diff.z);
            g = {rejected ? T(1) : T(0), len, T(0), T(0)};
            break;
        }
        case 3: g = {alpha, alpha, alpha, T(0)}; break;
        case 4: g = F.vel_at(iuv.x, iuv.y); break;
        case 5: g = {pixelSpeed, T(0), T(0), T(0)}; break;
        case 6: g = o.screen; break;
        case 7: g = o.history; break;
        case 8:  // the mask category (taa.comp:932-938); category 3 does not occur
            switch (segMaskValue & 3u) {
                case 0: g = {T(0), T(0), T(1), T(0)}; break;
                case 1: g = {T(1), T(0), T(0), T(0)}; break;
                case 2: g = {T(1), T(1), T(0), T(0)}; break;
                default: break;
            }
            break;
        default: break;
    }
    T ds = T(P.mDebugScale);
    g = {g.x * ds * T(P.mDebugMask[0]), g.y * ds * T(P.mDebugMask[1]), g.z * ds * T(P.mDebugMask[2]), g.w * ds * T(P.mDebugMask[3])};
    if (P.mDebugCenter) g = {g.x * T(0.5) + T(0.5), g.y * T(0.5) + T(0.5), g.z * T(0.5) + T(0.5), g.w * T(0.5) + T(0.5)};
    o.debug = g;

    if (rejected) o.inspect[TAA_INSPECT_FLAGS] |= TAA_FLAG_REJECTED;
    if (rectified) o.inspect[TAA_INSPECT_FLAGS] |= TAA_FLAG_RECTIFIED;
    put_f(o.inspect, TAA_INSPECT_COLMIN + 0, colMin.x);
    put_f(o.inspect, TAA_INSPECT_COLMIN + 1, colMin.y);
    put_f(o.inspect, TAA_INSPECT_COLMIN + 2, colMin.z);
    put_f(o.inspect, TAA_INSPECT_COLMAX + 0, colMax.x);
    put_f(o.inspect, TAA_INSPECT_COLMAX + 1, colMax.y);
    put_f(o.inspect, TAA_INSPECT_COLMAX + 2, colMax.z);
    put_f(o.inspect, TAA_INSPECT_BLEND, ab);
    return o;
}

// ------------------------------------------------------------------------------------------
// FidelityFX CAS, sharpen-only path

inline uint32_t f2u(float f) { uint32_t u; std::memcpy(&u, &f, 4); return u; }
inline float u2f(uint32_t u) { float f; std::memcpy(&f, &u, 4); return f; }

// ffx_a.h:1455-1457. These operate on the fp32 bit pattern; the FP64 instantiation rounds its
// argument to fp32, applies the same integer step and widens the result back.
template <class T> inline T prx_lo_sqrt(T a) { return T(u2f((f2u(float(a)) >> 1u) + 0x1fbc4639u)); }
template <class T> inline T prx_lo_rcp(T a) { return T(u2f(0x7ef07ebbu - f2u(float(a)))); }
template <class T> inline T prx_med_rcp(T a) {
    T b = T(u2f(0x7ef19fffu - f2u(float(a))));
    return b * (-b * a + T(2.0));
}
template <class T> inline T sat(T a) { return gclamp(a, T(0.0), T(1.0)); }  // ASatF1 ffx_a.h:732
template <class T> inline T min3(T x, T y, T z) { return gmin(x, gmin(y, z)); }  // AMin3F1 ffx_a.h:688
template <class T> inline T max3(T x, T y, T z) { return gmax(x, gmax(y, z)); }  // AMax3F1 ffx_a.h:660

// CasFilter(noScaling = true) (ffx_cas.h:424-537, no CAS_BETTER_DIAGONALS / CAS_GO_SLOWER /
// CAS_SLOW), for the pixel at (x, y) of an RGBA image `src` (clamped loads). Only the cross
// b, d, e, f, h contributes; the corner taps are loaded by the reference but unused.
template <class T, class Load>
inline v3<T> cas_filter(Load load, int x, int y, uint32_t const1_x) {
    v3<T> b = load(x, y - 1), d = load(x - 1, y), e = load(x, y), f = load(x + 1, y), h = load(x, y + 1);
    T mnR = min3(min3(d.x, e.x, f.x), b.x, h.x);
    T mnG = min3(min3(d.y, e.y, f.y), b.y, h.y);
    T mnB = min3(min3(d.z, e.z, f.z), b.z, h.z);
    T mxR = max3(max3(d.x, e.x, f.x), b.x, h.x);
    T mxG = max3(max3(d.y, e.y, f.y), b.y, h.y);
    T mxB = max3(max3(d.z, e.z, f.z), b.z, h.z);
    T rcpMR = prx_lo_rcp(mxR), rcpMG = prx_lo_rcp(mxG), rcpMB = prx_lo_rcp(mxB);
    T ampR = sat(gmin(mnR, T(1.0) - mxR) * rcpMR);
    T ampG = sat(gmin(mnG, T(1.0) - mxG) * rcpMG);
    T ampB = sat(gmin(mnB, T(1.0) - mxB) * rcpMB);
    ampR = prx_lo_sqrt(ampR);
    ampG = prx_lo_sqrt(ampG);
    ampB = prx_lo_sqrt(ampB);
    (void)ampR; (void)ampB;  // only the green weight is used (ffx_cas.h:508-522)
    T peak = T(u2f(const1_x));
    T wG = ampG * peak;
    T rcpWeight = prx_med_rcp(T(1.0) + T(4.0) * wG);
    v3<T> o;
    o.x = sat((b.x * wG + d.x * wG + f.x * wG + h.x * wG + e.x) * rcpWeight);
    o.y = sat((b.y * wG + d.y * wG + f.y * wG + h.y * wG + e.y) * rcpWeight);
    o.z = sat((b.z * wG + d.z * wG + f.z * wG + h.z * wG + e.z) * rcpWeight);
    return o;
}

// ------------------------------------------------------------------------------------------
// host helpers

// CasSetup (ffx_cas.h:375-394) for the CPU: ALerpF1(a,b,c) = b*c + (-a*c + a) (ffx_a.h:302),
// ARcpF1 = 1/x, the packed-half constant truncates the mantissa like AU1_AH1_AF1's tables
// (ffx_a.h:475-541): overflow, Inf and NaN saturate to +-65504, denormals are kept.
inline uint32_t half_bits_trunc(float f) {
    uint32_t u = f2u(f);
    uint32_t sign = (u >> 16) & 0x8000u;
    int e = int((u >> 23) & 0xff) - 127;
    uint32_t m = u & 0x7fffffu;
    if (e > 15) return sign | 0x7bffu;
    if (e >= -14) return sign | (uint32_t(e + 15) << 10) | (m >> 13);
    if (e >= -24) return sign | ((m | 0x800000u) >> (uint32_t)(-e - 14 + 13));
    return sign;
}
inline void cas_setup(uint32_t c0[4], uint32_t c1[4], float sharpness, float inX, float inY, float outX, float outY) {
    c0[0] = f2u(inX * (1.0f / outX));
    c0[1] = f2u(inY * (1.0f / outY));
    c0[2] = f2u(0.5f * inX * (1.0f / outX) - 0.5f);
    c0[3] = f2u(0.5f * inY * (1.0f / outY) - 0.5f);
    float s = sharpness < 0.0f ? 0.0f : (sharpness > 1.0f ? 1.0f : sharpness);
    float lerp = 5.0f * s + (-8.0f * s + 8.0f);
    float sharp = -(1.0f / lerp);
    c1[0] = f2u(sharp);
    c1[1] = half_bits_trunc(sharp) + (half_bits_trunc(0.0f) << 16);
    c1[2] = f2u(8.0f * inX * (1.0f / outX));
    c1[3] = 0;
}

// helpers::halton (source/helper_functions.hpp:10-17)
inline float halton(int i, int b) {
    float f = 1.0f, r = 0.0f;
    while (i > 0) {
        f = f / float(b);
        r = r + f * float(i % b);
        i = i / b;
    }
    return r;
}

// ---- edge-AA fallback (the library's own filter for category-1 texels) ------------------------
// luma of the prepare pass (antialias_fxaa_prepare.comp:21-23)
template <class T> inline T aa_luma(v3<T> c) { return c.x * T(0.299) + c.y * T(0.587) + c.z * T(0.114); }

// One texel of taa_edge_aa_apply (include/taastar/taa_abi.h states the filter). `at(dx, dy)` gives
// the clamped neighbour's rgb; `category1` says whether the mask selects the texel.
template <class T, class At>
inline v4<T> edge_aa_texel(At at, bool category1, T subpix, T edge_threshold, T edge_threshold_min) {
    const v3<T> cC = at(0, 0);
    const T lC = aa_luma(cC);
    if (!category1) return {cC.x, cC.y, cC.z, lC};
    const T lN = aa_luma(at(0, -1)), lS = aa_luma(at(0, 1)), lW = aa_luma(at(-1, 0)), lE = aa_luma(at(1, 0));
    const T lNW = aa_luma(at(-1, -1)), lNE = aa_luma(at(1, -1)), lSW = aa_luma(at(-1, 1)), lSE = aa_luma(at(1, 1));
    const T lmin = gmin(lC, gmin(gmin(lN, lS), gmin(lW, lE)));
    const T lmax = gmax(lC, gmax(gmax(lN, lS), gmax(lW, lE)));
    const T range = lmax - lmin;
    if (range < gmax(edge_threshold_min, lmax * edge_threshold)) return {cC.x, cC.y, cC.z, lC};
    const T gx = (lNE + T(2) * lE + lSE) - (lNW + T(2) * lW + lSW);
    const T gy = (lSW + T(2) * lS + lSE) - (lNW + T(2) * lN + lNE);
    int dx = 0, dy = 0;
    if (std::fabs(gy) >= std::fabs(gx)) dy = std::fabs(lN - lC) >= std::fabs(lS - lC) ? -1 : 1;
    else dx = std::fabs(lW - lC) >= std::fabs(lE - lC) ? -1 : 1;
    const T avg = (((lN + lS) + (lW + lE)) * T(2) + ((lNW + lNE) + (lSW + lSE))) / T(12);
    const T s = gclamp(std::fabs(avg - lC) / range, T(0), T(1));
    const T blend = subpix * T(0.5) * s;
    const v3<T> cX = at(dx, dy);
    return {cC.x * (T(1) - blend) + cX.x * blend, cC.y * (T(1) - blend) + cX.y * blend,
            cC.z * (T(1) - blend) + cX.z * blend, lC};
}

// ---- velocity producer ----------------------------------------------------------------------
// The motion vector of one texel as the geometry pass writes it (fwd_geometry.frag:288-295),
// with the current clip position taken from the texel centre and its depth instead of from the
// interpolated vertex position:
//   clip      = (uv * 2 - 1, depth, 1)                                  (taa.comp:426)
//   world     = inverse_view_proj * clip           (homogeneous, not divided; taa.comp:427)
//   world     = object_prev_from_current[id - 1] * world   when 1 <= id <= n_objects
//               (prev_modelMatrix, transform_and_pass_on.vert:99 / animatedObject.vert:100)
//   prev      = prev_view_proj * world                                   (main.cpp:423)
//   ndc       = clip.xyz,  ndc_prev = prev.xyz * (1 / prev.w)            (fwd_geometry.frag:289-290)
//   ndc.xy   -= jitter.xy, ndc_prev.xy -= jitter.zw                      (:292-293)
//   v.xyz     = (ndc - ndc_prev) * (0.5, 0.5, 1),  v.w = id              (:294-295)
// Matrices are column-major; M * v is evaluated row by row as
// M[r] v0 + M[4 + r] v1 + M[8 + r] v2 + M[12 + r] v3, left to right (the resolve's matrix path
// uses the same order). GLSL's vector-by-scalar division is taken as one correctly rounded
// reciprocal and three products, the usual lowering of that division (within GLSL's 2.5 ulp).
template <class T>
inline void mat4_apply(const float* M, const T (&v)[4], T (&o)[4]) {
    for (int r = 0; r < 4; ++r) o[r] = T(M[r]) * v[0] + T(M[4 + r]) * v[1] + T(M[8 + r]) * v[2] + T(M[12 + r]) * v[3];
}

template <class T>
inline v4<T> velocity_texel(const taa_reproject_t& R, T u, T v, T depth, uint32_t id) {
    T clip[4] = {u * T(2.0) - T(1.0), v * T(2.0) - T(1.0), depth, T(1)};
    T world[4], prev[4];
    mat4_apply<T>(R.inverse_view_proj, clip, world);
    if (id >= 1u && id <= (uint32_t)R.n_objects) {
        T moved[4];
        mat4_apply<T>(R.object_prev_from_current + 16 * (id - 1u), world, moved);
        for (int r = 0; r < 4; ++r) world[r] = moved[r];
    }
    mat4_apply<T>(R.prev_view_proj, world, prev);
    const float* j = R.jitter_current_prev;
    T nx = clip[0] - T(j[0]), ny = clip[1] - T(j[1]), nz = clip[2];
    const T inv = T(1) / prev[3];
    T px = prev[0] * inv - T(j[2]), py = prev[1] * inv - T(j[3]), pz = prev[2] * inv;
    return v4<T>{(nx - px) * T(0.5), (ny - py) * T(0.5), (nz - pz) * T(1.0), T(float(id))};
}

}  // namespace oracle
