// gfx950 throughput kernel of the temporal resolve (TAA_ARITH_FP16): the taa.comp configuration
// of variance clipping + Catmull-Rom history + simple or closest-depth velocity, with optional
// clamp/clip, reject-outside, depth rejection (raw or linear), Lottes luma weighting and fused CAS.
//
// Arithmetic: texels stay fp16 (storage format) and enter the Catmull-Rom sums through mixed
// fp16 x fp32 FMAs, variance moments, rectification and the blend run in fp32 with contraction
// and the hardware reciprocal / square root. Everything that selects a texel -- the closest-depth
// tap, historyUv = uv - v, the Catmull-Rom base texel floor(uv*size - 0.5), the depth-test texel,
// the out-of-screen and depth-rejection decisions -- is computed exactly as in the FP32 path
// (IEEE division for uv and for the linear depths), so indices and rejections are bit-exact.
//
// Per workgroup (256 threads): stage the tile + apron once (YCoCg fp32, depth, velocity bits, and
// the tile's uv columns/rows), resolve the region, sharpen from LDS when CAS is fused.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#include "taa_device.hpp"
#include "taa_fast_math.hpp"
#include "taa_internal.hpp"

namespace taastar {
namespace {

// staged velocity texel: RG16F / RG32F (VEL4 false) or RGBA16F / RGBA32F (VEL4 true), with the
// colour storage's precision
template <bool S16, bool VEL4> struct VelSlot;
template <> struct VelSlot<true, false> { typedef uint32_t T; };  // RG16F bits
template <> struct VelSlot<false, false> { typedef float2 T; };   // RG32F
template <> struct VelSlot<true, true> { typedef uint2 T; };      // RGBA16F bits
template <> struct VelSlot<false, true> { typedef float4 T; };    // RGBA32F

template <int TW_, int TH_, bool CAS_, bool S16_, bool VEL4_ = false>
struct FGeo {
    static constexpr int TW = TW_, TH = TH_;
    static constexpr bool CAS = CAS_, S16 = S16_, VEL4 = VEL4_;
    static constexpr int NT = 256;
    static constexpr int A = CAS ? 2 : 1;
    static constexpr int SW = TW + 2 * A, SH = TH + 2 * A, SP = SW + 1;
    static constexpr int RA = CAS ? 1 : 0;
    static constexpr int RW = TW + 2 * RA, RH = TH + 2 * RA, RP = RW + 1;
    typedef typename VelSlot<S16, VEL4>::T VelT;
    // resolved screen texel kept for the fused CAS: the value the stored image would hold
    typedef typename std::conditional<S16, uint2, float4>::type ResT;
    static constexpr int NIT = (RW * RH + NT - 1) / NT;  // resolve iterations per thread
    // LDS: the staged region (Y, Co, Cg and depth in one 16-byte entry per texel, velocity, the uv
    // of the staged columns and rows) during the resolve; afterwards the same bytes hold the
    // resolved region that the fused CAS reads.
    struct Smem {
        static constexpr int kCd = SH * SP * 16, kVel = SH * SP * (int)sizeof(VelT), kUv = (SW + SH) * 4;
        static constexpr int kStage = kCd + kVel + kUv, kRes = CAS ? RH * RP * (int)sizeof(ResT) : 0;
        alignas(16) unsigned char raw[kStage > kRes ? kStage : kRes];
        __device__ float4* cd() { return reinterpret_cast<float4*>(raw); }
        __device__ VelT* v() { return reinterpret_cast<VelT*>(raw + kCd); }
        __device__ float* ux() { return reinterpret_cast<float*>(raw + kCd + kVel); }
        __device__ float* uy() { return ux() + SW; }
        __device__ ResT* r() { return reinterpret_cast<ResT*>(raw); }
    };
};

template <bool S16, bool VEL4>
__device__ __forceinline__ float2 vel_xy(typename VelSlot<S16, VEL4>::T v) {
    if constexpr (VEL4) {
        if constexpr (S16) return make_float2(h2f(v.x & 0xffffu), h2f(v.x >> 16));
        else return make_float2(v.x, v.y);
    } else {
        if constexpr (S16) return make_float2(h2f(v & 0xffffu), h2f(v >> 16));
        else return v;
    }
}
// the whole texel as the resolve reads it: RG formats as (r, g, 0, 1)
template <bool S16, bool VEL4>
__device__ __forceinline__ float4 vel_xyzw(typename VelSlot<S16, VEL4>::T v) {
    if constexpr (VEL4) {
        if constexpr (S16) return make_float4(h2f(v.x & 0xffffu), h2f(v.x >> 16), h2f(v.y & 0xffffu), h2f(v.y >> 16));
        else return v;
    } else {
        const float2 q = vel_xy<S16, false>(v);
        return make_float4(q.x, q.y, 0.0f, 1.0f);
    }
}

// GEN: the general configuration (variance or plain min/max box, Catmull-Rom or bilinear history,
// chosen per launch from the parameter block); without it the kernel is the variance box with
// Catmull-Rom history and no branch on either (the production configuration keeps the schedule
// measured for it: the general form runs it 15% slower)
// EXT (implies GEN): the reference's other per-texel options -- RGBA velocity (.z into the
// expected history depth, .w for movers only with matrix reprojection for the rest), the 3x3-longest
// velocity quirk, dynamic anti-ghosting, velocity-based alpha, the near-clamp blend reduction and
// the debug image -- in a form of their own, so that blocks without them keep the general form's
// register budget
template <class G, bool INSP, bool GEN, bool EXT>
__device__ void fast_tile(const FrameDesc& fd, const taa_uniforms_t& U, int tile, typename G::Smem& sm) {
    constexpr bool S16 = G::S16;
    const taa_params_t& P = U.param[fd.param_idx];
    const int W = fd.W, H = fd.H;
    const uint32_t pitch = (uint32_t)fd.pitch;
    const int ty = tile / fd.tiles_x, tx = tile - ty * fd.tiles_x;
    const int x0 = fd.x_begin + tx * G::TW, y0 = ty * G::TH;
    const int gx0 = x0 - G::A, gy0 = y0 - G::A;
    const int tid = threadIdx.x;

    const bool ycocg = P.mUseYCoCg != 0;
    const bool closest_mode = P.mVelocitySampleMode == 2;
    const int rect = P.mColorClampingOrClipping;
    const bool lottes = P.mLumaWeightingLottes != 0;
    const bool reject_out = P.mRejectOutside != 0;
    const bool depth_cull = P.mDepthCulling != 0;
    const bool lin_depth = U.pad1 > 0.0f;
    const bool reset = U.mResetHistory != 0;
    const float gamma = P.mVarClipGamma;
    const bool var_clip = !GEN || P.mVarianceClipping != 0;   // else the plain min/max box of the nine taps
    const bool cr_hist = !GEN || P.mInterpolationMode == 2;  // else bilinear history
    const float fW = (float)W, fH = (float)H;
    typedef typename G::VelT VelT;

    float4* const scd = sm.cd();
    VelT* const sv = sm.v();
    float* const sux = sm.ux();
    float* const suy = sm.uy();

    // 1. stage the region: YCoCg (fp32), depth, velocity; uv of the staged columns and rows. The
    // staging loads issue at the highest wave priority, ahead of other workgroups' arithmetic
    // (4K: +0.5% in three same-box triples; priority 2 here: +0.1%)
    __builtin_amdgcn_s_setprio(3);
    {
        constexpr int DY = G::NT / G::SW, DX = G::NT % G::SW;
        int sy = tid / G::SW, sx = tid - sy * G::SW;
        for (int i = tid; i < G::SW * G::SH; i += G::NT) {
            const int gx = clampi(gx0 + sx, 0, W - 1), gy = clampi(gy0 + sy, 0, H - 1);
            const uint32_t gi = (uint32_t)gy * pitch + (uint32_t)gx;
            const float4 c = ld_rgba<S16>(fd.color, gi);
            float Y = c.x, Co = c.y, Cg = c.z;
            if (ycocg) {
                const float t = 0.25f * (c.x + c.z);
                Y = __builtin_fmaf(0.5f, c.y, t);
                Cg = __builtin_fmaf(0.5f, c.y, -t);
                Co = 0.5f * (c.x - c.z);
            }
            const int li = sy * G::SP + sx;
            scd[li] = make_float4(Y, Co, Cg, ld_f32(fd.depth, gi));
            sv[li] = gload<VelT>(fd.velocity, gi * (uint32_t)sizeof(VelT));
            sx += DX;
            sy += DY;
            if (sx >= G::SW) { sx -= G::SW; sy += 1; }
        }
    }
    __builtin_amdgcn_s_setprio(0);
    if (tid < G::SW) sux[tid] = ((float)clampi(gx0 + tid, 0, W - 1) + 0.5f) / fW;              // taa.comp:131
    else if (tid < G::SW + G::SH) suy[tid - G::SW] = ((float)clampi(gy0 + tid - G::SW, 0, H - 1) + 0.5f) / fH;
    __syncthreads();

    constexpr float inv9 = 1.0f / 9.0f;
    constexpr int RDY = G::NT / G::RW, RDX = G::NT % G::RW;
    int ry = tid / G::RW, rx = tid - ry * G::RW;
    typename G::ResT keep[G::CAS ? G::NIT : 1];  // resolved texels for the fused CAS, written to LDS afterwards
#pragma unroll
    for (int it = 0; it < G::NIT; ++it) {
        if (it > 0) {
            rx += RDX;
            ry += RDY;
            if (rx >= G::RW) { rx -= G::RW; ry += 1; }
        }
        if (tid + it * G::NT >= G::RW * G::RH) break;
        // raised wave priority from the start of a texel until its history loads are issued: the
        // chain LDS -> closest depth -> velocity -> history address gets issue slots ahead of
        // other waves' arithmetic (measured +1.5% at 4K)
        __builtin_amdgcn_s_setprio(2);
        const int px = x0 - G::RA + rx, py = y0 - G::RA + ry;
        const int cx = clampi(px, 0, W - 1), cy = clampi(py, 0, H - 1);
        const bool owned = rx >= G::RA && rx < G::RA + G::TW && ry >= G::RA && ry < G::RA + G::TH && px < fd.x_end && py < H;
        const int lx = cx - gx0, ly = cy - gy0, l = ly * G::SP + lx;

        // ---- 3x3 neighbourhood: variance box (fp32 moments) and closest-depth tap
        // (first minimum in scan order, strict <), from the same nine LDS reads
        f3 m1, m2;
        float dC;
        int vo = -G::SP - 1, closest = 0;
        {
            const float4 t = scd[l - G::SP - 1];
            m1 = {t.x, t.y, t.z};
            m2 = {t.x * t.x, t.y * t.y, t.z * t.z};
            dC = t.w;
        }
#pragma unroll
        for (int k = 1; k < 9; ++k) {
            const int o = (k / 3 - 1) * G::SP + (k % 3 - 1);
            const float4 t = scd[l + o];
            m1.x += t.x; m1.y += t.y; m1.z += t.z;
            m2.x = __builtin_fmaf(t.x, t.x, m2.x);
            m2.y = __builtin_fmaf(t.y, t.y, m2.y);
            m2.z = __builtin_fmaf(t.z, t.z, m2.z);
            if (t.w < dC) {
                dC = t.w;
                vo = o;
                if constexpr (INSP) closest = k;
            }
        }
        const f3 mean = {m1.x * inv9, m1.y * inv9, m1.z * inv9};
        const f3 sig = {sqrt_fast(fmaxf(__builtin_fmaf(-mean.x, mean.x, m2.x * inv9), 0.f)),
                        sqrt_fast(fmaxf(__builtin_fmaf(-mean.y, mean.y, m2.y * inv9), 0.f)),
                        sqrt_fast(fmaxf(__builtin_fmaf(-mean.z, mean.z, m2.z * inv9), 0.f))};
        const float4 ctr = scd[l];
        const f3 cur = {ctr.x, ctr.y, ctr.z};
        const float depth = ctr.w;
        const float uvx = sux[lx], uvy = suy[ly];
        float hu, hv, expected = depth;  // history position and expected history depth (taa.comp:391-438)
        bool used_velocity = true;
        if (!EXT) {
            if (!closest_mode) {  // velocity of the centre tap
                vo = 0;
                closest = -1;
            }
            const float2 v = vel_xy<S16, G::VEL4>(sv[l + vo]);
            hu = uvx - v.x;  // taa.comp:421
            hv = uvy - v.y;
        } else {
            // movers only (mode 1) reads .w of the centre tap; RG formats read w = 1
            const float4 vc = vel_xyzw<S16, G::VEL4>(sv[l]);
            used_velocity = !(P.mUseVelocityVectors == 0 || (P.mUseVelocityVectors == 1 && vc.w < 0.5f));
            if (used_velocity) {
                float4 vs = vc;
                if (closest_mode) {
                    vs = vel_xyzw<S16, G::VEL4>(sv[l + vo]);
                } else {
                    closest = -1;
                    if (P.mVelocitySampleMode == 1) {  // taa.comp:398-412: the last tap's xy, not the longest
                        const float2 q = vel_xy<S16, G::VEL4>(sv[l + G::SP + 1]);
                        vs.x = q.x;
                        vs.y = q.y;
                    }
                }
                hu = uvx - vs.x;
                hv = uvy - vs.y;
                expected = depth - vs.z;
            } else {  // reprojection through the matrices (taa.comp:424-431), IEEE divisions as the FP32 path
                closest = -1;
                const float* Mi = U.mInverseViewProjMatrix;
                const float* Mh = U.mHistoryViewProjMatrix;
                const float cs[4] = {uvx * 2.0f - 1.0f, uvy * 2.0f - 1.0f, depth, 1.0f};
                float wv[4], hc[4];
#pragma unroll
                for (int r = 0; r < 4; ++r) wv[r] = Mi[r] * cs[0] + Mi[4 + r] * cs[1] + Mi[8 + r] * cs[2] + Mi[12 + r] * cs[3];
#pragma unroll
                for (int r = 0; r < 4; ++r) hc[r] = Mh[r] * wv[0] + Mh[4 + r] * wv[1] + Mh[8 + r] * wv[2] + Mh[12 + r] * wv[3];
                hu = (hc[0] / hc[3]) * 0.5f + 0.5f;
                hv = (hc[1] / hc[3]) * 0.5f + 0.5f;
                expected = hc[2] / hc[3];
            }
        }

        // previous depth under the history position (taa.comp:818), fetched ahead of the history
        // gather so that the two latencies overlap
        int dtx = 0, dty = 0;
        float hd = 0.0f;
        if (depth_cull) {
            dtx = to_int(hu * fW);
            dty = to_int(hv * fH);
            hd = ld_f32(fd.hdepth, (uint32_t)clampi(dty, 0, H - 1) * pitch + (uint32_t)clampi(dtx, 0, W - 1));
        }

        // ---- Catmull-Rom history (taa.comp:441-514)
        const float tX = hu * fW - 0.5f, tY = hv * fH - 0.5f;
        const float fbx = floorf(tX), fby = floorf(tY);
        const int kx = to_int(fbx), ky = to_int(fby);
        const bool ghost = EXT && P.mDynamicAntiGhosting != 0;
        const Hist hs = !cr_hist ? history_bilinear_fast<S16>(fd.history, W, H, (int)pitch, kx, ky, tX - fbx, tY - fby)
                        : ghost  ? history_cr_fast<S16, true>(fd.history, W, H, (int)pitch, kx, ky, tX - fbx, tY - fby)
                                 : history_cr_fast<S16, false>(fd.history, W, H, (int)pitch, kx, ky, tX - fbx, tY - fby);
        __builtin_amdgcn_s_setprio(0);
        const f3 hrgb = hs.c;
        f3 h = hrgb;
        if (ycocg) {
            const float t = 0.25f * (hrgb.x + hrgb.z);
            h = {__builtin_fmaf(0.5f, hrgb.y, t), 0.5f * (hrgb.x - hrgb.z), __builtin_fmaf(0.5f, hrgb.y, -t)};
        }

        // ---- rejection (taa.comp:787-823), decisions as in the FP32 path
        bool rejected = false;
        int flags = used_velocity ? TAA_FLAG_USED_VELOCITY : 0;
        if (reject_out && (hu < 0.0f || hv < 0.0f || hu >= 1.0f || hv >= 1.0f)) {
            rejected = true;
            flags |= TAA_FLAG_OUTSIDE;
        }
        float write_mask = 0.0f;  // history alpha: the dynamic mask of the anti-ghosting (taa.comp:796-811)
        if (ghost) {
            constexpr float eps = 1e-5f;
            auto mov = [&](int o) {
                const float4 q = vel_xyzw<S16, G::VEL4>(sv[l + o]);
                return (fabsf(q.x) > eps || fabsf(q.y) > eps) && (fabsf(q.w) >= 0.5f);
            };
            const bool movC = mov(0);
            if (!(mov(-1) || mov(1) || mov(-G::SP) || mov(G::SP) || movC) && hs.a > 0.0f) {
                rejected = true;
                flags |= TAA_FLAG_GHOST_REJECT;
            }
            write_mask = movC ? 1.0f : 0.0f;
        }
        if (depth_cull) {
            bool rej;
            if (lin_depth) {
                const float n = U.mCamNearPlane, f = U.mCamFarPlane;
                rej = linear_depth_reject(hd, expected, n, f, U.pad1);
            } else {
                rej = fabsf(hd - expected) > 0.1f * (1.0f - hd);
            }
            if (rej) {
                rejected = true;
                flags |= TAA_FLAG_DEPTH_REJECT;
            }
        }

        // ---- rectification: clamp to / clip towards the variance box (centre = mean) or the plain
        // min/max box of the nine taps (taa.comp:288-292, centre = its midpoint)
        f3 mn = {__builtin_fmaf(-gamma, sig.x, mean.x), __builtin_fmaf(-gamma, sig.y, mean.y), __builtin_fmaf(-gamma, sig.z, mean.z)};
        f3 mx = {__builtin_fmaf(gamma, sig.x, mean.x), __builtin_fmaf(gamma, sig.y, mean.y), __builtin_fmaf(gamma, sig.z, mean.z)};
        if (!var_clip) {
            mn = cur;
            mx = cur;
#pragma unroll
            for (int k = 0; k < 9; ++k) {
                const float4 t = scd[l + (k / 3 - 1) * G::SP + (k % 3 - 1)];
                mn = vmin(mn, f3{t.x, t.y, t.z});
                mx = vmax(mx, f3{t.x, t.y, t.z});
            }
        }
        const f3 h_orig = h;
        if (rect == 2) {
            // clipAabb (taa.comp:323-345): ma = max_i |v_i| / e_i, clip when ma > 1, with one
            // reciprocal: ma = M / D for M = max_i |v_i| * prod_{j != i} e_j and D = e_x e_y e_z
            const f3 bc = var_clip ? mean : f3{0.5f * (mx.x + mn.x), 0.5f * (mx.y + mn.y), 0.5f * (mx.z + mn.z)};
            const f3 vc = h - bc;
            const float ex = var_clip ? __builtin_fmaf(gamma, sig.x, 1e-7f) : __builtin_fmaf(0.5f, mx.x - mn.x, 1e-7f);
            const float ey = var_clip ? __builtin_fmaf(gamma, sig.y, 1e-7f) : __builtin_fmaf(0.5f, mx.y - mn.y, 1e-7f);
            const float ez = var_clip ? __builtin_fmaf(gamma, sig.z, 1e-7f) : __builtin_fmaf(0.5f, mx.z - mn.z, 1e-7f);
            const float exy = ex * ey, D = exy * ez;
            const float M = fmaxf(fabsf(vc.x) * (ey * ez), fmaxf(fabsf(vc.y) * (ex * ez), fabsf(vc.z) * exy));
            if (M > D) {
                const float s = D * rcp_fast(M);
                h = {__builtin_fmaf(vc.x, s, bc.x), __builtin_fmaf(vc.y, s, bc.y), __builtin_fmaf(vc.z, s, bc.z)};
            }
        } else if (rect == 1) {
            h = vmin(vmax(h, mn), mx);
        }

        // ---- blend (taa.comp:850-900)
        auto luma = [&](f3 c) { return ycocg ? c.x : 0.25f * (c.x + c.z) + 0.5f * c.y; };
        float pixel_speed = 0.0f;
        if (EXT) {
            const float vux = uvx - hu, vuy = uvy - hv;
            pixel_speed = sqrt_fast(vux * vux + vuy * vuy);  // taa.comp:434-435
        }
        float alpha = P.mAlpha;
        if (rejected) {
            alpha = P.mRejectionAlpha;
        } else {
            if (EXT && P.mVelBasedAlpha) {  // taa.comp:855-859
                const float t = fminf(fmaxf(pixel_speed * P.mVelBasedAlphaFactor, 0.0f), 1.0f);
                alpha = fmaxf(alpha, __builtin_fmaf(t, P.mVelBasedAlphaMax - alpha, alpha));
            }
            if (lottes) {
                const float lc = luma(cur), lh = luma(h);
                const float diff = fabsf(lc - lh) * rcp_fast(fmaxf(fmaxf(lc, lh), 0.2f));
                const float w = 1.0f - diff;
                alpha = __builtin_fmaf(w * w, P.mMinAlpha - P.mMaxAlpha, P.mMaxAlpha);
            }
            if (EXT && P.mReduceBlendNearClamp) {  // taa.comp:873-894 (the Unreal-based form)
                const float mnl = luma(mn), mxl = luma(mx), hl = luma(h_orig);
                float dist = 2.0f * fabsf(fminf(hl - mnl, mxl - hl)) * rcp_fast(mxl - mnl);
                if (mxl - mnl < 0.001f) dist = 1.0f;
                alpha *= fminf(fmaxf(4.0f * dist, 0.0f), 1.0f);
            }
        }
        if (reset) alpha = 1.0f;
        f3 aa = {__builtin_fmaf(alpha, cur.x - h.x, h.x), __builtin_fmaf(alpha, cur.y - h.y, h.y), __builtin_fmaf(alpha, cur.z - h.z, h.z)};
        if (reset) aa = cur;  // exactly the current colour, whatever the history holds
        if (ycocg) {
            const float t = aa.x - aa.z;
            aa = {t + aa.y, aa.x + aa.z, t - aa.y};
        }

        const uint32_t gi = (uint32_t)cy * pitch + (uint32_t)cx;
        // fp16 bits of the result, converted once for the history, the screen and the CAS stage
        uint2 hb = make_uint2(0u, 0u);
        if constexpr (S16) hb = make_uint2(pack_f2h(aa.x, aa.y), f2h(aa.z));
        if (owned) {
            if constexpr (S16) {
                // history alpha = the dynamic mask (0 or 1: 0x3c00 in fp16); the screen's alpha is 1
                gstore<uint2>(fd.history_out, gi * 8u, make_uint2(hb.x, hb.y | (write_mask != 0.0f ? 0x3c000000u : 0u)));
                if constexpr (!G::CAS) gstore<uint2>(fd.screen, gi * 8u, make_uint2(hb.x, hb.y | 0x3c000000u));
            } else {
                st_rgba<S16>(fd.history_out, gi, make_float4(aa.x, aa.y, aa.z, write_mask));
                if constexpr (!G::CAS) st_rgba<S16>(fd.screen, gi, make_float4(aa.x, aa.y, aa.z, 1.0f));
            }
            if (EXT && fd.debug) {  // taa.comp:912-942 (mode 8: no mask here, category 0)
                float4 g = make_float4(0.f, 0.f, 0.f, 0.f);
                switch (P.mDebugMode) {
                    case 0: g = make_float4(mx.x - mn.x, mx.y - mn.y, mx.z - mn.z, 0.f); break;
                    case 1: { const float q = (mx.x - mn.x) * (mx.y - mn.y) * (mx.z - mn.z); g = make_float4(q, q, q, 0.f); break; }
                    case 2: {
                        const f3 d = h - h_orig;
                        g = make_float4(rejected ? 1.f : 0.f, sqrt_fast(d.x * d.x + d.y * d.y + d.z * d.z), 0.f, 0.f);
                        break;
                    }
                    case 3: g = make_float4(alpha, alpha, alpha, 0.f); break;
                    case 4: g = vel_xyzw<S16, G::VEL4>(sv[l]); break;
                    case 5: g = make_float4(pixel_speed, 0.f, 0.f, 0.f); break;
                    case 6: g = make_float4(aa.x, aa.y, aa.z, 1.0f); break;
                    case 7: g = make_float4(aa.x, aa.y, aa.z, write_mask); break;
                    case 8: g = make_float4(0.f, 0.f, 1.f, 0.f); break;
                    default: break;
                }
                const float ds = P.mDebugScale;
                g = make_float4(g.x * ds * P.mDebugMask[0], g.y * ds * P.mDebugMask[1], g.z * ds * P.mDebugMask[2],
                                g.w * ds * P.mDebugMask[3]);
                if (P.mDebugCenter) g = make_float4(g.x * 0.5f + 0.5f, g.y * 0.5f + 0.5f, g.z * 0.5f + 0.5f, g.w * 0.5f + 0.5f);
                st_rgba16(fd.debug, gi, g);
            }
            if constexpr (INSP) {
                if (rejected) flags |= TAA_FLAG_REJECTED;
                int32_t w[TAA_INSPECT_WORDS];
                w[TAA_INSPECT_HIST_TC_X] = kx;
                w[TAA_INSPECT_HIST_TC_Y] = ky;
                w[TAA_INSPECT_CLOSEST_TAP] = closest;
                w[TAA_INSPECT_FLAGS] = flags;
                w[TAA_INSPECT_COLMIN + 0] = __float_as_int(mn.x);
                w[TAA_INSPECT_COLMIN + 1] = __float_as_int(mn.y);
                w[TAA_INSPECT_COLMIN + 2] = __float_as_int(mn.z);
                w[TAA_INSPECT_COLMAX + 0] = __float_as_int(mx.x);
                w[TAA_INSPECT_COLMAX + 1] = __float_as_int(mx.y);
                w[TAA_INSPECT_COLMAX + 2] = __float_as_int(mx.z);
                w[TAA_INSPECT_BLEND] = __float_as_int(alpha);
                w[TAA_INSPECT_DEPTH_TC_X] = dtx;
                w[TAA_INSPECT_DEPTH_TC_Y] = dty;
                w[TAA_INSPECT_HIST_UV_X] = __float_as_int(hu);
                w[TAA_INSPECT_HIST_UV_Y] = __float_as_int(hv);
                w[15] = 0;
                st_inspect(fd.inspect, gi, w);
            }
        }
        if constexpr (G::CAS) {
            if constexpr (S16) keep[it] = hb;
            else keep[it] = make_float4(aa.x, aa.y, aa.z, 0.0f);
        }
    }

    if constexpr (G::CAS) {
        __syncthreads();  // every thread is done with the staged region
        typename G::ResT* const sr = sm.r();
#pragma unroll
        for (int it = 0; it < G::NIT; ++it) {
            const int i = tid + it * G::NT;
            if (i >= G::RW * G::RH) break;
            const int qy = i / G::RW, qx = i - qy * G::RW;
            sr[qy * G::RP + qx] = keep[it];
        }
        __syncthreads();
        const float peak = __uint_as_float(fd.cas_peak_bits);
        constexpr int CDY = G::NT / G::TW, CDX = G::NT % G::TW;
        int oy = tid / G::TW, ox = tid - oy * G::TW;
        for (int i = tid; i < G::TW * G::TH; i += G::NT, ox += CDX, oy += CDY) {
            if (ox >= G::TW) { ox -= G::TW; oy += 1; }
            const int px = x0 + ox, py = y0 + oy;
            if (px >= fd.x_end || py >= H) continue;
            const int li = (oy + 1) * G::RP + (ox + 1);
            const uint32_t go = (uint32_t)py * pitch + (uint32_t)px;
            if constexpr (S16) {
                gstore<uint2>(fd.screen, go * 8u,
                    cas_filter_h2(sr[li - G::RP], sr[li - 1], sr[li], sr[li + 1], sr[li + G::RP], peak));
            } else {
                auto at = [&](int k) {
                    const float4 q = sr[k];
                    return f3{q.x, q.y, q.z};
                };
                const f3 o = cas_filter_fma(at(li - G::RP), at(li - 1), at(li), at(li + 1), at(li + G::RP), peak);
                st_rgba<S16>(fd.screen, go, make_float4(o.x, o.y, o.z, 1.0f));
            }
        }
    }
}

struct FastArgs {
    FrameDesc fd;
    taa_uniforms_t U;
};

// Occupancy hint: five waves per SIMD for the extended forms (the RGBA-velocity form on FP16
// storage then holds 96 VGPRs and 24 bytes of spills instead of 101 VGPRs at four waves: 4K with
// the reference's defaults 56.7 -> 57.7 Gpix/s); the production form reaches 92 VGPRs unhinted.
template <int TW, int TH, bool CAS, bool S16, bool INSP, int FORM>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(FORM == 2 || (FORM == 3 && S16) ? 5 : 1))) void k_resolve_fast(const FastArgs a) {
    using G = FGeo<TW, TH, CAS, S16, FORM == 3>;
    __shared__ typename G::Smem sm;
    const int n = a.fd.tiles_x * a.fd.tiles_y;
    fast_tile<G, INSP, FORM != 0, FORM >= 2>(a.fd, a.U, xcd_tile(blockIdx.x, n), sm);
}

template <int TW, int TH, bool CAS, bool S16, int FORM>
__global__ __launch_bounds__(256) void k_resolve_fast_batch(const BatchEntry* __restrict__ entries, int n_entries, int total_tiles) {
    using G = FGeo<TW, TH, CAS, S16, FORM == 3>;
    __shared__ typename G::Smem sm;
    const int t = batch_tile(total_tiles);
    if (t < 0) return;
    const BatchEntry& e = batch_entry(entries, n_entries, t);
    fast_tile<G, false, FORM != 0, FORM >= 2>(e.fd, e.U, t - e.tile_begin, sm);
}

// FORM: 0 the production form (variance box, Catmull-Rom, RG velocity, no branch on either),
// 1 the general form (min/max box or bilinear history), 2 the extended form with RG velocity,
// 3 the extended form with RGBA velocity
template <bool CAS, bool S16, bool INSP, int FORM>
hipError_t launch_fast_t(const FrameDesc& fd, const taa_uniforms_t& U, int shape, hipStream_t st) {
    FastArgs a;
    a.fd = fd;
    a.U = U;
    const int n = fd.tiles_x * fd.tiles_y;
    constexpr const TileShapeDef* S = CAS ? kShapesCas : kShapesPlain;
    switch (shape) {
        case 0: hipLaunchKernelGGL((k_resolve_fast<S[0].tw, S[0].th, CAS, S16, INSP, FORM>), dim3(n), dim3(256), 0, st, a); break;
        case 1: hipLaunchKernelGGL((k_resolve_fast<S[1].tw, S[1].th, CAS, S16, INSP, FORM>), dim3(n), dim3(256), 0, st, a); break;
        case 2: hipLaunchKernelGGL((k_resolve_fast<S[2].tw, S[2].th, CAS, S16, INSP, FORM>), dim3(n), dim3(256), 0, st, a); break;
        default: hipLaunchKernelGGL((k_resolve_fast<S[3].tw, S[3].th, CAS, S16, INSP, FORM>), dim3(n), dim3(256), 0, st, a); break;
    }
    return hipGetLastError();
}

template <bool CAS, bool S16, int FORM>
hipError_t launch_fast_batch_t(const BatchEntry* e, int n, int total, int shape, hipStream_t st) {
    const int grid = batch_grid(total);
    constexpr const TileShapeDef* S = CAS ? kShapesCas : kShapesPlain;
    switch (shape) {
        case 0: hipLaunchKernelGGL((k_resolve_fast_batch<S[0].tw, S[0].th, CAS, S16, FORM>), dim3(grid), dim3(256), 0, st, e, n, total); break;
        case 1: hipLaunchKernelGGL((k_resolve_fast_batch<S[1].tw, S[1].th, CAS, S16, FORM>), dim3(grid), dim3(256), 0, st, e, n, total); break;
        case 2: hipLaunchKernelGGL((k_resolve_fast_batch<S[2].tw, S[2].th, CAS, S16, FORM>), dim3(grid), dim3(256), 0, st, e, n, total); break;
        default: hipLaunchKernelGGL((k_resolve_fast_batch<S[3].tw, S[3].th, CAS, S16, FORM>), dim3(grid), dim3(256), 0, st, e, n, total); break;
    }
    return hipGetLastError();
}

bool rgba_velocity(const FrameDesc& fd) { return fd.vel_fmt == TAA_VEL_RGBA16F || fd.vel_fmt == TAA_VEL_RGBA32F; }

}  // namespace

bool fast_path_supported(const FrameDesc& fd, const taa_uniforms_t& U, int param_idx, const char** why) {
    const taa_params_t& P = U.param[param_idx];
    const bool s16 = fd.color_fmt == TAA_FMT_RGBA16F;
    const char* r = nullptr;
    if (s16 ? (fd.vel_fmt != TAA_VEL_RG16F && fd.vel_fmt != TAA_VEL_RGBA16F) : (fd.vel_fmt != TAA_VEL_RG32F && fd.vel_fmt != TAA_VEL_RGBA32F))
        r = "velocity format of another precision than the colour storage (RG16F / RGBA16F with RGBA16F, RG32F / RGBA32F with "
            "RGBA32F)";
    else if (fd.up || U.mUpsampling) r = "upsampling (input size != output size, or mUpsampling)";
    else if (U.mBypassHistoryUpdate) r = "mBypassHistoryUpdate";
    else if (P.mPassThrough) r = "mPassThrough";
    else if (P.mToneMapLumaKaris) r = "mToneMapLumaKaris";
    else if (P.mUnjitterNeighbourhood || P.mUnjitterCurrentSample) r = "unjitter";
    else if (!P.mVarianceClipping && P.mShapedNeighbourhood) r = "mShapedNeighbourhood";
    else if (P.mShrinkChromaAxis) r = "mShrinkChromaAxis";
    else if (P.mColorClampingOrClipping < 0 || P.mColorClampingOrClipping > 2) r = "mColorClampingOrClipping 3 (slow clip)";
    else if (P.mInterpolationMode != 2 && P.mInterpolationMode != 0) r = "B-spline history (mInterpolationMode 1)";
    else if (P.mVelocitySampleMode < 0 || P.mVelocitySampleMode > 2) r = "unknown mVelocitySampleMode";
    else if (P.mAddNoise) r = "mAddNoise";
    if (why) *why = r;
    return r == nullptr;
}

int fast_path_form(const FrameDesc& fd, const taa_uniforms_t& U, int param_idx) {
    const taa_params_t& P = U.param[param_idx];
    if (rgba_velocity(fd)) return 3;
    // with RG velocity .w reads 1, so "movers only" (mode 1) uses every velocity, as mode 2 does
    if (P.mUseVelocityVectors == 0 || P.mVelocitySampleMode == 1 || P.mDynamicAntiGhosting || P.mVelBasedAlpha ||
        P.mReduceBlendNearClamp || fd.debug != nullptr)
        return 2;
    return (!P.mVarianceClipping || P.mInterpolationMode != 2) ? 1 : 0;
}

template <bool CAS, bool S16, bool INSP>
hipError_t launch_fast_g(const FrameDesc& fd, const taa_uniforms_t& U, int shape, hipStream_t st) {
    switch (fast_path_form(fd, U, fd.param_idx)) {
        case 0: return launch_fast_t<CAS, S16, INSP, 0>(fd, U, shape, st);
        case 1: return launch_fast_t<CAS, S16, INSP, 1>(fd, U, shape, st);
        case 2: return launch_fast_t<CAS, S16, INSP, 2>(fd, U, shape, st);
        default: return launch_fast_t<CAS, S16, INSP, 3>(fd, U, shape, st);
    }
}

int launch_resolve_fast(const FrameDesc& fd, const taa_uniforms_t& U, int shape, void* stream) {
    hipStream_t st = static_cast<hipStream_t>(stream);
    const bool s16 = fd.color_fmt == TAA_FMT_RGBA16F, cas = fd.cas_on != 0, insp = fd.inspect != nullptr;
    hipError_t e;
    if (insp) {
        if (cas) e = s16 ? launch_fast_g<true, true, true>(fd, U, shape, st) : launch_fast_g<true, false, true>(fd, U, shape, st);
        else e = s16 ? launch_fast_g<false, true, true>(fd, U, shape, st) : launch_fast_g<false, false, true>(fd, U, shape, st);
    } else {
        if (cas) e = s16 ? launch_fast_g<true, true, false>(fd, U, shape, st) : launch_fast_g<true, false, false>(fd, U, shape, st);
        else e = s16 ? launch_fast_g<false, true, false>(fd, U, shape, st) : launch_fast_g<false, false, false>(fd, U, shape, st);
    }
    return (int)e;
}

template <bool CAS, bool S16>
hipError_t launch_fast_batch_f(const BatchEntry* e, int n, int total, int shape, int form, hipStream_t st) {
    switch (form) {
        case 0: return launch_fast_batch_t<CAS, S16, 0>(e, n, total, shape, st);
        case 1: return launch_fast_batch_t<CAS, S16, 1>(e, n, total, shape, st);
        case 2: return launch_fast_batch_t<CAS, S16, 2>(e, n, total, shape, st);
        default: return launch_fast_batch_t<CAS, S16, 3>(e, n, total, shape, st);
    }
}

int launch_batch_fast(const BatchEntry* d_entries, int n_entries, int total_tiles, int shape, bool cas, bool s16, int form,
                      void* stream) {
    hipStream_t st = static_cast<hipStream_t>(stream);
    hipError_t e;
    if (cas) e = s16 ? launch_fast_batch_f<true, true>(d_entries, n_entries, total_tiles, shape, form, st)
                     : launch_fast_batch_f<true, false>(d_entries, n_entries, total_tiles, shape, form, st);
    else e = s16 ? launch_fast_batch_f<false, true>(d_entries, n_entries, total_tiles, shape, form, st)
                 : launch_fast_batch_f<false, false>(d_entries, n_entries, total_tiles, shape, form, st);
    return (int)e;
}


}  // namespace taastar
