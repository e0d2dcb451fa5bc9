// Host-side helpers of the TAA driver: CAS constants, Halton jitter and the jitter selection of
// taa<CF>::get_jitter_offset_for_frame (source/taa.hpp:150-233).
#include <cmath>
#include <cstdint>
#include <cstring>

#include "taastar/taa_abi.h"

namespace {

uint32_t bits(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    return u;
}

// fp32 -> fp16 bit pattern by mantissa truncation, saturating overflow/Inf/NaN to +-65504 and
// keeping denormals: the conversion CasSetup uses on the CPU (ffx_a.h:475-541).
uint32_t half_trunc(float f) {
    const uint32_t u = bits(f);
    const uint32_t sign = (u >> 16) & 0x8000u;
    const int exp = (int)((u >> 23) & 0xffu) - 127;
    const uint32_t man = u & 0x7fffffu;
    if (exp > 15) return sign | 0x7bffu;
    if (exp >= -14) return sign | ((uint32_t)(exp + 15) << 10) | (man >> 13);
    if (exp >= -24) return sign | ((man | 0x800000u) >> (uint32_t)(-exp - 1));
    return sign;
}

float sat(float v) { return v < 0.0f ? 0.0f : (v > 1.0f ? 1.0f : v); }

}  // namespace

extern "C" {

// CasSetup (shaders/ffx_cas.h:375-394): scaling terms, then peak = -1 / lerp(8, 5, sat(s)) with
// ALerpF1(a, b, t) = b*t + (-a*t + a) (ffx_a.h:302).
int32_t taa_cas_setup(uint32_t const0[4], uint32_t const1[4], float sharpness, float in_w, float in_h, float out_w,
                      float out_h) {
    if (!const0 || !const1) return TAA_ERR_INVALID_ARG;
    const float rx = 1.0f / out_w, ry = 1.0f / out_h;
    const0[0] = bits(in_w * rx);
    const0[1] = bits(in_h * ry);
    const0[2] = bits(0.5f * in_w * rx - 0.5f);
    const0[3] = bits(0.5f * in_h * ry - 0.5f);
    const float t = sat(sharpness);
    const float peak = -(1.0f / (5.0f * t + (-8.0f * t + 8.0f)));
    const1[0] = bits(peak);
    const1[1] = half_trunc(peak) | (half_trunc(0.0f) << 16);
    const1[2] = bits(8.0f * in_w * rx);
    const1[3] = 0;
    return TAA_OK;
}

// helpers::halton (source/helper_functions.hpp:10-17), same float recurrence.
float taa_halton(int32_t index, int32_t base) {
    if (base < 2) return 0.0f;
    float scale = 1.0f, acc = 0.0f;
    for (int32_t i = index; i > 0; i /= base) {
        scale = scale / (float)base;
        acc = acc + scale * (float)(i % base);
    }
    return acc;
}

int32_t taa_jitter_offset(const taa_jitter_config_t* cfg, uint64_t frame_id, uint32_t width, uint32_t height,
                          float out_ndc[2], uint32_t* out_num_samples) {
    if (!cfg || !out_ndc || width == 0 || height == 0) return TAA_ERR_INVALID_ARG;
    // pixel size in NDC for this resolution (the reference caches the first call's, taa.hpp:154-155)
    const float pxw = 2.0f / (float)width, pxh = 2.0f / (float)height;
    float ox = 0.f, oy = 0.f;  // offset in pixel units
    uint32_t n = 0;
    bool user = false;
    uint64_t f = frame_id;
    if (cfg->slow_motion > 1) f /= (uint64_t)cfg->slow_motion;
    if (cfg->fixed_index >= 0) f = (uint64_t)cfg->fixed_index;
    switch (cfg->sample_distribution) {
        case 0:
        case 1: {  // circular quad / uniform-4 helix (taa.hpp:158-169)
            static const float quad[4][2] = {{-0.25f, -0.25f}, {0.25f, -0.25f}, {0.25f, 0.25f}, {-0.25f, 0.25f}};
            static const float helix[4][2] = {{-0.25f, -0.25f}, {0.25f, 0.25f}, {0.25f, -0.25f}, {-0.25f, 0.25f}};
            n = 4;
            const float(*t)[2] = cfg->sample_distribution == 0 ? quad : helix;
            ox = t[f % 4][0];
            oy = t[f % 4][1];
            break;
        }
        case 2:
        case 3: {  // Halton(2,3) x8 / x16 (taa.hpp:170-171, helper_functions.hpp:19-26)
            n = cfg->sample_distribution == 2 ? 8u : 16u;
            const int32_t i = (int32_t)(f % n) + 1;
            ox = taa_halton(i, 2) - 0.5f;
            oy = taa_halton(i, 3) - 0.5f;
            break;
        }
        case 4: {  // regular 4x4 grid at odd eighths (taa.hpp:173-179)
            n = 16;
            const int k = (int)(f % 16);
            const float eighth = 1.f / 8.f;
            ox = (float)(2 * (k % 4) - 3) * eighth;
            oy = (float)(2 * (k / 4) - 3) * eighth;
            break;
        }
        case 5: {  // user offsets in pixel units (taa.hpp:207-211)
            if (cfg->num_user_offsets <= 0 || !cfg->user_offsets) return TAA_ERR_INVALID_ARG;
            n = (uint32_t)cfg->num_user_offsets;
            const uint64_t k = f % n;
            ox = cfg->user_offsets[2 * k];
            oy = cfg->user_offsets[2 * k + 1];
            user = true;
            break;
        }
        default: return TAA_ERR_INVALID_ARG;
    }
    // the tables hold pxSize * offset; distribution 5 holds the offset and scales at selection
    float x = user ? ox * pxw : pxw * ox;
    float y = user ? oy * pxh : pxh * oy;
    if (cfg->rotate_degrees != 0.f) {
        const float rad = cfg->rotate_degrees * 0.01745329251994329576923690768489f;
        const float s = std::sin(rad), c = std::cos(rad);
        const float nx = x * c - y * s, ny = x * s + y * c;
        x = nx;
        y = ny;
    }
    out_ndc[0] = x * cfg->extra_scale;
    out_ndc[1] = y * cfg->extra_scale;
    if (out_num_samples) *out_num_samples = n;
    return TAA_OK;
}

}  // extern "C"
