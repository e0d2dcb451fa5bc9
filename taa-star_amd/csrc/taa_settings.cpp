// The TAA sections of the reference's settings file (include/taastar/taa_settings.h): the key
// lists of source/taa.hpp:1198-1333 and the value formats of source/IniUtil.cpp:50-110, printed
// and parsed the way mINI's generate() and reader treat them (lower-cased names, "key=value").
#include <cctype>
#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <unordered_map>
#include <vector>

#include "taa_internal.hpp"
#include "taastar/taa_settings.h"

namespace {

// ---- one typed field of the settings, by key -----------------------------------------------
enum class Kind { Bool32, Int, Float, Bool, UInt };

struct Field {
    const char* key;  // the reference's key, as written before lower-casing
    Kind kind;
    size_t offset;    // byte offset in taa_settings_t
};

#define PF(name, kind) Field{#name, Kind::kind, offsetof(taa_params_t, name)}
// Parameters keys in the order writeSettingsToIni writes them (taa.hpp:1205-1237)
const Field kParamWrite[] = {
    PF(mPassThrough, Bool32), PF(mColorClampingOrClipping, Int), PF(mShapedNeighbourhood, Bool32),
    PF(mVarianceClipping, Bool32), PF(mVarClipGamma, Float), PF(mUseYCoCg, Bool32),
    PF(mLumaWeightingLottes, Bool32), PF(mDepthCulling, Bool32), PF(mRejectOutside, Bool32),
    PF(mUnjitterNeighbourhood, Bool32), PF(mUnjitterCurrentSample, Bool32), PF(mUnjitterFactor, Float),
    PF(mAlpha, Float), PF(mMinAlpha, Float), PF(mMaxAlpha, Float), PF(mRejectionAlpha, Float),
    PF(mUseVelocityVectors, Int), PF(mVelocitySampleMode, Int), PF(mInterpolationMode, Int),
    PF(mToneMapLumaKaris, Bool32), PF(mAddNoise, Bool32), PF(mNoiseFactor, Float),
    PF(mReduceBlendNearClamp, Bool32), PF(mDynamicAntiGhosting, Bool32),
    Field{"mDebugMask.x", Kind::Float, offsetof(taa_params_t, mDebugMask) + 0},
    Field{"mDebugMask.y", Kind::Float, offsetof(taa_params_t, mDebugMask) + 4},
    Field{"mDebugMask.z", Kind::Float, offsetof(taa_params_t, mDebugMask) + 8},
    Field{"mDebugMask.w", Kind::Float, offsetof(taa_params_t, mDebugMask) + 12},
    PF(mDebugMode, Int), PF(mDebugScale, Float), PF(mDebugCenter, Bool32), PF(mDebugToScreenOutput, Bool32),
    PF(mVelBasedAlpha, Bool32), PF(mVelBasedAlphaMax, Float), PF(mVelBasedAlphaFactor, Float),
    PF(mRayTraceAugment, Bool32),
};
#undef PF

#define SF(key, member, kind) Field{key, Kind::kind, offsetof(taa_settings_t, member)}
// [TAA_Primary] keys before the sample offsets (taa.hpp:1241-1251)
const Field kPrimary[] = {
    SF("mTaaEnabled", taa_enabled, Bool), SF("mSampleDistribution", sample_distribution, Int),
    SF("mSharpener", sharpener, Int), SF("mSharpenFactor", sharpen_factor, Float),
    SF("mSplitScreen", split_screen, Bool), SF("mSplitX", split_x, Int),
    SF("mFixedJitterIndex", fixed_jitter_index, Int), SF("mJitterExtraScale", jitter_extra_scale, Float),
    SF("mJitterSlowMotion", jitter_slow_motion, Int), SF("mJitterRotateDegrees", jitter_rotate_degrees, Float),
    SF("mResetHistoryOnChange", reset_history_on_change, Bool),
};
// [TAA_Postprocess] (taa.hpp:1258-1262)
const Field kPost[] = {
    SF("mPostProcessEnabled", postprocess_enabled, Bool), SF("zoom", zoom, UInt),
    SF("showZoomBox", show_zoom_box, UInt),
    Field{"zoomSrcLTWH.x", Kind::Int, offsetof(taa_settings_t, zoom_src_ltwh) + 0},
    Field{"zoomSrcLTWH.y", Kind::Int, offsetof(taa_settings_t, zoom_src_ltwh) + 4},
    Field{"zoomSrcLTWH.z", Kind::Int, offsetof(taa_settings_t, zoom_src_ltwh) + 8},
    Field{"zoomSrcLTWH.w", Kind::Int, offsetof(taa_settings_t, zoom_src_ltwh) + 12},
    Field{"zoomDstLTWH.x", Kind::Int, offsetof(taa_settings_t, zoom_dst_ltwh) + 0},
    Field{"zoomDstLTWH.y", Kind::Int, offsetof(taa_settings_t, zoom_dst_ltwh) + 4},
    Field{"zoomDstLTWH.z", Kind::Int, offsetof(taa_settings_t, zoom_dst_ltwh) + 8},
    Field{"zoomDstLTWH.w", Kind::Int, offsetof(taa_settings_t, zoom_dst_ltwh) + 12},
};
#undef SF

std::string lower(std::string s) {
    for (char& c : s) c = (char)std::tolower((unsigned char)c);
    return s;
}
std::string trim(const std::string& s) {
    const char* ws = " \t\n\r\f\v";
    const size_t b = s.find_first_not_of(ws);
    if (b == std::string::npos) return std::string();
    return s.substr(b, s.find_last_not_of(ws) - b + 1);
}

// ---- writing (IniUtil.cpp:50-57: std::to_string, "1"/"0") ----------------------------------
void put(std::string& out, const char* key, const std::string& value) {
    out += lower(key);
    out += '=';
    out += value;  // mINI trims values; these have no surrounding whitespace
    out += '\n';
}
std::string fmt_float(float v) {
    char b[512];
    std::snprintf(b, sizeof(b), "%f", (double)v);
    return b;
}
void put_field(std::string& out, const Field& f, const char* base) {
    const char* p = base + f.offset;
    switch (f.kind) {
        case Kind::Bool: put(out, f.key, *(const int32_t*)p ? "1" : "0"); break;
        case Kind::Bool32:
        case Kind::UInt: put(out, f.key, std::to_string(*(const uint32_t*)p)); break;
        case Kind::Int: put(out, f.key, std::to_string(*(const int32_t*)p)); break;
        case Kind::Float: put(out, f.key, fmt_float(*(const float*)p)); break;
    }
}

// ---- reading (IniUtil.cpp:101-110) ----------------------------------------------------------
using Section = std::unordered_map<std::string, std::string>;  // lower-cased key -> value

// parse a value the way the reference's std::stof / stol / stoul + narrowing store it;
// false when no number can be read
bool parse_field(const Field& f, const std::string& v, char* base) {
    char* p = base + f.offset;
    const char* s = v.c_str();
    char* end = nullptr;
    errno = 0;
    switch (f.kind) {
        case Kind::Bool: *(int32_t*)p = (v == "1" || v == "true") ? 1 : 0; return true;
        case Kind::Bool32:
        case Kind::UInt: {
            const unsigned long x = std::strtoul(s, &end, 10);
            if (end == s || errno == ERANGE) return false;
            *(uint32_t*)p = (uint32_t)x;
            return true;
        }
        case Kind::Int: {
            const long x = std::strtol(s, &end, 10);
            if (end == s || errno == ERANGE) return false;
            *(int32_t*)p = (int32_t)x;
            return true;
        }
        case Kind::Float: {
            const float x = std::strtof(s, &end);
            if (end == s || errno == ERANGE) return false;
            *(float*)p = x;
            return true;
        }
    }
    return false;
}

// apply a field if its key is present with a non-empty value (iniRead*: `if (ini[s][k] != "")`)
bool apply(const Section& sec, const Field& f, char* base, std::string& bad) {
    auto it = sec.find(lower(f.key));
    if (it == sec.end() || it->second.empty()) return true;
    if (!parse_field(f, it->second, base)) {
        bad = f.key;
        return false;
    }
    return true;
}

}  // namespace

extern "C" {

int32_t taa_settings_defaults(taa_settings_t* out) {
    if (!out) return taastar::set_last_error(TAA_ERR_INVALID_ARG, "out is NULL");
    std::memset(out, 0, sizeof(*out));
    for (int i = 0; i < 2; ++i) {
        int32_t rc = taa_params_defaults(&out->param[i]);
        if (rc) return rc;
    }
    out->taa_enabled = 1;
    out->sample_distribution = 1;
    out->sharpener = 0;
    out->sharpen_factor = 0.5f;
    out->split_screen = 0;
    out->split_x = 0;
    out->fixed_jitter_index = -1;
    out->jitter_extra_scale = 1.0f;
    out->jitter_slow_motion = 1;
    out->jitter_rotate_degrees = 0.0f;
    out->reset_history_on_change = 1;
    out->num_sample_offsets = 1;  // { {0, 0} }
    out->postprocess_enabled = 1;
    out->zoom = 0;
    out->show_zoom_box = 1;
    const int32_t src[4] = {960 - 10, 540 - 10, 20, 20}, dst[4] = {1920 - 200 - 10, 10, 200, 200};
    std::memcpy(out->zoom_src_ltwh, src, sizeof(src));
    std::memcpy(out->zoom_dst_ltwh, dst, sizeof(dst));
    return TAA_OK;
}

int32_t taa_settings_read_ini(const char* text, size_t len, taa_settings_t* settings) {
    if (!settings || (!text && len)) return taastar::set_last_error(TAA_ERR_INVALID_ARG, "NULL argument");
    // mINI's reader: trimmed lines, ';' comments, "[name]" sections (a ';' on the line ends it),
    // "key=value" split at the first '=' not written as "\="
    std::unordered_map<std::string, Section> ini;
    std::string current;
    size_t pos = 0;
    while (pos < len) {
        size_t e = pos;
        while (e < len && text[e] != '\n') ++e;
        std::string line = trim(std::string(text + pos, e - pos));
        pos = e + 1;
        if (line.empty() || line[0] == ';') continue;
        if (line[0] == '[') {
            std::string l = line;
            const size_t c = l.find(';');
            if (c != std::string::npos) l = l.substr(0, c);
            const size_t close = l.rfind(']');
            if (close != std::string::npos) {
                current = lower(trim(l.substr(1, close - 1)));
                continue;
            }
        }
        std::string norm = line;
        for (size_t i = 0; (i = norm.find("\\=", i)) != std::string::npos; i += 2) norm.replace(i, 2, "  ");
        const size_t eq = norm.find('=');
        if (eq == std::string::npos) continue;
        std::string key = trim(line.substr(0, eq));
        for (size_t i = 0; (i = key.find("\\=", i)) != std::string::npos; ++i) key.replace(i, 2, "=");
        ini[current][lower(key)] = trim(line.substr(eq + 1));
    }

    taa_settings_t s = *settings;  // all or nothing
    char* base = reinterpret_cast<char*>(&s);
    std::string bad;
    for (int ip = 0; ip < 2; ++ip) {
        auto it = ini.find("taa_param_" + std::to_string(ip));
        if (it == ini.end()) continue;
        char* pb = reinterpret_cast<char*>(&s.param[ip]);
        for (const Field& f : kParamWrite)
            if (!apply(it->second, f, pb, bad))
                return taastar::set_last_error(TAA_ERR_INVALID_ARG, ("TAA_Param_" + std::to_string(ip) + "." + bad).c_str());
    }
    auto pr = ini.find("taa_primary");
    if (pr != ini.end()) {
        for (const Field& f : kPrimary)
            if (!apply(pr->second, f, base, bad)) return taastar::set_last_error(TAA_ERR_INVALID_ARG, ("TAA_Primary." + bad).c_str());
        // mDebugSampleOffsets (taa.hpp:1315-1320): resize to max(1, n) (new entries 0), read n
        int32_t n = 0;
        const Field cnt{"mDebugSampleOffsets.size", Kind::Int, 0};
        if (!apply(pr->second, cnt, reinterpret_cast<char*>(&n), bad))
            return taastar::set_last_error(TAA_ERR_INVALID_ARG, "TAA_Primary.mDebugSampleOffsets.size");
        const int32_t keep = n > 1 ? n : 1;
        if (keep > TAA_SETTINGS_MAX_SAMPLE_OFFSETS)
            return taastar::set_last_error(TAA_ERR_CAPACITY, "more than TAA_SETTINGS_MAX_SAMPLE_OFFSETS sample offsets");
        for (int32_t i = s.num_sample_offsets; i < keep; ++i) s.sample_offsets[i][0] = s.sample_offsets[i][1] = 0.0f;
        s.num_sample_offsets = keep;
        for (int32_t i = 0; i < n; ++i) {
            const std::string k = "mDebugSampleOffsets_" + std::to_string(i), kx = k + ".x", ky = k + ".y";
            const size_t at = offsetof(taa_settings_t, sample_offsets) + (size_t)i * 8;
            const Field gx{kx.c_str(), Kind::Float, at}, gy{ky.c_str(), Kind::Float, at + 4};
            if (!apply(pr->second, gx, base, bad) || !apply(pr->second, gy, base, bad))
                return taastar::set_last_error(TAA_ERR_INVALID_ARG, ("TAA_Primary." + bad).c_str());
        }
    }
    auto pp = ini.find("taa_postprocess");
    if (pp != ini.end())
        for (const Field& f : kPost)
            if (!apply(pp->second, f, base, bad)) return taastar::set_last_error(TAA_ERR_INVALID_ARG, ("TAA_Postprocess." + bad).c_str());
    *settings = s;
    return TAA_OK;
}

int32_t taa_settings_write_ini(const taa_settings_t* settings, char* out, size_t capacity, size_t* length) {
    if (!settings || !length) return taastar::set_last_error(TAA_ERR_INVALID_ARG, "NULL argument");
    const int32_t n = settings->num_sample_offsets;
    if (n < 0 || n > TAA_SETTINGS_MAX_SAMPLE_OFFSETS) return taastar::set_last_error(TAA_ERR_INVALID_ARG, "num_sample_offsets out of range");
    std::string t;
    // mINI generate(): "[section]\n" + "key=value" lines joined by '\n', sections joined by '\n'
    for (int ip = 0; ip < 2; ++ip) {
        t += "[taa_param_" + std::to_string(ip) + "]\n";
        const char* pb = reinterpret_cast<const char*>(&settings->param[ip]);
        for (const Field& f : kParamWrite) put_field(t, f, pb);
    }
    const char* base = reinterpret_cast<const char*>(settings);
    t += "[taa_primary]\n";
    for (const Field& f : kPrimary) put_field(t, f, base);
    put(t, "mDebugSampleOffsets.size", std::to_string(n));
    for (int32_t i = 0; i < n; ++i) {
        const std::string k = "mDebugSampleOffsets_" + std::to_string(i);
        put(t, (k + ".x").c_str(), fmt_float(settings->sample_offsets[i][0]));
        put(t, (k + ".y").c_str(), fmt_float(settings->sample_offsets[i][1]));
    }
    t += "[taa_postprocess]\n";
    for (const Field& f : kPost) put_field(t, f, base);
    t.pop_back();  // the generator ends the last line without '\n'
    *length = t.size();
    if (!out || capacity < t.size() + 1) return taastar::set_last_error(TAA_ERR_CAPACITY, "output buffer too small");
    std::memcpy(out, t.c_str(), t.size() + 1);
    return TAA_OK;
}

int32_t taa_settings_apply(const taa_settings_t* s, taa_uniforms_t* u, taa_jitter_config_t* j) {
    if (!s) return taastar::set_last_error(TAA_ERR_INVALID_ARG, "settings is NULL");
    if (u) {  // taa.hpp:938-951
        u->param[0] = s->param[0];
        u->param[1] = s->param[1];
        u->splitScreen = s->split_screen;
        u->splitX = s->split_x;
    }
    if (j) {  // the members get_jitter_offset_for_frame reads (taa.hpp:150-233)
        j->sample_distribution = s->sample_distribution;
        j->slow_motion = s->jitter_slow_motion;
        j->fixed_index = s->fixed_jitter_index;
        j->rotate_degrees = s->jitter_rotate_degrees;
        j->extra_scale = s->jitter_extra_scale;
        j->num_user_offsets = s->num_sample_offsets;
        j->user_offsets = &s->sample_offsets[0][0];
    }
    return TAA_OK;
}

}  // extern "C"
