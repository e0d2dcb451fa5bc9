// Host runtime behind the C ABI (include/taastar/taa_abi.h): argument checking, the per-frame
// launch, sessions with HBM-budget admission, batched and graph-captured launches, autotuner.
// The MI355X counterpart of the host driver taa<CF> (source/taa.hpp).
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <new>
#include <string>
#include <tuple>
#include <vector>

#include "taa_internal.hpp"

// ---- layout of the reference structs ------------------------------------------------------
static_assert(sizeof(taa_params_t) == 176, "Parameters must be 176 bytes (source/taa.hpp:30-77)");
static_assert(offsetof(taa_params_t, mVarClipGamma) == 48, "layout");
static_assert(offsetof(taa_params_t, mRejectOutside) == 64, "layout");
static_assert(offsetof(taa_params_t, mInterpolationMode) == 76, "layout");
static_assert(offsetof(taa_params_t, mRayTraceHistoryCount) == 140, "layout");
static_assert(offsetof(taa_params_t, mDebugMask) == 144, "layout");
static_assert(offsetof(taa_params_t, mDebugToScreenOutput) == 172, "layout");
static_assert(sizeof(taa_uniforms_t) == 544, "uniforms_for_taa must be 544 bytes (source/taa.hpp:113-129)");
static_assert(offsetof(taa_uniforms_t, param) == 128, "layout");
static_assert(offsetof(taa_uniforms_t, mJitterNdc) == 480, "layout");
static_assert(offsetof(taa_uniforms_t, mSinTime) == 496, "layout");
static_assert(offsetof(taa_uniforms_t, splitScreen) == 512, "layout");
static_assert(offsetof(taa_uniforms_t, mResetHistory) == 528, "layout");
static_assert(offsetof(taa_uniforms_t, mCamNearPlane) == 532, "layout");
static_assert(offsetof(taa_uniforms_t, pad1) == 540, "layout");
static_assert(sizeof(taa_cas_consts_t) == 32, "push_constants_for_cas must be 32 bytes (source/taa.hpp:88-91)");
static_assert(sizeof(taa_edge_aa_consts_t) == 20, "push_constants_for_fxaa is 20 bytes (source/taa.hpp:93-98)");
static_assert(sizeof(taastar::BatchEntry) % 16 == 0, "BatchEntry alignment");

namespace {

thread_local std::string g_err;

int fail(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}
int ok() { return TAA_OK; }
int hip_fail(hipError_t e, const char* what) { return fail(TAA_ERR_HIP, "%s: %s", what, hipGetErrorString(e)); }

size_t color_bpp(int fmt) { return fmt == TAA_FMT_RGBA16F ? 8 : 16; }
size_t vel_bpp(int fmt) {
    switch (fmt) {
        case TAA_VEL_RG32F: return 8;
        case TAA_VEL_RG16F: return 4;
        case TAA_VEL_RGBA32F: return 16;
        default: return 8;
    }
}
bool valid_color_fmt(int f) { return f == TAA_FMT_RGBA32F || f == TAA_FMT_RGBA16F; }
bool valid_vel_fmt(int f) { return f >= TAA_VEL_RG32F && f <= TAA_VEL_RGBA16F; }
size_t align256(size_t v) { return (v + 255) & ~size_t(255); }

// ---- process-wide tile shape table ----------------------------------------------------------
using ClassKey = std::tuple<int, int, int, int, int>;  // W, H, fmt, arith, cas
std::mutex g_tune_mu;
std::map<ClassKey, int> g_tune;

// The built-in shape of an untuned class. With CAS, a frame of fewer 30x30 tiles than the device
// holds workgroups at once (1,280) leaves CUs idle: smaller tiles spread it wider (single FP16
// frames, per-frame launches: 640x360 at 15.4 Gpix/s on 30x30, 21.9 on 30x14, 26.0 on 14x14;
// 1280x720 22.8 µs on 30x14 against 23.3 on 30x30; from 1920x1080 up 30x30 is the fastest).
int default_shape(int W, int H, bool cas) {
    static_assert(taastar::kShapesCas[0].tw == 14 && taastar::kShapesCas[0].th == 14, "shape 0 is 14x14");
    static_assert(taastar::kShapesCas[1].tw == 30 && taastar::kShapesCas[1].th == 14, "shape 1 is 30x14");
    if (!cas) return taastar::kDefaultShapePlain;
    const long long n30 = (long long)taastar::tiles_for(W, 30) * taastar::tiles_for(H, 30);
    if (n30 < 512) return 0;
    if (n30 < 1280) return 1;
    return taastar::kDefaultShapeCas;
}

int selected_shape(int W, int H, int fmt, int arith, bool cas) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto it = g_tune.find(ClassKey{W, H, fmt, arith, cas ? 1 : 0});
    if (it != g_tune.end()) return it->second;
    return default_shape(W, H, cas);
}

// The shape for a launch over the tiles of several frames: such a launch behaves like one frame of
// their total size, so it takes the shape tuned for the largest class of at most that many pixels
// (the default when none is tuned). Measured with every class tuned on single frames: 64 sessions
// of 640x360 took 14x14 tiles from their own class, 48.2 Gpix/s batched, against 61.9 with 30x30.
int selected_shape_for_pixels(long long pixels, int fmt, int arith, bool cas) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    long long best_px = -1;
    int shape = cas ? taastar::kDefaultShapeCas : taastar::kDefaultShapePlain;
    for (const auto& e : g_tune) {
        const ClassKey& k = e.first;
        if (std::get<2>(k) != fmt || std::get<3>(k) != arith || std::get<4>(k) != (cas ? 1 : 0)) continue;
        const long long px = (long long)std::get<0>(k) * std::get<1>(k);
        if (px <= pixels && px > best_px) {
            best_px = px;
            shape = e.second;
        }
    }
    return shape;
}

taastar::TileShapeDef shape_def(int idx, bool cas) { return cas ? taastar::kShapesCas[idx] : taastar::kShapesPlain[idx]; }

// Fill the kernel-side frame descriptor; validates the images.
// The kernels address every image with 32-bit byte offsets: the largest image (16 bytes per
// texel; 64 for the inspect image) must stay below 4 GiB.
bool fits_offsets(int pitch, int height, int bytes_per_texel) {
    return (uint64_t)pitch * (uint64_t)height * (uint64_t)bytes_per_texel <= 0xFFFFFFFFull;
}

int make_desc(const taa_frame_images_t* img, const taa_cas_consts_t* cas, taastar::FrameDesc& fd) {
    if (!img) return fail(TAA_ERR_INVALID_ARG, "images is NULL");
    if (img->width <= 0 || img->height <= 0 || img->width > 32768 || img->height > 32768)
        return fail(TAA_ERR_INVALID_ARG, "bad frame size %dx%d", img->width, img->height);
    const int pitch = img->pitch == 0 ? img->width : img->pitch;
    if (pitch < img->width) return fail(TAA_ERR_INVALID_ARG, "pitch %d < width %d", pitch, img->width);
    if (!fits_offsets(pitch, img->height, 16) || (img->inspect && !fits_offsets(pitch, img->height, 4 * TAA_INSPECT_WORDS)))
        return fail(TAA_ERR_INVALID_ARG, "frame of %d x %d texels: an image would exceed 4 GiB", pitch, img->height);
    const int wl = img->input_width == 0 ? img->width : img->input_width;
    const int hl = img->input_height == 0 ? img->height : img->input_height;
    const int pl = img->input_pitch != 0 ? img->input_pitch : (img->input_width == 0 ? pitch : wl);
    if (wl <= 0 || hl <= 0 || wl > 32768 || hl > 32768 || pl < wl)
        return fail(TAA_ERR_INVALID_ARG, "bad input size %dx%d (pitch %d)", wl, hl, pl);
    if (!fits_offsets(pl, hl, 16)) return fail(TAA_ERR_INVALID_ARG, "input of %d x %d texels exceeds 4 GiB", pl, hl);
    if (img->reserved[0] != 0 || img->reserved[1] != 0) return fail(TAA_ERR_INVALID_ARG, "reserved must be 0");
    if (!valid_color_fmt(img->color_format)) return fail(TAA_ERR_INVALID_ARG, "bad color_format %d", img->color_format);
    if (!valid_vel_fmt(img->velocity_format)) return fail(TAA_ERR_INVALID_ARG, "bad velocity_format %d", img->velocity_format);
    if (!img->color || !img->depth || !img->history || !img->history_depth || !img->result_screen || !img->velocity ||
        !img->result_history)
        return fail(TAA_ERR_INVALID_ARG, "a required image pointer is NULL");
    // every workgroup reads its apron and its history footprint from other tiles' texels while
    // other workgroups write theirs, so no output may share a buffer with an input or another output
    if (img->history == img->result_history) return fail(TAA_ERR_INVALID_ARG, "history and result_history alias");
    const void* const ins[] = {img->color, img->depth, img->history, img->history_depth, img->velocity};
    const void* const outs[] = {img->result_screen, img->result_history, img->debug, img->inspect};
    const char* const out_names[] = {"result_screen", "result_history", "debug", "inspect"};
    for (int o = 0; o < 4; ++o) {
        if (!outs[o]) continue;
        for (const void* in : ins)
            if (outs[o] == in) return fail(TAA_ERR_INVALID_ARG, "%s aliases an input image", out_names[o]);
        for (int p = o + 1; p < 4; ++p)
            if (outs[o] == outs[p]) return fail(TAA_ERR_INVALID_ARG, "%s and %s alias", out_names[o], out_names[p]);
    }
    std::memset(&fd, 0, sizeof(fd));
    fd.color = img->color;
    fd.depth = img->depth;
    fd.history = img->history;
    fd.hdepth = img->history_depth;
    fd.screen = img->result_screen;
    fd.debug = img->debug;
    fd.velocity = img->velocity;
    fd.history_out = img->result_history;
    fd.inspect = img->inspect;
    fd.W = img->width;
    fd.H = img->height;
    fd.pitch = pitch;
    fd.color_fmt = img->color_format;
    fd.vel_fmt = img->velocity_format;
    fd.x_begin = 0;
    fd.x_end = img->width;
    fd.param_idx = 0;
    fd.cas_on = cas ? 1 : 0;
    fd.cas_peak_bits = cas ? cas->const1[0] : 0u;
    fd.Wl = wl;
    fd.Hl = hl;
    fd.pitch_l = pl;
    fd.up = (wl != img->width || hl != img->height) ? 1 : 0;
    return TAA_OK;
}

int check_uniforms(const taa_uniforms_t* u, bool segmented = false) {
    if (!u) return fail(TAA_ERR_INVALID_ARG, "uniforms is NULL");
    for (int i = 0; i < 2; ++i) {
        if (u->param[i].mRayTraceAugment && !segmented)
            return fail(TAA_ERR_UNSUPPORTED, "mRayTraceAugment needs the segmentation images (taa_resolve_frame_segmented)");
        if (u->param[i].mInterpolationMode < 0 || u->param[i].mInterpolationMode > 2)
            return fail(TAA_ERR_INVALID_ARG, "mInterpolationMode %d", u->param[i].mInterpolationMode);
    }
    return TAA_OK;
}

// The kernel a resolve of [x_begin, x_end) with parameter block param_idx runs (TAA_KERNEL_*);
// an FP16 request the FP16 kernel does not cover records the reason in *why.
int kernel_for(const taastar::FrameDesc& fd, const taa_uniforms_t& U, int param_idx, int arith, const char** why) {
    *why = nullptr;
    if (fd.up || U.mUpsampling) return arith == TAA_ARITH_FP16 ? TAA_KERNEL_UPSAMPLE_FP16 : TAA_KERNEL_UPSAMPLE;
    if (arith == TAA_ARITH_FP16 && taastar::fast_path_supported(fd, U, param_idx, why)) return TAA_KERNEL_FP16;
    return TAA_KERNEL_EXACT;
}

// Launch the resolve over [x_begin, x_end) with the given parameter block.
int launch_range(taastar::FrameDesc fd, const taa_uniforms_t& U, int x_begin, int x_end, int param_idx, int arith, void* stream,
                 const taastar::SegDesc* sg = nullptr) {
    if (x_end <= x_begin) return TAA_OK;
    const bool cas = fd.cas_on != 0;
    if (sg) {  // segmentation mask: its own kernels, tiled by the launcher
        fd.x_begin = x_begin;
        fd.x_end = x_end;
        fd.param_idx = param_idx;
        if (U.mUpsampling) fd.up = 1;
        const int e = taastar::launch_resolve_seg(fd, U, *sg, arith, stream);
        if (e != 0) return hip_fail((hipError_t)e, "resolve launch");
        return TAA_OK;
    }
    const int shape = selected_shape(fd.W, fd.H, fd.color_fmt, arith, cas);
    const taastar::TileShapeDef s = shape_def(shape, cas);
    fd.x_begin = x_begin;
    fd.x_end = x_end;
    fd.param_idx = param_idx;
    fd.tiles_x = taastar::tiles_for(x_end - x_begin, s.tw);
    fd.tiles_y = taastar::tiles_for(fd.H, s.th);
    if (fd.up || U.mUpsampling) {  // upsampling path: the launcher tiles by its form's own shape
        fd.up = 1;
        const int e = taastar::launch_resolve_up(fd, U, arith, stream);
        if (e != 0) return hip_fail((hipError_t)e, "resolve launch");
        return TAA_OK;
    }
    int e = (arith == TAA_ARITH_FP16 && taastar::fast_path_supported(fd, U, param_idx))
                ? taastar::launch_resolve_fast(fd, U, shape, stream)
                : taastar::launch_resolve(fd, U, shape, arith, stream);
    if (e != 0) return hip_fail((hipError_t)e, "resolve launch");
    return TAA_OK;
}

// Argument checks of a resolve and its frame descriptor (and segmentation descriptor).
int prepare_resolve(const taa_uniforms_t* u, const taa_frame_images_t* img, const taa_cas_consts_t* cas, int arith,
                    const taa_segmentation_images_t* seg, taastar::FrameDesc& fd, taastar::SegDesc& sgd) {
    int rc = check_uniforms(u, seg != nullptr);
    if (rc) return rc;
    if (arith != TAA_ARITH_FP32 && arith != TAA_ARITH_FP16) return fail(TAA_ERR_INVALID_ARG, "bad arith %d", arith);
    rc = make_desc(img, cas, fd);
    if (rc) return rc;
    if (seg) {
        if (!seg->material || !seg->prev_material || !seg->uv_normal || !seg->prev_segmask || !seg->segmask)
            return fail(TAA_ERR_INVALID_ARG, "a segmentation image pointer is NULL");
        if (seg->reserved[0] || seg->reserved[1]) return fail(TAA_ERR_INVALID_ARG, "reserved must be 0");
        const int sp = seg->pitch == 0 ? fd.pitch : seg->pitch;
        if (sp < fd.W) return fail(TAA_ERR_INVALID_ARG, "segmentation pitch %d < width %d", sp, fd.W);
        if (!fits_offsets(sp, fd.H, 16)) return fail(TAA_ERR_INVALID_ARG, "a segmentation image would exceed 4 GiB");
        // the renderer's images: input-sized; at one size with the masks' pitch unless input_pitch says otherwise
        const bool one_size = fd.Wl == fd.W && fd.Hl == fd.H;
        if (seg->input_pitch < 0) return fail(TAA_ERR_INVALID_ARG, "segmentation input_pitch %d < 0", seg->input_pitch);
        const int spl = seg->input_pitch != 0 ? seg->input_pitch : (one_size ? sp : fd.pitch_l);
        if (spl < fd.Wl) return fail(TAA_ERR_INVALID_ARG, "segmentation input pitch %d < input width %d", spl, fd.Wl);
        if (!fits_offsets(spl, fd.Hl, 16)) return fail(TAA_ERR_INVALID_ARG, "a segmentation input image would exceed 4 GiB");
        const void* m = seg->segmask;
        if (m == seg->prev_segmask || m == seg->material || m == seg->prev_material || m == seg->uv_normal || m == fd.color ||
            m == fd.depth || m == fd.history || m == fd.hdepth || m == fd.velocity || m == fd.history_out || m == fd.screen)
            return fail(TAA_ERR_INVALID_ARG, "segmask aliases another image");
        sgd = taastar::SegDesc{seg->material, seg->prev_material, seg->uv_normal, seg->prev_segmask, seg->segmask, sp, spl};
    }
    return TAA_OK;
}

// first column of parameter block 1: paramsIdx = (splitScreen && x > splitX) ? 1 : 0 (taa.comp:717)
int split_column(const taa_uniforms_t& u, int W) { return std::min(std::max(u.splitX + 1, 0), W); }

// Bytes of the intermediate screen of split screen with CAS.
size_t split_bytes(const taastar::FrameDesc& fd) { return (size_t)fd.pitch * (size_t)fd.H * color_bpp(fd.color_fmt); }

// The intermediate screen of split screen with CAS taa_resolve_frame uses: one buffer per device,
// allocated on first use and grown on demand, kept until the process ends. Calls on different
// streams are ordered through an event recorded after each call's CAS pass.
struct SplitScratch {
    std::mutex mu;
    void* ptr = nullptr;
    size_t bytes = 0;
    hipEvent_t ev = nullptr;
};
constexpr int kMaxDevices = 64;
SplitScratch g_split[kMaxDevices];

// Split screen with CAS, in the reference's order (taa.hpp:1008-1027, 1111-1139): both parts of
// the resolve write the unsharpened screen to `tmp`, then the CAS pass reads it into the result.
int resolve_split_cas(taastar::FrameDesc fd, const taa_uniforms_t& u, const taa_cas_consts_t* cas, int arith, void* stream,
                      const taastar::SegDesc* sg, void* tmp) {
    void* const screen = fd.screen;
    fd.screen = tmp;
    fd.cas_on = 0;
    fd.cas_peak_bits = 0u;
    const int split = split_column(u, fd.W);
    int rc = launch_range(fd, u, 0, split, 0, arith, stream, sg);
    if (rc) return rc;
    rc = launch_range(fd, u, split, fd.W, 1, arith, stream, sg);
    if (rc) return rc;
    const int e = taastar::launch_cas(tmp, screen, fd.W, fd.H, fd.pitch, fd.color_fmt, cas->const1[0], stream);
    if (e) return hip_fail((hipError_t)e, "cas launch");
    return TAA_OK;
}

int split_cas_library(const taastar::FrameDesc& fd, const taa_uniforms_t& u, const taa_cas_consts_t* cas, int arith,
                      void* stream, const taastar::SegDesc* sg) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
    if (dev < 0 || dev >= kMaxDevices) return fail(TAA_ERR_NO_DEVICE, "device %d", dev);
    SplitScratch& sc = g_split[dev];
    std::lock_guard<std::mutex> lk(sc.mu);
    if (!sc.ev && (e = hipEventCreateWithFlags(&sc.ev, hipEventDisableTiming)) != hipSuccess) {
        sc.ev = nullptr;
        return hip_fail(e, "hipEventCreate");
    }
    const size_t need = split_bytes(fd);
    if (sc.bytes < need) {
        if (sc.ptr) {
            if ((e = hipEventSynchronize(sc.ev)) != hipSuccess) return hip_fail(e, "split scratch wait");
            hipFree(sc.ptr);
            sc.ptr = nullptr;
            sc.bytes = 0;
        }
        if ((e = hipMalloc(&sc.ptr, need)) != hipSuccess) {
            sc.ptr = nullptr;
            return fail(TAA_ERR_OUT_OF_MEMORY, "split screen intermediate (%zu bytes): %s", need, hipGetErrorString(e));
        }
        sc.bytes = need;
    }
    hipStream_t st = static_cast<hipStream_t>(stream);
    if ((e = hipStreamWaitEvent(st, sc.ev, 0)) != hipSuccess) return hip_fail(e, "split scratch order");
    const int rc = resolve_split_cas(fd, u, cas, arith, stream, sg, sc.ptr);
    hipEventRecord(sc.ev, st);
    return rc;
}

// split_scratch: NULL = the library's per-device intermediate for split screen with CAS; else a
// device buffer of split_scratch_bytes the caller provides (refused when too small).
int resolve_impl(const taa_uniforms_t* u, const taa_frame_images_t* img, const taa_cas_consts_t* cas, int arith, void* stream,
                 const taa_segmentation_images_t* seg = nullptr, void* split_scratch = nullptr,
                 size_t split_scratch_bytes = 0) {
    taastar::FrameDesc fd;
    taastar::SegDesc sgd;
    int rc = prepare_resolve(u, img, cas, arith, seg, fd, sgd);
    if (rc) return rc;
    const taastar::SegDesc* sg = seg ? &sgd : nullptr;
    if (!u->splitScreen) return launch_range(fd, *u, 0, fd.W, 0, arith, stream, sg);
    if (cas && split_scratch) {
        const size_t need = split_bytes(fd);
        if (need > split_scratch_bytes)
            return fail(TAA_ERR_CAPACITY, "split screen with CAS needs a %zu-byte intermediate screen; the scratch holds %zu",
                        need, split_scratch_bytes);
        return resolve_split_cas(fd, *u, cas, arith, stream, sg, split_scratch);
    }
    if (cas) return split_cas_library(fd, *u, cas, arith, stream, sg);
    const int split = split_column(*u, fd.W);
    rc = launch_range(fd, *u, 0, split, 0, arith, stream, sg);
    if (rc) return rc;
    return launch_range(fd, *u, split, fd.W, 1, arith, stream, sg);
}

}  // namespace

// ---- contexts and sessions ----------------------------------------------------------------

struct taa_session;

struct taa_ctx {
    int device = 0;
    uint64_t budget = 0;
    std::atomic<uint64_t> used{0};
    std::mutex mu;
    // batch staging ring: pinned host mirror + device copy of entries and uniforms
    static constexpr int kSlots = 16;
    static constexpr int kCap = TAA_BATCH_MAX_ITEMS;
    struct Slot {
        void* host = nullptr;
        void* dev = nullptr;
        hipEvent_t done = nullptr;
        bool used = false;
    } slots[kSlots];
    int next_slot = 0;
    hipStream_t tune_stream = nullptr;
};

struct taa_session {
    taa_ctx* ctx = nullptr;
    int W = 0, H = 0, pitch = 0, cf = 0, vf = 0;
    int Wl = 0, Hl = 0;  // input (render) size of the colour, depth and velocity slots
    uint64_t bytes = 0;
    void* base = nullptr;
    void* hist[2] = {nullptr, nullptr};
    float* depth[2] = {nullptr, nullptr};
    void* color = nullptr;
    void* vel = nullptr;
    uint64_t frames = 0;
    bool force_reset = true;
};

namespace {

size_t staging_bytes() { return sizeof(taastar::BatchEntry) * taa_ctx::kCap; }

// Frame descriptor for the session's next frame (frame k reads depth[k&1], history[(k+1)&1] and
// writes history[k&1]; the previous depth is depth[(k+1)&1]).
void session_desc(taa_session* s, void* screen, const taa_cas_consts_t* cas, taastar::FrameDesc& fd) {
    const int k = (int)(s->frames & 1);
    std::memset(&fd, 0, sizeof(fd));
    fd.color = s->color;
    fd.depth = s->depth[k];
    fd.history = s->hist[k ^ 1];
    fd.hdepth = s->depth[k ^ 1];
    fd.screen = screen;
    fd.velocity = s->vel;
    fd.history_out = s->hist[k];
    fd.W = s->W;
    fd.H = s->H;
    fd.pitch = s->pitch;
    fd.color_fmt = s->cf;
    fd.vel_fmt = s->vf;
    fd.x_begin = 0;
    fd.x_end = s->W;
    fd.cas_on = cas ? 1 : 0;
    fd.cas_peak_bits = cas ? cas->const1[0] : 0u;
    fd.Wl = s->Wl;
    fd.Hl = s->Hl;
    fd.pitch_l = s->Wl;
    fd.up = (s->Wl != s->W || s->Hl != s->H) ? 1 : 0;
}

// frames that take the upsampling resolve (a launch of their own in a batch, in the batch's arithmetic)
bool session_upsamples(const taa_session* s, const taa_uniforms_t* u) {
    return s->Wl != s->W || s->Hl != s->H || u->mUpsampling;
}

void session_advance(taa_session* s) {
    s->frames++;
    s->force_reset = false;
}

// After work that reads a pinned staging buffer has been enqueued on `st`: mark the buffer as in
// use until `done` passes. If the event cannot be recorded the stream is drained instead, so that
// the buffer is never rewritten while a copy from it may still be pending.
void staging_enqueued(hipEvent_t done, bool& used, hipStream_t st) {
    if (hipEventRecord(done, st) == hipSuccess) {
        used = true;
    } else {
        hipStreamSynchronize(st);
        used = false;
    }
}

}  // namespace

namespace taastar {
// What a batched frame's uniforms must satisfy (also checked by the server when a frame is queued).
int check_batch_uniforms(const taa_uniforms_t* u) {
    if (int rc = check_uniforms(u)) return rc;
    if (u->splitScreen) return fail(TAA_ERR_UNSUPPORTED, "split screen in a batch");
    return TAA_OK;
}
}  // namespace taastar

extern "C" {

int32_t taa_abi_version(void) { return TAASTAR_ABI_VERSION; }
const char* taa_last_error(void) { return g_err.c_str(); }

int32_t taa_params_defaults(taa_params_t* p) {
    if (!p) return fail(TAA_ERR_INVALID_ARG, "out is NULL");
    std::memset(p, 0, sizeof(*p));
    // source/taa.hpp:31-76
    p->mAlpha = 0.05f;
    p->mColorClampingOrClipping = 1;
    p->mUnjitterFactor = 1.0f;
    p->mVarClipGamma = 1.0f;
    p->mMinAlpha = 1.0f - 0.97f;
    p->mMaxAlpha = 1.0f - 0.88f;
    p->mRejectionAlpha = 1.0f;
    p->mUseVelocityVectors = 1;
    p->mNoiseFactor = 1.f / 510.f;
    p->mVelBasedAlphaMax = 0.2f;
    p->mVelBasedAlphaFactor = 1.f / 40.f;
    p->mRayTraceAugmentFlags = 0xffffffffu & ~(0x0080u | 0x0100u);  // ~(TAA_RTFLAG_ALL | TAA_RTFLAG_FXD)
    p->mRayTraceAugment_WNrm = 0.5f;
    p->mRayTraceAugment_WDpt = 0.015f;
    p->mRayTraceAugment_WMId = 0.25f;
    p->mRayTraceAugment_WLum = 0.5f;
    p->mRayTraceAugment_Thresh = 0.5f;
    p->mRayTraceHistoryCount = -1;
    p->mDebugMask[0] = p->mDebugMask[1] = p->mDebugMask[2] = 1.0f;
    p->mDebugScale = 1.0f;
    return ok();
}

int32_t taa_resolve_frame(const taa_uniforms_t* uniforms, const taa_frame_images_t* images, const taa_cas_consts_t* cas,
                          int32_t arith, void* stream) {
    return resolve_impl(uniforms, images, cas, arith, stream);
}

int32_t taa_resolve_frame_segmented(const taa_uniforms_t* uniforms, const taa_frame_images_t* images,
                                    const taa_segmentation_images_t* seg, const taa_cas_consts_t* cas, int32_t arith,
                                    void* stream) {
    if (!seg) return fail(TAA_ERR_INVALID_ARG, "seg is NULL");
    return resolve_impl(uniforms, images, cas, arith, stream, seg);
}

int32_t taa_resolve_kernel(const taa_uniforms_t* uniforms, const taa_frame_images_t* images, const taa_cas_consts_t* cas,
                           int32_t arith, int32_t kernels[2]) {
    if (!kernels) return fail(TAA_ERR_INVALID_ARG, "kernels is NULL");
    taastar::FrameDesc fd;
    taastar::SegDesc sgd;
    int rc = prepare_resolve(uniforms, images, cas, arith, nullptr, fd, sgd);
    if (rc) return rc;
    const char* why[2] = {nullptr, nullptr};
    const bool split = uniforms->splitScreen != 0;
    kernels[0] = kernel_for(fd, *uniforms, 0, arith, &why[0]);
    kernels[1] = split ? kernel_for(fd, *uniforms, 1, arith, &why[1]) : kernels[0];
    if (!split) why[1] = nullptr;
    g_err.clear();
    if (why[0] || why[1]) {
        char buf[384];
        if (why[0] && why[1]) std::snprintf(buf, sizeof(buf), "FP16 request resolved by the FP32 path: %s (param[0]); %s (param[1])", why[0], why[1]);
        else std::snprintf(buf, sizeof(buf), "FP16 request resolved by the FP32 path: %s (param[%d])", why[0] ? why[0] : why[1], why[0] ? 0 : 1);
        g_err = buf;
    }
    return ok();
}

int32_t taa_cas_apply(const void* src, void* dst, int32_t width, int32_t height, int32_t pitch, int32_t format,
                      const taa_cas_consts_t* cas, void* stream) {
    if (!src || !dst || !cas) return fail(TAA_ERR_INVALID_ARG, "NULL argument");
    if (src == dst) return fail(TAA_ERR_INVALID_ARG, "src and dst alias");
    if (width <= 0 || height <= 0) return fail(TAA_ERR_INVALID_ARG, "bad size");
    if (!valid_color_fmt(format)) return fail(TAA_ERR_INVALID_ARG, "bad format");
    if (pitch == 0) pitch = width;
    if (pitch < width) return fail(TAA_ERR_INVALID_ARG, "pitch < width");
    if (!fits_offsets(pitch, height, 16)) return fail(TAA_ERR_INVALID_ARG, "image exceeds 4 GiB");
    int e = taastar::launch_cas(src, dst, width, height, pitch, format, cas->const1[0], stream);
    if (e) return hip_fail((hipError_t)e, "cas launch");
    return ok();
}

// the prepare + FXAA passes of taa.hpp:1061-1107, with the library's own edge filter
int32_t taa_edge_aa_apply(const void* src, const uint32_t* segmask, void* dst, int32_t width, int32_t height,
                          int32_t pitch, int32_t seg_pitch, int32_t format, const taa_edge_aa_consts_t* k, void* stream) {
    if (!src || !segmask || !dst || !k) return fail(TAA_ERR_INVALID_ARG, "NULL argument");
    if (src == dst) return fail(TAA_ERR_INVALID_ARG, "src and dst alias");
    if (static_cast<const void*>(segmask) == dst) return fail(TAA_ERR_INVALID_ARG, "segmask and dst alias");
    if (width <= 0 || height <= 0) return fail(TAA_ERR_INVALID_ARG, "bad size %dx%d", width, height);
    if (!valid_color_fmt(format)) return fail(TAA_ERR_INVALID_ARG, "bad format %d", format);
    if (pitch == 0) pitch = width;
    if (seg_pitch == 0) seg_pitch = pitch;
    if (pitch < width || seg_pitch < width) return fail(TAA_ERR_INVALID_ARG, "pitch < width");
    if (!fits_offsets(pitch, height, 16) || !fits_offsets(seg_pitch, height, 4))
        return fail(TAA_ERR_INVALID_ARG, "image exceeds 4 GiB");
    int e = taastar::launch_edge_aa(src, segmask, dst, width, height, pitch, seg_pitch, format, k->subpix, k->edge_threshold,
                                    k->edge_threshold_min, stream);
    if (e) return hip_fail((hipError_t)e, "edge-AA launch");
    return ok();
}

int32_t taa_edge_aa_defaults(int32_t width, int32_t height, taa_edge_aa_consts_t* out) {
    if (!out) return fail(TAA_ERR_INVALID_ARG, "out is NULL");
    if (width <= 0 || height <= 0) return fail(TAA_ERR_INVALID_ARG, "bad size %dx%d", width, height);
    // source/taa.hpp:94-97, 953
    out->rcp_frame[0] = 1.0f / (float)width;
    out->rcp_frame[1] = 1.0f / (float)height;
    out->subpix = 0.75f;
    out->edge_threshold = 0.116f;
    out->edge_threshold_min = 0.0833f;
    return ok();
}

// fwd_geometry.frag:288-295 from depth (include/taastar/taa_velocity.h)
int32_t taa_velocity_from_depth(const taa_reproject_t* r, const float* depth, const uint32_t* object_id,
                                void* velocity, int32_t velocity_format, int32_t width, int32_t height, int32_t pitch,
                                void* stream) {
    if (!r || !depth || !velocity) return fail(TAA_ERR_INVALID_ARG, "NULL argument");
    if (width <= 0 || height <= 0) return fail(TAA_ERR_INVALID_ARG, "bad size %dx%d", width, height);
    if (!valid_vel_fmt(velocity_format)) return fail(TAA_ERR_INVALID_ARG, "bad velocity_format %d", velocity_format);
    if (pitch == 0) pitch = width;
    if (pitch < width) return fail(TAA_ERR_INVALID_ARG, "pitch < width");
    if (!fits_offsets(pitch, height, 16)) return fail(TAA_ERR_INVALID_ARG, "image exceeds 4 GiB");
    if (r->reserved[0] || r->reserved[1] || r->reserved[2]) return fail(TAA_ERR_INVALID_ARG, "reserved fields must be 0");
    if (r->n_objects < 0) return fail(TAA_ERR_INVALID_ARG, "n_objects < 0");
    if (r->n_objects > 0 && !r->object_prev_from_current)
        return fail(TAA_ERR_INVALID_ARG, "n_objects > 0 needs object_prev_from_current");
    if (r->n_objects > (1 << 20)) return fail(TAA_ERR_INVALID_ARG, "n_objects above 2^20");
    if (reinterpret_cast<uintptr_t>(r->object_prev_from_current) % 16u)
        return fail(TAA_ERR_INVALID_ARG, "object_prev_from_current must be 16-byte aligned");
    if (static_cast<const void*>(depth) == velocity || static_cast<const void*>(object_id) == velocity)
        return fail(TAA_ERR_INVALID_ARG, "velocity aliases an input");
    int e = taastar::launch_velocity(*r, depth, object_id, velocity, velocity_format, width, height, pitch, stream);
    if (e) return hip_fail((hipError_t)e, "velocity launch");
    return ok();
}

int32_t taa_ctx_create(int32_t device, uint64_t hbm_budget_bytes, taa_ctx_t** out) {
    if (!out) return fail(TAA_ERR_INVALID_ARG, "out is NULL");
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return fail(TAA_ERR_NO_DEVICE, "no HIP device");
    if (device < 0 || device >= n) return fail(TAA_ERR_INVALID_ARG, "device %d of %d", device, n);
    hipDeviceProp_t prop;
    hipError_t e = hipGetDeviceProperties(&prop, device);
    if (e != hipSuccess) return hip_fail(e, "hipGetDeviceProperties");
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return fail(TAA_ERR_NO_DEVICE, "device %d is %s; this build targets gfx950 only", device, prop.gcnArchName);
    if ((e = hipSetDevice(device)) != hipSuccess) return hip_fail(e, "hipSetDevice");
    taa_ctx* c = new (std::nothrow) taa_ctx();
    if (!c) return fail(TAA_ERR_OUT_OF_MEMORY, "host allocation");
    c->device = device;
    if (hbm_budget_bytes == 0) {
        size_t fr = 0, tot = 0;
        if ((e = hipMemGetInfo(&fr, &tot)) != hipSuccess) { delete c; return hip_fail(e, "hipMemGetInfo"); }
        hbm_budget_bytes = (uint64_t)(0.9 * (double)fr);
    }
    c->budget = hbm_budget_bytes;
    for (auto& s : c->slots) {
        if (hipHostMalloc(&s.host, staging_bytes(), hipHostMallocDefault) != hipSuccess ||
            hipMalloc(&s.dev, staging_bytes()) != hipSuccess || hipEventCreateWithFlags(&s.done, hipEventDisableTiming) != hipSuccess) {
            taa_ctx_destroy(c);
            return fail(TAA_ERR_OUT_OF_MEMORY, "batch staging allocation");
        }
    }
    *out = c;
    return ok();
}

int32_t taa_ctx_destroy(taa_ctx_t* c) {
    if (!c) return fail(TAA_ERR_INVALID_ARG, "ctx is NULL");
    hipSetDevice(c->device);
    for (auto& s : c->slots) {
        if (s.done) { hipEventSynchronize(s.done); hipEventDestroy(s.done); }
        if (s.dev) hipFree(s.dev);
        if (s.host) hipHostFree(s.host);
    }
    if (c->tune_stream) hipStreamDestroy(c->tune_stream);
    delete c;
    return ok();
}

int32_t taa_ctx_usage(const taa_ctx_t* c, uint64_t* used, uint64_t* budget) {
    if (!c) return fail(TAA_ERR_INVALID_ARG, "ctx is NULL");
    if (used) *used = c->used.load();
    if (budget) *budget = c->budget;
    return ok();
}

uint64_t taa_session_bytes_scaled(int32_t w, int32_t h, int32_t wl, int32_t hl, int32_t cf, int32_t vf) {
    if (w <= 0 || h <= 0 || wl <= 0 || hl <= 0 || !valid_color_fmt(cf) || !valid_vel_fmt(vf)) return 0;
    const size_t px = (size_t)w * (size_t)h, pl = (size_t)wl * (size_t)hl;
    return 2 * align256(px * color_bpp(cf)) + align256(pl * color_bpp(cf)) + 2 * align256(pl * 4) +
           align256(pl * vel_bpp(vf));
}

uint64_t taa_session_bytes(int32_t w, int32_t h, int32_t cf, int32_t vf) { return taa_session_bytes_scaled(w, h, w, h, cf, vf); }

int32_t taa_session_open(taa_ctx_t* c, int32_t w, int32_t h, int32_t cf, int32_t vf, taa_session_t** out) {
    return taa_session_open_scaled(c, w, h, w, h, cf, vf, out);
}

int32_t taa_session_open_scaled(taa_ctx_t* c, int32_t w, int32_t h, int32_t wl, int32_t hl, int32_t cf, int32_t vf,
                                taa_session_t** out) {
    if (!c || !out) return fail(TAA_ERR_INVALID_ARG, "NULL argument");
    *out = nullptr;
    if (w <= 0 || h <= 0 || w > 32768 || h > 32768) return fail(TAA_ERR_INVALID_ARG, "bad size %dx%d", w, h);
    if (wl <= 0 || hl <= 0 || wl > 32768 || hl > 32768) return fail(TAA_ERR_INVALID_ARG, "bad input size %dx%d", wl, hl);
    if (!valid_color_fmt(cf) || !valid_vel_fmt(vf)) return fail(TAA_ERR_INVALID_ARG, "bad format");
    if (!fits_offsets(w, h, 16) || !fits_offsets(wl, hl, 16))
        return fail(TAA_ERR_INVALID_ARG, "session of %dx%d: an image would exceed 4 GiB", w, h);
    const uint64_t bytes = taa_session_bytes_scaled(w, h, wl, hl, cf, vf);
    // admission: reserve the bytes against the budget first
    uint64_t cur = c->used.load();
    do {
        if (cur + bytes > c->budget)
            return fail(TAA_ERR_BUDGET, "HBM budget: %llu + %llu > %llu bytes", (unsigned long long)cur,
                        (unsigned long long)bytes, (unsigned long long)c->budget);
    } while (!c->used.compare_exchange_weak(cur, cur + bytes));
    taa_session* s = new (std::nothrow) taa_session();
    if (!s) { c->used -= bytes; return fail(TAA_ERR_OUT_OF_MEMORY, "host allocation"); }
    hipSetDevice(c->device);
    hipError_t e = hipMalloc(&s->base, bytes);
    if (e != hipSuccess) {
        c->used -= bytes;
        delete s;
        return fail(TAA_ERR_OUT_OF_MEMORY, "hipMalloc(%llu): %s", (unsigned long long)bytes, hipGetErrorString(e));
    }
    // zero everything once so the first (reset) frame never blends non-finite garbage
    if ((e = hipMemset(s->base, 0, bytes)) != hipSuccess) {
        hipFree(s->base);
        c->used -= bytes;
        delete s;
        return hip_fail(e, "hipMemset");
    }
    const size_t px = (size_t)w * (size_t)h, pl = (size_t)wl * (size_t)hl;
    char* p = static_cast<char*>(s->base);
    s->hist[0] = p; p += align256(px * color_bpp(cf));
    s->hist[1] = p; p += align256(px * color_bpp(cf));
    s->color = p; p += align256(pl * color_bpp(cf));
    s->depth[0] = reinterpret_cast<float*>(p); p += align256(pl * 4);
    s->depth[1] = reinterpret_cast<float*>(p); p += align256(pl * 4);
    s->vel = p;
    s->ctx = c;
    s->W = w; s->H = h; s->pitch = w; s->cf = cf; s->vf = vf;
    s->Wl = wl; s->Hl = hl;
    s->bytes = bytes;
    *out = s;
    return ok();
}

int32_t taa_session_close(taa_session_t* s) {
    if (!s) return fail(TAA_ERR_INVALID_ARG, "session is NULL");
    hipSetDevice(s->ctx->device);
    hipDeviceSynchronize();  // the session's buffers may still be read by queued work
    hipFree(s->base);
    s->ctx->used -= s->bytes;
    delete s;
    return ok();
}

int32_t taa_session_io(taa_session_t* s, taa_session_io_t* out) {
    if (!s || !out) return fail(TAA_ERR_INVALID_ARG, "NULL argument");
    const int k = (int)(s->frames & 1);
    out->color = s->color;
    out->depth = s->depth[k];
    out->velocity = s->vel;
    out->history = s->hist[k ^ 1];
    out->width = s->Wl; out->height = s->Hl; out->pitch = s->Wl;  // the input slots
    out->color_format = s->cf; out->velocity_format = s->vf;
    out->frames_resolved = s->frames;
    return ok();
}

int32_t taa_session_reset(taa_session_t* s) {
    if (!s) return fail(TAA_ERR_INVALID_ARG, "session is NULL");
    s->force_reset = true;
    return ok();
}

int32_t taa_session_resolve(taa_session_t* s, const taa_uniforms_t* u, const taa_cas_consts_t* cas, void* screen,
                            int32_t arith, void* stream) {
    if (!s || !u || !screen) return fail(TAA_ERR_INVALID_ARG, "NULL argument");
    taa_uniforms_t U = *u;
    if (s->force_reset) U.mResetHistory = 1;
    const int k = (int)(s->frames & 1);
    taa_frame_images_t img;
    std::memset(&img, 0, sizeof(img));
    img.color = s->color;
    img.depth = s->depth[k];
    img.history = s->hist[k ^ 1];
    img.history_depth = s->depth[k ^ 1];
    img.result_screen = screen;
    img.velocity = s->vel;
    img.result_history = s->hist[k];
    img.width = s->W; img.height = s->H; img.pitch = s->pitch;
    img.input_width = s->Wl; img.input_height = s->Hl; img.input_pitch = s->Wl;
    img.color_format = s->cf; img.velocity_format = s->vf;
    int rc = resolve_impl(&U, &img, cas, arith, stream);
    if (rc) return rc;
    session_advance(s);
    return ok();
}

// ---- batches ------------------------------------------------------------------------------

// The items' buffers: no NULL field, sessions of this context and distinct, screens distinct and
// outside every session block of the call (a screen there would be written while that session's
// images are read).
static int check_item_buffers(taa_ctx_t* c, int n, const taa_batch_item_t* items) {
    for (int i = 0; i < n; ++i)
        if (!items[i].session || !items[i].uniforms || !items[i].result_screen)
            return fail(TAA_ERR_INVALID_ARG, "item %d: NULL field", i);
    for (int i = 0; i < n; ++i) {
        const taa_batch_item_t& it = items[i];
        if (it.session->ctx != c) return fail(TAA_ERR_INVALID_ARG, "item %d: session of another context", i);
        for (int j = 0; j < i; ++j) {
            if (items[j].session == it.session) return fail(TAA_ERR_INVALID_ARG, "item %d: session repeated", i);
            if (items[j].result_screen == it.result_screen)
                return fail(TAA_ERR_INVALID_ARG, "items %d and %d: result_screen alias", j, i);
        }
        const uintptr_t p = reinterpret_cast<uintptr_t>(it.result_screen);
        for (int j = 0; j < n; ++j) {
            const uintptr_t b = reinterpret_cast<uintptr_t>(items[j].session->base);
            if (p >= b && p < b + items[j].session->bytes)
                return fail(TAA_ERR_INVALID_ARG, "item %d: result_screen lies inside the buffers of item %d's session", i, j);
        }
    }
    return TAA_OK;
}

// What the items of one batched launch must satisfy whatever their uniforms hold: checked by
// every submit and once when a plan is created.
static int check_items(taa_ctx_t* c, int n, const taa_batch_item_t* items) {
    if (n <= 0 || n > taa_ctx::kCap) return fail(TAA_ERR_CAPACITY, "batch of %d items (capacity %d)", n, taa_ctx::kCap);
    if (int rc = check_item_buffers(c, n, items)) return rc;
    const bool cas = items[0].cas != nullptr;
    const bool s16 = items[0].session->cf == TAA_FMT_RGBA16F;
    for (int i = 0; i < n; ++i) {
        if ((items[i].cas != nullptr) != cas) return fail(TAA_ERR_INVALID_ARG, "item %d: CAS on/off must match the batch", i);
        if ((items[i].session->cf == TAA_FMT_RGBA16F) != s16)
            return fail(TAA_ERR_INVALID_ARG, "item %d: colour format must match", i);
    }
    return TAA_OK;
}

// The descriptor table of one batched call: one-size frames first (entries [0, n_main), the
// main batched kernel), then the upsampling frames (entries [n_main, n_main + n_up), the
// batched upsampling kernel); each part numbers its tiles from 0.
struct BatchLayout {
    int shape = 0;  // tile shape of the main launch
    bool cas = false, s16 = false, fast = false;
    int form = 0;  // the FP16 kernel's form (taastar::fast_path_form): the largest over the frames
    bool prod_main = false, prod_up = false;  // every frame of that launch has the production feature set
    int n_main = 0, total_main = 0;
    int n_up = 0, total_up = 0;
};

static int build_batch(taa_ctx_t* c, int n, const taa_batch_item_t* items, int arith, void* host, BatchLayout& L) {
    if (int rc = check_items(c, n, items)) return rc;
    auto* ents = static_cast<taastar::BatchEntry*>(host);
    L = BatchLayout{};
    L.cas = items[0].cas != nullptr;
    L.s16 = items[0].session->cf == TAA_FMT_RGBA16F;
    // one tile shape per main launch: a single frame's tuned shape; for several frames, the shape
    // of one frame of their total size (selected_shape_for_pixels)
    long long total_px = 0;
    for (int i = 0; i < n; ++i) {
        const taa_batch_item_t& it = items[i];
        if (int rc = taastar::check_batch_uniforms(it.uniforms)) return rc;
        if (session_upsamples(it.session, it.uniforms)) {
            ++L.n_up;
            continue;
        }
        ++L.n_main;
        total_px += (long long)it.session->W * it.session->H;
        if (L.n_main == 1) L.shape = selected_shape(it.session->W, it.session->H, it.session->cf, arith, L.cas);
    }
    if (L.n_main > 1) {
        L.shape = selected_shape_for_pixels(total_px, items[0].session->cf, arith, L.cas);
        // a shape tuned on single frames can resolve more texels over a batch's frames (edge tiles,
        // CAS ring): the default shape wins when its resolved texels are 2% fewer (with 62x14 tuned
        // at 4K, where it ties 30x30: 64 x 640x360 57.2 against 61.7 Gpix/s, the 64-session mix
        // 68.2 against 71.3)
        const int def = L.cas ? taastar::kDefaultShapeCas : taastar::kDefaultShapePlain;
        if (L.shape != def) {
            auto work = [&](int shape) {
                const taastar::TileShapeDef t = shape_def(shape, L.cas);
                const long long per_tile = (long long)(t.tw + (L.cas ? 2 : 0)) * (t.th + (L.cas ? 2 : 0));
                long long w = 0;
                for (int i = 0; i < n; ++i)
                    if (!session_upsamples(items[i].session, items[i].uniforms))
                        w += (long long)taastar::tiles_for(items[i].session->W, t.tw) *
                             taastar::tiles_for(items[i].session->H, t.th) * per_tile;
                return w;
            };
            if (work(def) * 100 < work(L.shape) * 98) L.shape = def;
        }
    }
    const taastar::TileShapeDef sd = shape_def(L.shape, L.cas);
    const taastar::TileShapeDef su = shape_def(L.cas ? taastar::kUpShapeCas : taastar::kUpShapePlain, L.cas);
    int k_main = 0, k_up = L.n_main;
    for (int i = 0; i < n; ++i) {
        taa_session* s = items[i].session;
        const bool up = session_upsamples(s, items[i].uniforms);
        taastar::BatchEntry& e = ents[up ? k_up++ : k_main++];
        std::memset(&e, 0, sizeof(e));
        session_desc(s, items[i].result_screen, items[i].cas, e.fd);
        const taastar::TileShapeDef& t = up ? su : sd;
        if (up) e.fd.up = 1;  // also a one-size frame with mUpsampling
        e.fd.tiles_x = taastar::tiles_for(s->W, t.tw);
        e.fd.tiles_y = taastar::tiles_for(s->H, t.th);
        e.U = *items[i].uniforms;
        if (s->force_reset) e.U.mResetHistory = 1;
        int& total = up ? L.total_up : L.total_main;
        e.tile_begin = total;
        total += e.fd.tiles_x * e.fd.tiles_y;
    }
    L.fast = arith == TAA_ARITH_FP16 && L.n_main > 0;
    for (int i = 0; i < L.n_main && L.fast; ++i) L.fast = taastar::fast_path_supported(ents[i].fd, ents[i].U, 0);
    // one launch, one staged velocity layout: RGBA-velocity frames (form 3) batch on the FP16 kernel
    // only with each other
    int n_rgba = 0;
    for (int i = 0; i < L.n_main && L.fast; ++i) {
        const int f = taastar::fast_path_form(ents[i].fd, ents[i].U, 0);
        n_rgba += f == 3;
        L.form = std::max(L.form, f);
    }
    if (L.fast && n_rgba != 0 && n_rgba != L.n_main) L.fast = false;
    L.prod_main = !L.fast && L.n_main > 0;
    for (int i = 0; i < L.n_main && L.prod_main; ++i) L.prod_main = taastar::exact_prod_form(ents[i].fd, ents[i].U);
    L.prod_up = L.n_up > 0;
    for (int i = L.n_main; i < L.n_main + L.n_up && L.prod_up; ++i) L.prod_up = taastar::prod_params(ents[i].fd, ents[i].U);
    if (L.prod_up && L.cas) {  // the production form's own CAS tile, as in taa_resolve_frame
        const taastar::TileShapeDef sp = shape_def(taastar::kUpShapeCasProd, true);
        L.total_up = 0;
        for (int i = L.n_main; i < L.n_main + L.n_up; ++i) {
            taastar::BatchEntry& e = ents[i];
            e.fd.tiles_x = taastar::tiles_for(e.fd.W, sp.tw);
            e.fd.tiles_y = taastar::tiles_for(e.fd.H, sp.th);
            e.tile_begin = L.total_up;
            L.total_up += e.fd.tiles_x * e.fd.tiles_y;
        }
    }
    return TAA_OK;
}

// The launches of one batched call on `stream` (descriptors already on the device at d_ents).
static int launch_layout(const taastar::BatchEntry* d_ents, const BatchLayout& L, int arith, void* stream) {
    if (L.n_main) {
        int le = L.fast ? taastar::launch_batch_fast(d_ents, L.n_main, L.total_main, L.shape, L.cas, L.s16, L.form, stream)
                        : taastar::launch_batch(d_ents, L.n_main, L.total_main, L.shape, L.cas, L.s16, arith, L.prod_main,
                                                stream);
        if (le) return le;
    }
    if (L.n_up)
        return taastar::launch_batch_up(d_ents + L.n_main, L.n_up, L.total_up, L.cas, L.s16, arith, L.prod_up, stream);
    return 0;
}

int32_t taa_batch_submit(taa_ctx_t* c, int32_t n, const taa_batch_item_t* items, int32_t arith, void* stream) {
    if (!c || !items) return fail(TAA_ERR_INVALID_ARG, "NULL argument");
    if (arith != TAA_ARITH_FP32 && arith != TAA_ARITH_FP16) return fail(TAA_ERR_INVALID_ARG, "bad arith");
    std::lock_guard<std::mutex> lk(c->mu);
    taa_ctx::Slot& slot = c->slots[c->next_slot];
    if (slot.used) {
        hipError_t e = hipEventSynchronize(slot.done);  // the slot's previous batch has been consumed
        if (e != hipSuccess) return hip_fail(e, "staging slot wait");
        slot.used = false;
    }
    BatchLayout L;
    int rc = build_batch(c, n, items, arith, slot.host, L);
    if (rc) return rc;
    hipStream_t st = static_cast<hipStream_t>(stream);
    hipError_t e = hipMemcpyAsync(slot.dev, slot.host, sizeof(taastar::BatchEntry) * (size_t)n, hipMemcpyHostToDevice, st);
    if (e != hipSuccess) return hip_fail(e, "descriptor upload");
    // the copy from the pinned slot is enqueued: the slot stays reserved whatever follows
    const int le = launch_layout(static_cast<const taastar::BatchEntry*>(slot.dev), L, arith, stream);
    staging_enqueued(slot.done, slot.used, st);
    c->next_slot = (c->next_slot + 1) % taa_ctx::kSlots;
    if (le) return hip_fail((hipError_t)le, "batch launch");
    for (int i = 0; i < n; ++i) session_advance(items[i].session);
    return ok();
}


}  // extern "C"

// ---- graph-captured batch plans -----------------------------------------------------------

struct taa_batch_plan {
    taa_ctx* ctx = nullptr;
    int n = 0, arith = 0;
    std::vector<taa_batch_item_t> items;
    // two descriptor buffers alternate between launches; each holds one captured graph per kernel
    // family of the one-size frames (FP32 kernel, FP16 kernel), captured the first time that family
    // is needed and again when the launch layout (tile shape, tile counts, split between one-size
    // and upsampling frames) changes
    struct Buf {
        void* host = nullptr;
        void* dev = nullptr;
        hipGraphExec_t exec[2] = {nullptr, nullptr};
        BatchLayout layout[2];
        hipEvent_t done = nullptr;
        bool used = false;
    } buf[2];
    int next = 0;
    hipStream_t cap = nullptr;
};

namespace {

void plan_free(taa_batch_plan* p) {
    for (auto& b : p->buf) {
        if (b.done) { hipEventSynchronize(b.done); hipEventDestroy(b.done); }
        for (auto& e : b.exec)
            if (e) hipGraphExecDestroy(e);
        if (b.dev) hipFree(b.dev);
        if (b.host) hipHostFree(b.host);
    }
    if (p->cap) hipStreamDestroy(p->cap);
    delete p;
}

bool same_launches(const BatchLayout& a, const BatchLayout& b) {
    return a.shape == b.shape && a.fast == b.fast && a.form == b.form && a.prod_main == b.prod_main &&
           a.prod_up == b.prod_up && a.n_main == b.n_main &&
           a.total_main == b.total_main &&
           a.n_up == b.n_up && a.total_up == b.total_up;
}

// Capture descriptor upload + the batched launches of a layout for one buffer.
int plan_capture(taa_batch_plan* p, taa_batch_plan::Buf& b, const BatchLayout& L) {
    const int fam = L.fast ? 1 : 0;
    hipGraph_t g = nullptr;
    hipError_t e;
    if ((e = hipStreamBeginCapture(p->cap, hipStreamCaptureModeThreadLocal)) != hipSuccess) return hip_fail(e, "begin capture");
    hipMemcpyAsync(b.dev, b.host, sizeof(taastar::BatchEntry) * (size_t)p->n, hipMemcpyHostToDevice, p->cap);
    const int le = launch_layout(static_cast<const taastar::BatchEntry*>(b.dev), L, p->arith, p->cap);
    if ((e = hipStreamEndCapture(p->cap, &g)) != hipSuccess) return hip_fail(e, "end capture");
    if (le) {
        hipGraphDestroy(g);
        return hip_fail((hipError_t)le, "batch launch (capture)");
    }
    e = hipGraphInstantiate(&b.exec[fam], g, nullptr, nullptr, 0);
    hipGraphDestroy(g);
    if (e != hipSuccess) return hip_fail(e, "graph instantiate");
    b.layout[fam] = L;
    return TAA_OK;
}

}  // namespace

extern "C" {

int32_t taa_batch_plan_create(taa_ctx_t* c, int32_t n, const taa_batch_item_t* items, int32_t arith, taa_batch_plan_t** out) {
    if (!c || !items || !out) return fail(TAA_ERR_INVALID_ARG, "NULL argument");
    *out = nullptr;
    if (arith != TAA_ARITH_FP32 && arith != TAA_ARITH_FP16) return fail(TAA_ERR_INVALID_ARG, "bad arith");
    if (int rc = check_items(c, n, items)) return rc;
    taa_batch_plan* p = new (std::nothrow) taa_batch_plan();
    if (!p) return fail(TAA_ERR_OUT_OF_MEMORY, "host allocation");
    p->ctx = c;
    p->n = n;
    p->arith = arith;
    p->items.assign(items, items + n);
    hipSetDevice(c->device);
    hipError_t e;
    if ((e = hipStreamCreateWithFlags(&p->cap, hipStreamNonBlocking)) != hipSuccess) { plan_free(p); return hip_fail(e, "stream"); }
    for (auto& b : p->buf) {
        if ((e = hipHostMalloc(&b.host, staging_bytes(), hipHostMallocDefault)) != hipSuccess ||
            (e = hipMalloc(&b.dev, staging_bytes())) != hipSuccess ||
            (e = hipEventCreateWithFlags(&b.done, hipEventDisableTiming)) != hipSuccess) {
            plan_free(p);
            return fail(TAA_ERR_OUT_OF_MEMORY, "plan buffers: %s", hipGetErrorString(e));
        }
    }
    *out = p;
    return ok();
}

int32_t taa_batch_plan_launch(taa_batch_plan_t* p, void* stream) {
    if (!p) return fail(TAA_ERR_INVALID_ARG, "plan is NULL");
    auto& b = p->buf[p->next];
    if (b.used) {
        hipError_t e = hipEventSynchronize(b.done);  // this buffer's previous replay has finished
        if (e != hipSuccess) return hip_fail(e, "plan buffer wait");
        b.used = false;
    }
    BatchLayout L;
    int rc = build_batch(p->ctx, p->n, p->items.data(), p->arith, b.host, L);
    if (rc) return rc;
    const int fam = L.fast ? 1 : 0;
    if (b.exec[fam] && !same_launches(b.layout[fam], L)) {  // launch layout changed: recapture
        hipGraphExecDestroy(b.exec[fam]);
        b.exec[fam] = nullptr;
    }
    if (!b.exec[fam]) {
        rc = plan_capture(p, b, L);
        if (rc) return rc;
    }
    hipStream_t st = static_cast<hipStream_t>(stream);
    hipError_t e = hipGraphLaunch(b.exec[fam], st);
    if (e != hipSuccess) return hip_fail(e, "graph launch");
    staging_enqueued(b.done, b.used, st);  // the graph copies from b.host
    p->next ^= 1;
    for (int i = 0; i < p->n; ++i) session_advance(p->items[i].session);
    return ok();
}

int32_t taa_batch_plan_destroy(taa_batch_plan_t* p) {
    if (!p) return fail(TAA_ERR_INVALID_ARG, "plan is NULL");
    plan_free(p);
    return ok();
}

// ---- autotuner ----------------------------------------------------------------------------

int32_t taa_tile_shape_count(void) { return taastar::kNumShapes; }

int32_t taa_tile_shape_get(int32_t index, taa_bool32 cas, taa_tile_shape_t* out) {
    if (!out || index < 0 || index >= taastar::kNumShapes) return fail(TAA_ERR_INVALID_ARG, "bad index");
    taastar::TileShapeDef d = shape_def(index, cas != 0);
    out->tile_w = d.tw;
    out->tile_h = d.th;
    out->lds_pad = d.pad;
    return ok();
}

int32_t taa_tile_shape_select(int32_t w, int32_t h, int32_t fmt, int32_t arith, taa_bool32 cas, int32_t index) {
    if (index < -1 || index >= taastar::kNumShapes) return fail(TAA_ERR_INVALID_ARG, "bad index");
    std::lock_guard<std::mutex> lk(g_tune_mu);
    ClassKey k{w, h, fmt, arith, cas ? 1 : 0};
    if (index < 0) g_tune.erase(k);
    else g_tune[k] = index;
    return ok();
}

int32_t taa_tile_shape_selected(int32_t w, int32_t h, int32_t fmt, int32_t arith, taa_bool32 cas) {
    return selected_shape(w, h, fmt, arith, cas != 0);
}

int32_t taa_autotune(taa_ctx_t* c, int32_t w, int32_t h, int32_t fmt, int32_t arith, taa_bool32 cas, int32_t iters,
                     int32_t* best_index, float* best_us) {
    if (!c) return fail(TAA_ERR_INVALID_ARG, "ctx is NULL");
    if (w <= 0 || h <= 0 || !valid_color_fmt(fmt)) return fail(TAA_ERR_INVALID_ARG, "bad class");
    if (iters <= 0) iters = 20;
    hipSetDevice(c->device);
    if (!c->tune_stream && hipStreamCreateWithFlags(&c->tune_stream, hipStreamNonBlocking) != hipSuccess)
        return fail(TAA_ERR_HIP, "stream");
    hipStream_t st = c->tune_stream;
    const size_t px = (size_t)w * h, cb = px * color_bpp(fmt);
    // synthetic frame: colour ramps and a depth step so every branch of the config runs
    std::vector<char> hc(cb), hv(px * 8);
    std::vector<float> hd(px);
    for (size_t i = 0; i < px; ++i) {
        const int x = (int)(i % w), y = (int)(i / w);
        float rgb[4] = {0.1f + 0.8f * (float)((x * 7 + y * 3) % 64) / 64.f, 0.3f + 0.01f * (float)(y % 17), 0.5f, 1.0f};
        if (fmt == TAA_FMT_RGBA16F) {
            for (int k = 0; k < 4; ++k) {  // fp16 of values in [0.1, 1]: exact enough for timing
                uint16_t hb = 0;
                float v = rgb[k];
                uint32_t u;
                std::memcpy(&u, &v, 4);
                hb = (uint16_t)(((u >> 16) & 0x8000u) | ((((u >> 23) & 0xff) - 112) << 10) | ((u >> 13) & 0x3ff));
                std::memcpy(&hc[i * 8 + 2 * k], &hb, 2);
            }
        } else {
            std::memcpy(&hc[i * 16], rgb, 16);
        }
        hd[i] = ((x / 37 + y / 23) & 1) ? 0.4f : 0.7f;
        float v2[2] = {(float)((x % 13) - 6) / (float)w, (float)((y % 11) - 5) / (float)h};
        if (fmt == TAA_FMT_RGBA16F) {  // RG16F: small normal values, truncated to fp16
            for (int k = 0; k < 2; ++k) {
                uint32_t u;
                std::memcpy(&u, &v2[k], 4);
                const int e = (int)((u >> 23) & 0xff) - 127;
                uint16_t hb = (e < -14) ? (uint16_t)((u >> 16) & 0x8000u)
                                        : (uint16_t)(((u >> 16) & 0x8000u) | ((uint32_t)(e + 15) << 10) | ((u >> 13) & 0x3ff));
                std::memcpy(&hv[i * 4 + 2 * k], &hb, 2);
            }
        } else {
            std::memcpy(&hv[i * 8], v2, 8);
        }
    }
    void *dc = nullptr, *dh = nullptr, *dho = nullptr, *ds = nullptr, *dv = nullptr;
    float *dd = nullptr, *dhd = nullptr;
    bool alloc_ok = hipMalloc(&dc, cb) == hipSuccess && hipMalloc(&dh, cb) == hipSuccess && hipMalloc(&dho, cb) == hipSuccess &&
                    hipMalloc(&ds, cb) == hipSuccess && hipMalloc(&dv, px * 8) == hipSuccess &&
                    hipMalloc((void**)&dd, px * 4) == hipSuccess && hipMalloc((void**)&dhd, px * 4) == hipSuccess;
    auto release = [&]() {
        hipStreamSynchronize(st);
        for (void* p : {dc, dh, dho, ds, dv, (void*)dd, (void*)dhd})
            if (p) hipFree(p);
    };
    if (!alloc_ok) { release(); return fail(TAA_ERR_OUT_OF_MEMORY, "autotune buffers"); }
    hipMemcpy(dc, hc.data(), cb, hipMemcpyHostToDevice);
    hipMemcpy(dh, hc.data(), cb, hipMemcpyHostToDevice);
    hipMemcpy(dv, hv.data(), px * 8, hipMemcpyHostToDevice);
    hipMemcpy(dd, hd.data(), px * 4, hipMemcpyHostToDevice);
    hipMemcpy(dhd, hd.data(), px * 4, hipMemcpyHostToDevice);
    // the full feature set (variance clip, Catmull-Rom, closest-depth velocity, rejections, Lottes)
    taa_uniforms_t U;
    std::memset(&U, 0, sizeof(U));
    taa_params_defaults(&U.param[0]);
    taa_params_t& P = U.param[0];
    P.mVarianceClipping = 1; P.mVarClipGamma = 1.0f; P.mColorClampingOrClipping = 2; P.mInterpolationMode = 2;
    P.mAlpha = 0.1f; P.mUseVelocityVectors = 2; P.mVelocitySampleMode = 2; P.mUseYCoCg = 1;
    P.mRejectOutside = 1; P.mDepthCulling = 1; P.mLumaWeightingLottes = 1;
    U.param[1] = P;
    U.mCamNearPlane = 0.1f; U.mCamFarPlane = 100.f; U.pad1 = 0.01f;
    taa_cas_consts_t K;
    taa_cas_setup(K.const0, K.const1, 0.5f, (float)w, (float)h, (float)w, (float)h);
    taastar::FrameDesc fd;
    taa_frame_images_t img;
    std::memset(&img, 0, sizeof(img));
    img.color = dc; img.depth = dd; img.history = dh; img.history_depth = dhd; img.result_screen = ds;
    img.velocity = dv; img.result_history = dho; img.width = w; img.height = h; img.pitch = w;
    img.color_format = fmt; img.velocity_format = fmt == TAA_FMT_RGBA16F ? TAA_VEL_RG16F : TAA_VEL_RG32F;
    int rc = make_desc(&img, cas ? &K : nullptr, fd);
    if (rc) { release(); return rc; }
    hipEvent_t e0, e1;
    hipEventCreate(&e0);
    hipEventCreate(&e1);
    int best = -1;
    float best_t = 1e30f;
    for (int s = 0; s < taastar::kNumShapes; ++s) {
        const taastar::TileShapeDef sd = shape_def(s, cas != 0);
        fd.tiles_x = taastar::tiles_for(w, sd.tw);
        fd.tiles_y = taastar::tiles_for(h, sd.th);
        const bool fast = arith == TAA_ARITH_FP16 && taastar::fast_path_supported(fd, U, 0);
        auto go = [&]() {
            if (fast) taastar::launch_resolve_fast(fd, U, s, st);
            else taastar::launch_resolve(fd, U, s, arith, st);
        };
        for (int i = 0; i < 3; ++i) go();
        hipEventRecord(e0, st);
        for (int i = 0; i < iters; ++i) go();
        hipEventRecord(e1, st);
        hipEventSynchronize(e1);
        float ms = 0.f;
        hipEventElapsedTime(&ms, e0, e1);
        const float us = 1000.f * ms / (float)iters;
        if (us < best_t) { best_t = us; best = s; }
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    hipError_t le = hipGetLastError();
    release();
    if (le != hipSuccess) return hip_fail(le, "autotune launch");
    taa_tile_shape_select(w, h, fmt, arith, cas, best);
    if (best_index) *best_index = best;
    if (best_us) *best_us = best_t;
    return ok();
}

}  // extern "C"

namespace taastar {
int set_last_error(int code, const char* msg) { return fail(code, "%s", msg); }

int resolve_frame_with_scratch(const taa_uniforms_t* uniforms, const taa_frame_images_t* images, const taa_cas_consts_t* cas,
                               int arith, void* stream, void* scratch, size_t scratch_bytes) {
    return resolve_impl(uniforms, images, cas, arith, stream, nullptr, scratch, scratch_bytes);
}

size_t split_cas_scratch_bytes(const taa_frame_images_t* images) {
    if (!images) return 0;
    const int pitch = images->pitch == 0 ? images->width : images->pitch;
    if (images->height <= 0 || pitch <= 0) return 0;
    return (size_t)pitch * (size_t)images->height * color_bpp(images->color_format);
}
}  // namespace taastar
