// Internal types shared by the gfx950 kernels (taa_kernels.hip) and the host runtime
// (taa_api.cpp). Not part of the C ABI.
#pragma once

#include <stddef.h>
#include <stdint.h>

#include "taastar/taa_abi.h"
#include "taastar/taa_velocity.h"

namespace taastar {

// One frame as the kernels see it. Mirrors the descriptor set of taa.comp
// (shaders/taa.comp:24-38) plus the tiling of this launch.
struct FrameDesc {
    const void* color;        // uCurrentFrame
    const float* depth;       // uCurrentDepth
    const void* history;      // uHistoryFrame
    const float* hdepth;      // uHistoryDepth
    void* screen;             // uResultScreen (sharpened when cas_on)
    void* debug;              // uDebug (RGBA16F) or null
    const void* velocity;     // uCurrentVelocity
    void* history_out;        // uResultHistory
    void* inspect;            // parity-test intermediates or null
    int32_t W, H, pitch;
    int32_t color_fmt, vel_fmt;
    int32_t x_begin, x_end;   // columns owned by this launch (split screen), [x_begin, x_end)
    int32_t tiles_x, tiles_y; // output tiles covering [x_begin, x_end) x [0, H)
    int32_t param_idx;        // which Parameters block of the uniforms applies
    int32_t cas_on;
    uint32_t cas_peak_bits;   // const1[0] of CasSetup
    int32_t Wl, Hl, pitch_l;  // input (render) size of colour, depth, velocity, previous depth
    int32_t up;               // upsampling path: input size != output size, or mUpsampling
};

// The segmentation mask's extra images (taa_segmentation_images_t), pitch resolved.
struct SegDesc {
    const uint32_t* material;
    const uint32_t* prev_material;
    const float* uv_normal;
    const uint32_t* prev_segmask;
    uint32_t* segmask;
    int32_t pitch;    // the masks' row pitch (output size)
    int32_t pitch_l;  // the renderer's images' row pitch (input size)
};

// One frame of a batched launch: frame + device copy of its uniforms + first flattened tile.
struct BatchEntry {
    FrameDesc fd;
    taa_uniforms_t U;         // the frame's uniforms, inline so the kernel reads them with scalar loads
    int32_t tile_begin;       // prefix sum of tiles of the entries before this one
    int32_t pad;
};

// Tile shapes the kernels are instantiated for (output tile, 256 threads per workgroup).
struct TileShapeDef {
    int tw, th, pad;
};
constexpr int kNumShapes = 4;
// Shapes used without fused CAS (resolve region == output tile)
constexpr TileShapeDef kShapesPlain[kNumShapes] = {{16, 16, 1}, {32, 8, 1}, {64, 4, 1}, {32, 16, 1}};
// Shapes used with fused CAS (resolve region == output tile + 1-px apron)
constexpr TileShapeDef kShapesCas[kNumShapes] = {{14, 14, 1}, {30, 14, 1}, {30, 30, 1}, {62, 14, 1}};
constexpr int kDefaultShapePlain = 0;
constexpr int kDefaultShapeCas = 2;
// tile shapes of the resolve with the segmentation mask: its LDS also holds the mask's inputs,
// so the small 14x14 CAS tile keeps more workgroups resident (1080p FP32 with the reference's
// default flags: 102 us against 105 us for 30x14 and 151 us for 30x30)
constexpr int kSegShapeCas = 0;
constexpr int kSegShapePlain = kDefaultShapePlain;
// tile shapes of the upsampling resolve (its LDS holds the staged input rectangle); its production
// form (fewer registers) takes 30x14 CAS tiles: 1080p -> 4K in the FP16 arithmetic 0.205 ms against
// 0.232 ms with 14x14 (30x30 and 62x14: 0.261 and 0.267 ms); the general form stays at 14x14
constexpr int kUpShapeCas = 0;
constexpr int kUpShapeCasProd = 1;
constexpr int kUpShapePlain = kDefaultShapePlain;

inline int tiles_for(int extent, int t) { return (extent + t - 1) / t; }

// Launchers (taa_kernels.hip). All return hipError_t as int.
int launch_resolve(const FrameDesc& fd, const taa_uniforms_t& U, int shape, int arith, void* stream);
// whether launch_resolve runs the exact kernel's production form for this frame (variance clip box,
// clip, YCoCg, Catmull-Rom history, closest-depth RG velocity of the colour's precision, depth
// culling, Lottes, outside rejection; no debug or inspect image): the general kernel's arithmetic
// with the other branches compiled out
bool exact_prod_form(const FrameDesc& fd, const taa_uniforms_t& U);
// the same switches, whatever the sizes: launch_resolve_up runs the upsampling kernel's production
// form when they hold; with seg, the same with mRayTraceAugment set (the masked resolve's
// production form, whose mask keeps its own parameters)
bool prod_params(const FrameDesc& fd, const taa_uniforms_t& U, bool seg = false);
// the upsampling (TAAU) path: exact FP32 arithmetic, or with TAA_ARITH_FP16 the FP16 path's where no
// selection depends on it; tiled by its form's shape (kUpShapeCas / kUpShapeCasProd / kUpShapePlain),
// whatever fd.tiles_x / tiles_y say
int launch_resolve_up(const FrameDesc& fd, const taa_uniforms_t& U, int arith, void* stream);
// the resolve with the segmentation mask (exact FP32 arithmetic, or with TAA_ARITH_FP16 the FP16 path's
// where no selection depends on it); fd.up: the upsampling path's kernel, which reads the mask's
// inputs from global memory. Tiled by its own shape, whatever fd.tiles_x / tiles_y say
int launch_resolve_seg(const FrameDesc& fd, const taa_uniforms_t& U, const SegDesc& sg, int arith, void* stream);
int launch_cas(const void* src, void* dst, int W, int H, int pitch, int fmt, uint32_t peak_bits, void* stream);
// Workgroups of a batched launch: one per tile, rounded to whole XCD slots; the batch kernels take
// exactly this grid and resolve one tile per workgroup (batch_tile). Measured on the 64-session
// mix, one workgroup per tile with the hardware dispatcher balancing the XCDs is faster (55.7 vs
// 48.1 Gpix/s) than an occupancy-sized persistent grid with static strides; and a kernel body
// without a loop over tiles keeps the single-frame kernel's register budget (the loop cost the
// FP16 batch kernel 12 VGPRs and a wave per SIMD: 62.6 against 70.9 Gpix/s on the 48-stream mix).
inline int batch_grid(int total_tiles) { return (total_tiles + 7) & ~7; }

// prod: every frame has the production feature set (exact_prod_form): the production form
int launch_batch(const BatchEntry* d_entries, int n_entries, int total_tiles, int shape, bool cas,
                 bool s16, int arith, bool prod, void* stream);
// the upsampling frames of a batch (the upsampling resolve in the batch's arithmetic, kUpShapeCas /
// kUpShapePlain tiles, kUpShapeCasProd with prod: every frame has the production feature set,
// prod_params)
int launch_batch_up(const BatchEntry* d_entries, int n_entries, int total_tiles, bool cas, bool s16, int arith, bool prod,
                    void* stream);

// Throughput kernel (taa_fast.hip): variance clip + Catmull-Rom + simple/closest-depth velocity
// with RG velocity, fp16-fed mixed arithmetic. fast_path_supported() says whether a parameter
// block is one it implements (else *why, when given, names the first reason it is not);
// everything else runs the FP32 kernel.
bool fast_path_supported(const FrameDesc& fd, const taa_uniforms_t& U, int param_idx, const char** why = nullptr);
// the form of the FP16 kernel a supported block runs: 0 the production form (variance box,
// Catmull-Rom history, RG velocity), 1 the general form (min/max box or bilinear history), 2 the
// extended form (matrix reprojection, 3x3-longest velocity, anti-ghosting, velocity-based alpha,
// near-clamp, debug image) with RG velocity, 3 the extended form with RGBA velocity (a batched
// launch's frames must all have RGBA velocity or none)
int fast_path_form(const FrameDesc& fd, const taa_uniforms_t& U, int param_idx);
int launch_resolve_fast(const FrameDesc& fd, const taa_uniforms_t& U, int shape, void* stream);
int launch_batch_fast(const BatchEntry* d_entries, int n_entries, int total_tiles, int shape, bool cas, bool s16,
                      int form, void* stream);

// Edge-AA fallback on category-1 texels of a segmentation mask (taa_edge_aa.hip).
int launch_edge_aa(const void* src, const uint32_t* mask, void* dst, int W, int H, int pitch, int seg_pitch, int fmt,
                   float subpix, float thr, float thr_min, void* stream);

// TAA_OK, or the status taa_batch_submit returns for a frame with these uniforms (taa_api.cpp).
int check_batch_uniforms(const taa_uniforms_t* u);

// Record `msg` as the calling thread's taa_last_error() and return `code` (taa_api.cpp).
int set_last_error(int code, const char* msg);

// taa_resolve_frame with the intermediate screen of split screen with CAS in `scratch` (a device
// buffer of scratch_bytes; TAA_ERR_CAPACITY when the frame needs more) instead of the library's
// per-device buffer; every other case is taa_resolve_frame itself (taa_api.cpp).
int resolve_frame_with_scratch(const taa_uniforms_t* uniforms, const taa_frame_images_t* images, const taa_cas_consts_t* cas,
                               int arith, void* stream, void* scratch, size_t scratch_bytes);
// Bytes that intermediate screen takes for these images (pitch x height texels of the colour format).
size_t split_cas_scratch_bytes(const taa_frame_images_t* images);

// Velocity producer (taa_velocity.hip): one launch over width x height; ids may be NULL.
int launch_velocity(const taa_reproject_t& r, const float* depth, const uint32_t* ids, void* velocity, int fmt, int W,
                    int H, int pitch, void* stream);

}  // namespace taastar
