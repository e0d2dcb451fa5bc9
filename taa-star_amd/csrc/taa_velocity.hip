// Velocity producer: the motion-vector image of the reference's geometry pass
// (shaders/fwd_geometry.frag:288-295), computed per texel from the depth image and the camera
// matrices instead of from interpolated vertex positions. FP32 without contraction (this file is
// compiled with -ffp-contract=off) in the operation order of the oracle's velocity_texel: IEEE
// division for the texel-centre uv and for the reciprocal of the previous clip w; half formats
// store with the colour path's rounding (nearest even, saturating).
//
// A workgroup covers 64x16 texels: each lane walks one column through four rows 4 apart, so a
// wave reads 64 consecutive depths and writes 64 consecutive velocities per row, and the column's
// uv and its x term of the inverse view-projection are formed once per lane. The two camera
// matrices and the jitter arrive as kernel arguments (scalar registers); per-object matrices are
// read from device memory for texels that carry an object id.
#include <hip/hip_runtime.h>

#include "taa_device.hpp"
#include "taa_internal.hpp"

namespace taastar {
namespace {

struct VelocityArgs {
    float Mi[16];   // inverse(P_jittered_current * V_current), column-major
    float Mp[16];   // P_jittered_previous * V_previous
    float jit[4];   // current jitter xy, previous jitter zw (NDC)
    const float* depth;
    const uint32_t* ids;
    const float* objects;
    void* velocity;
    int W, H, pitch, n_objects;
};

// M * v, row r = M[r] v0 + M[4 + r] v1 + M[8 + r] v2 + M[12 + r] v3, left to right
__device__ __forceinline__ void mat4_apply(const float* M, const float (&v)[4], float (&o)[4]) {
#pragma unroll
    for (int r = 0; r < 4; ++r) o[r] = M[r] * v[0] + M[4 + r] * v[1] + M[8 + r] * v[2] + M[12 + r] * v[3];
}

template <int FMT, bool IDS>
__global__ __launch_bounds__(256) void k_velocity(VelocityArgs a) {
    constexpr int kRows = 4;
    const int x = (int)blockIdx.x * 64 + (int)(threadIdx.x & 63u);
    const int y0 = (int)blockIdx.y * (4 * kRows) + (int)(threadIdx.x >> 6);
    // The wave's four rows y0, y0 + 4, y0 + 8, y0 + 12 share one v each: lane k (k < 4) divides
    // for row y0 + 4k and every lane reads the four results back with readlane, before any lane
    // leaves (one division instead of four per lane).
    const float v_lane = ((float)(y0 + 4 * (int)(threadIdx.x & 3u)) + 0.5f) / (float)a.H;
    float v_row[kRows];
#pragma unroll
    for (int k = 0; k < kRows; ++k)
        v_row[k] = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v_lane), k));
    if (x >= a.W) return;
    const float u = ((float)x + 0.5f) / (float)a.W;
    const float cx = u * 2.0f - 1.0f;
    float mx[4];  // first product of each row of inverse_view_proj * clip
#pragma unroll
    for (int r = 0; r < 4; ++r) mx[r] = a.Mi[r] * cx;

#pragma unroll
    for (int k = 0; k < kRows; ++k) {
        const int y = y0 + 4 * k;
        if (y >= a.H) break;
        const uint32_t i = (uint32_t)y * (uint32_t)a.pitch + (uint32_t)x;
        const float v = v_row[k];
        const float depth = ld_f32(a.depth, i);
        const uint32_t id = IDS ? gload<uint32_t>(a.ids, i * 4u) : 0u;
        const float cy = v * 2.0f - 1.0f;
        float world[4], prev[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) world[r] = mx[r] + a.Mi[4 + r] * cy + a.Mi[8 + r] * depth + a.Mi[12 + r] * 1.0f;
        if (IDS && id >= 1u && id <= (uint32_t)a.n_objects) {
            const uint32_t o = (id - 1u) * 64u;
            float M[16];
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                const float4 col = gload<float4>(a.objects, o + 16u * (uint32_t)c);
                M[4 * c] = col.x, M[4 * c + 1] = col.y, M[4 * c + 2] = col.z, M[4 * c + 3] = col.w;
            }
            float moved[4];
            mat4_apply(M, world, moved);
#pragma unroll
            for (int r = 0; r < 4; ++r) world[r] = moved[r];
        }
        mat4_apply(a.Mp, world, prev);

        const float inv = 1.0f / prev[3];
        const float nx = cx - a.jit[0], ny = cy - a.jit[1], nz = depth;
        const float px = prev[0] * inv - a.jit[2], py = prev[1] * inv - a.jit[3], pz = prev[2] * inv;
        const float vx = (nx - px) * 0.5f, vy = (ny - py) * 0.5f, vz = nz - pz, vw = (float)id;

        if constexpr (FMT == TAA_VEL_RG32F) gstore<float2>(a.velocity, i * 8u, make_float2(vx, vy));
        else if constexpr (FMT == TAA_VEL_RG16F) gstore<uint32_t>(a.velocity, i * 4u, pack_f2h(vx, vy));
        else if constexpr (FMT == TAA_VEL_RGBA32F) gstore<float4>(a.velocity, i * 16u, make_float4(vx, vy, vz, vw));
        else gstore<uint2>(a.velocity, i * 8u, make_uint2(pack_f2h(vx, vy), pack_f2h(vz, vw)));
    }
}

template <int FMT>
void launch_fmt(const VelocityArgs& a, bool ids, dim3 grid, hipStream_t st) {
    if (ids) hipLaunchKernelGGL((k_velocity<FMT, true>), grid, dim3(256), 0, st, a);
    else hipLaunchKernelGGL((k_velocity<FMT, false>), grid, dim3(256), 0, st, a);
}

}  // namespace

int launch_velocity(const taa_reproject_t& r, const float* depth, const uint32_t* ids, void* velocity, int fmt, int W,
                    int H, int pitch, void* stream) {
    VelocityArgs a;
    for (int k = 0; k < 16; ++k) a.Mi[k] = r.inverse_view_proj[k], a.Mp[k] = r.prev_view_proj[k];
    for (int k = 0; k < 4; ++k) a.jit[k] = r.jitter_current_prev[k];
    a.depth = depth, a.ids = ids, a.objects = r.object_prev_from_current, a.velocity = velocity;
    a.W = W, a.H = H, a.pitch = pitch, a.n_objects = ids ? r.n_objects : 0;
    const dim3 grid((unsigned)((W + 63) / 64), (unsigned)((H + 15) / 16));
    hipStream_t st = static_cast<hipStream_t>(stream);
    const bool with_ids = ids != nullptr;
    switch (fmt) {
        case TAA_VEL_RG32F: launch_fmt<TAA_VEL_RG32F>(a, with_ids, grid, st); break;
        case TAA_VEL_RG16F: launch_fmt<TAA_VEL_RG16F>(a, with_ids, grid, st); break;
        case TAA_VEL_RGBA32F: launch_fmt<TAA_VEL_RGBA32F>(a, with_ids, grid, st); break;
        default: launch_fmt<TAA_VEL_RGBA16F>(a, with_ids, grid, st); break;
    }
    return (int)hipGetLastError();
}

}  // namespace taastar
