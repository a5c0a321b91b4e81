// HDR -> LDR output of a float4 colour buffer: hdr_to_ldr of
// include/cuda/color.cuh:82-121 as a gfx950 kernel.
//
// Per pixel: ACES (color.cuh:35-41) or sRGB (color.cuh:14-31), then
// quantize_unsigned_8_bits (color.cuh:43-56) of each channel, packed as
// x | y << 8 | z << 16 | 255 << 24 (to_ui32, include/cuda/cast.cuh:16-20), and
// written to the vertically flipped row (color.cuh:92-93).  clamp is
// fmaxf(a, fminf(f, b)) (include/cuda/math.cuh:101-104), so a NaN channel
// clamps to 1 and quantizes to 255, as in the reference.
//
// HBM-bound: 16 B read + 4 B written per pixel.  Each lane converts one pixel;
// a wavefront reads 1 KB and writes 256 B of one row, both contiguous.
#include <hip/hip_runtime.h>

#include "krt_internal.h"

namespace krt {

namespace {

__device__ __forceinline__ float clampf(float f, float a, float b) { return fmaxf(a, fminf(f, b)); }

__device__ __forceinline__ float aces1(float c) {
    float a = c * (2.51f * c + 0.03f);
    float b = c * (2.43f * c + 0.59f) + 0.14f;
    return clampf(a / b, 0.0f, 1.0f);
}

__device__ __forceinline__ float srgb1(float c) {
    const float inv_gamma = 1.0f / 2.4f;
    float powed = powf(c, inv_gamma);
    float s = c < 0.0031308f ? 12.92f * c : 1.055f * powed - 0.055f;
    return clampf(s, 0.0f, 1.0f);
}

__device__ __forceinline__ uint32_t quantize8(float x) {
    x = clampf(x, 0.0f, 1.0f);
    return min((uint32_t)(x * 256.0f), 255u);
}

__global__ __launch_bounds__(256) void hdr_to_ldr_kernel(const float4 *__restrict__ pixels,
                                                         uint32_t *__restrict__ target, uint32_t width,
                                                         uint32_t height, int srgb) {
    const uint32_t x = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t y = blockIdx.y;
    if (x >= width) return;
    const float4 p = pixels[(size_t)y * width + x];
    float r, g, b;
    if (srgb) {
        r = srgb1(p.x);
        g = srgb1(p.y);
        b = srgb1(p.z);
    } else {
        r = aces1(p.x);
        g = aces1(p.y);
        b = aces1(p.z);
    }
    const uint32_t v = quantize8(r) | (quantize8(g) << 8) | (quantize8(b) << 16) | (255u << 24);
    target[(size_t)(height - y - 1) * width + x] = v;
}

}  // namespace

void hdr_to_ldr(Context &ctx, const float4 *pixels, uint32_t *target, uint32_t width, uint32_t height,
                int tonemapping) {
    KRT_REQUIRE(tonemapping == 0 || tonemapping == 1, KRT_ERR_INVALID, "tonemapping must be 0 (ACES) or 1 (sRGB)");
    KRT_REQUIRE(height <= 65535u, KRT_ERR_INVALID, "height must be at most 65535");
    if (width == 0 || height == 0) return;
    dim3 grid((width + 255) / 256, height);
    hdr_to_ldr_kernel<<<grid, 256, 0, ctx.stream>>>(pixels, target, width, height, tonemapping);
    KRT_HIP(hipGetLastError());
}

}  // namespace krt
