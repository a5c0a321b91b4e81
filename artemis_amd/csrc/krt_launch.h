// Launch-shape constants shared by the traversal-based kernels.
#pragma once

#include <cstdint>

namespace krt {

constexpr int kTraceBlock = 128;        // lanes per workgroup (2 wavefronts)
constexpr int kLdsDepth = 32;           // per-lane stack entries held in LDS
// Shadow-ray list kernels: a 16-entry LDS stack (8 KB per workgroup) lets
// eight waves per SIMD stay resident; deeper entries spill to HBM.
constexpr int kLdsDepthLean = 16;
constexpr uint32_t kMaxResidentBlocks = 256 * 16;  // persistent grid: 16 workgroups per CU

struct Context;
struct Scene;
int *spill_buffer(Context &ctx, Scene &s, uint32_t threads, int lds_depth = kLdsDepth);
uint32_t persistent_blocks(uint32_t work, uint32_t block, uint32_t max_blocks);

}  // namespace krt
