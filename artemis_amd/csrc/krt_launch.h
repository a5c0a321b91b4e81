// Launch-shape constants shared by the traversal-based kernels.
#pragma once

#include <cstdint>

namespace krt {

constexpr int kTraceBlock = 128;        // lanes per workgroup (2 wavefronts)
constexpr int kLdsDepth = 32;           // per-lane stack entries held in LDS
constexpr uint32_t kMaxResidentBlocks = 256 * 16;  // persistent grid: 16 workgroups per CU

struct Context;
struct Scene;
int *spill_buffer(Context &ctx, Scene &s, uint32_t threads);
uint32_t persistent_blocks(uint32_t work, uint32_t block, uint32_t max_blocks);

}  // namespace krt
