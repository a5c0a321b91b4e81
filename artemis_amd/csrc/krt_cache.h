// Radiance-cache plumbing shared by the path kernels that stop at a cache
// vertex (krt_cache.hip for the path tracer, krt_restir.hip for the indirect
// path of ReSTIR DI): the per-frame view of the query and training buffers,
// the query record, and the host steps around the path kernel.
#pragma once

#include "krt_device.h"
#include "krt_internal.h"

namespace krt {

struct CacheView {
    int enabled = 0;
    int depth = 1;           // level of the cache vertex
    int learn_emission = 0;  // 1: the vertex's emission is learned, not added
    float *query = nullptr;  // npix x 13 query records
    float4 *theta = nullptr; // npix: {Theta, 1 when the cache term applies}
    float *slot_raw = nullptr, *slot_tgt = nullptr;  // M training slots
    uint32_t *slot_ok = nullptr;
    uint32_t M = 0, stride = 1, offset = 0;

    // Training pixels: M of them, `stride` apart from a per-frame offset.
    KRT_DEV bool training(uint32_t pix, uint32_t &slot) const {
        if (M == 0 || pix < offset) return false;
        const uint32_t d = pix - offset;
        if (d % stride) return false;
        slot = d / stride;
        return slot < M;
    }
};

// Query record: {x, w, n, albedo, roughness}.
KRT_DEV void put_query(float *q, f3 x, f3 w, f3 n, const Material &mat) {
    const float v[kMlpRaw] = {x.x, x.y, x.z, w.x, w.y, w.z, n.x, n.y, n.z,
                              mat.diffuse.x, mat.diffuse.y, mat.diffuse.z, mat.roughness};
#pragma unroll
    for (int a = 0; a < kMlpRaw; a++) q[a] = v[a];
}

KRT_DEV bool finite3(f3 v) { return isfinite(v.x) && isfinite(v.y) && isfinite(v.z); }

// Host side (krt_cache.hip).  cache_begin sizes the buffers, initialises the
// network on first use, maps positions into the scene's bounding cube and
// picks this frame's training pixels.
CacheView cache_begin(Context &ctx, Scene &s, const CacheOptions &opt, CacheState &cs, uint32_t npix,
                      uint32_t user_seed);
// MLP inference over every pixel's query record, on `stream`.
void cache_infer(CacheState &cs, uint32_t npix, hipStream_t stream);
// intermediate[pix].xyz += Theta * max(prediction, 0) where the term applies.
void cache_fold_intermediate(CacheState &cs, uint32_t npix, float4 *intermediate, hipStream_t stream);
// Compaction of the valid training slots and one optimiser step.
void cache_train(Context &ctx, CacheState &cs, const CacheView &cv);

}  // namespace krt
