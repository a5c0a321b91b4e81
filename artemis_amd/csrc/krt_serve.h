// Render-request server (krt_serve.hip): continuous batching of path-tracer
// and ReSTIR DI requests over resident scenes, reservoirs resident per
// request, admission by HBM budget.
#pragma once

#include "krt_internal.h"

namespace krt {

constexpr int kServeMaxActive = 64;           // requests sharing one step
constexpr int kServeMaxDepth = 8;             // path depth cap (history slab)
constexpr int kServeSpillDepth = 64;          // traversal stack levels beyond LDS
constexpr uint32_t kServeBlocks = 256 * 8;    // persistent grid: 8 workgroups per CU
constexpr uint32_t kServeMaxFrames = 64;      // frames of one request in one step

struct ServeConfig {
    size_t hbm_budget = (size_t)8 << 30;  // fixed scratch + per-request AOV buffers
    uint32_t max_active = 16;
    int use_graph = 1;
    double max_queue_ms = 0.0;  // > 0: queued requests older than this are rejected
    uint64_t step_pixels = 1ull << 22;  // pixel-frames one step aims at, shared fairly
    // path-tracer requests of at least this many pixels run the wavefront
    // passes on a lane; smaller ones share the persistent batched launch
    uint64_t wave_pixels = 1ull << 18;
};

constexpr int kServeLanes = 3;  // streams that run lane requests' frames side by side

struct Server;
Server *serve_create(Context &ctx, const ServeConfig &cfg);
void serve_destroy(Server *s);
// kind KRT_REQUEST_PATH_TRACER or KRT_REQUEST_RESTIR (ropt: the restir
// attachment's options)
uint64_t serve_submit(Server *s, Scene &scene, const krt_camera &cam, uint32_t w, uint32_t h, uint32_t spp,
                      int max_depth, uint32_t seed, int kind, const RestirOptions &ropt);
// KRT_OK with the colour buffer (W x H x float4) or KRT_ERR_BUDGET when the
// request was rejected; the request is forgotten afterwards.
int serve_wait(Server *s, uint64_t id, float *color, double *latency_ms, uint32_t *w, uint32_t *h);
void serve_drain(Server *s);
void serve_stats(Server *s, krt_server_stats &out);

}  // namespace krt
