// Render-request server: continuous batching of path-tracer and ReSTIR DI
// requests.
//
// A request is {scene, camera, resolution, spp, depth, seed, kind}; its
// answer is the colour buffer the path-tracer (or ReSTIR DI) attachment would
// hold after `spp` accumulated frames (ArmadaRTX::render,
// source/amadeus/armada.cu:733-780, with the estimators of
// source/optix/amadeus_path_tracer.cu:73-228 and amadeus_restir.cu).
//
// One host thread runs the loop.  Each step advances every active request by
// its fair share of `step_pixels` pixel-frames (at least one frame):
//  * small path-tracer requests (fewer than `wave_pixels` pixels) in a single
//    persistent launch: the grid strides over the concatenated pixels of all
//    of them and finds each pixel's request by binary search in a device
//    table {LaunchInfo, SceneView, first pixel, frames}.  Small requests run
//    several consecutive frames of a pixel in one thread (the accumulation is
//    per pixel, so the result is the one a launch per frame gives) and finish
//    in one step.  The step (table upload + kernel) is a hipGraph captured
//    once; only the table's contents change between steps.  A request this
//    small cannot fill the device on its own, so batching them is what keeps
//    it busy.
//  * large path-tracer requests frame by frame through the attachment's own
//    wavefront passes (pt_wave: ray lists sorted by 8x8 tile, one list walk
//    per level), and ReSTIR DI requests through theirs (restir_render).  Each
//    such request runs on one of kServeLanes streams for its whole life (its
//    frames stay in order), the lanes side by side with the batched launch.
//    Its wavefront buffers (path tracer) or reservoirs, G-buffer and ray lists
//    (ReSTIR) are allocated when it is admitted and stay resident until its
//    last frame; the path tracer's are swapped into the lane's context for
//    the frames it enqueues.
// Requests arriving while others run join at the next step as soon as a slot
// and HBM budget are free (FIFO admission, optional queueing deadline);
// finished requests leave after their last frame and their buffers are
// released.  Scenes (vertex data and BVHs) stay resident and cannot be freed
// or rebuilt while a request holds them; every request keeps its own
// buffers, so a pixel's value never depends on what else shares the step.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "krt_device.h"
#include "krt_internal.h"
#include "krt_launch.h"
#include "krt_pt.h"
#include "krt_serve.h"

namespace krt {

namespace {

struct ServeDesc {
    LaunchInfo li;  // li.samples: the first frame of this step
    SceneView sv;
    uint32_t pix_begin;
    uint32_t frames;  // consecutive frames this step
};

struct ServeTable {
    uint32_t count, total;
    ServeDesc d[kServeMaxActive];
};

__global__ __launch_bounds__(kTraceBlock) void serve_kernel(const ServeTable *__restrict__ tab, float4 *hist_base,
                                                            int *spill) {
    __shared__ int lds_stack[kLdsDepth * kTraceBlock];
    const uint32_t gtid = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t nthreads = gridDim.x * blockDim.x;
    TraceStack stk{lds_stack + threadIdx.x, kTraceBlock, kLdsDepth, spill + gtid, (int)nthreads};
    History hist{hist_base + gtid, nthreads};
    TravStats ts;
    RayCount rc;
    const uint32_t count = tab->count, total = tab->total;
    for (uint32_t g = gtid; g < total; g += nthreads) {
        int lo = 0, hi = (int)count - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (tab->d[mid].pix_begin <= g) lo = mid;
            else hi = mid - 1;
        }
        const ServeDesc &D = tab->d[lo];
        const int first = D.li.samples, frames = (int)D.frames;
        for (int f = 0; f < frames; f++) pt_pixel(D.li, D.sv, g - D.pix_begin, first + f, stk, hist, ts, rc, nullptr);
    }
}

double now_ms() {
    using namespace std::chrono;
    return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

}  // namespace

struct ServeRequest {
    uint64_t id = 0;
    Scene *scene = nullptr;
    uint32_t w = 0, h = 0, spp = 0, seed = 0;
    int max_depth = 5;
    int kind = KRT_REQUEST_PATH_TRACER;
    RestirOptions ropt;
    RestirState rst;  // ReSTIR: reservoirs, G-buffer and ray lists, resident while active
    bool on_lane = false;  // ReSTIR, or a path-tracer request of at least wave_pixels pixels
    int lane = 0;          // the stream its frames run on
    WaveScratch wave;      // path tracer on a lane: its wavefront buffers, resident while active
    krt_camera cam{};
    int state = 0;  // 0 queued, 1 active, 2 done, 3 rejected
    uint32_t frames = 0, step_frames = 1;
    double t_submit = 0, t_admit = 0, t_done = 0;
    size_t bytes = 0;
    DevBuf aov;  // colour, normal, albedo, position planes (float4 x W x H each)
    std::vector<float> result;
    ~ServeRequest() {
        if (scene && state < 2) scene->serve_refs--;
    }
    void finish(int st) {  // done (2) or rejected (3): the scene is free again
        if (state < 2 && scene) scene->serve_refs--;
        state = st;
    }
};

struct Server {
    Context *ctx = nullptr;
    ServeConfig cfg;
    hipStream_t stream = nullptr;
    // lanes: a context shell per stream (its stream, traversal-stack spill
    // slab and counters), so that requests on different lanes share no
    // scratch
    std::unique_ptr<Context> lanes[kServeLanes];
    hipEvent_t lane_done[kServeLanes] = {};
    int next_lane = 0;
    std::thread th;
    std::mutex mu;
    std::condition_variable cv, done_cv;
    std::deque<ServeRequest *> queue;
    std::vector<ServeRequest *> active;
    std::map<uint64_t, std::unique_ptr<ServeRequest>> all;
    uint64_t next_id = 1;
    bool stop = false;
    std::string fault;  // first error raised on the server thread
    ServeTable *table_host = nullptr;
    DevBuf table_dev, hist, spill;
    uint32_t nblocks = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipEvent_t ev[2] = {nullptr, nullptr};
    size_t scratch = 0, in_use = 0, peak = 0;
    size_t fixed_scratch = 0;  // the batched launch's table, history and spill slabs
    // statistics
    std::vector<double> latency, step_ms;
    uint64_t submitted = 0, completed = 0, rejected = 0, steps = 0, max_concurrent = 0;
    double samples = 0, busy_ms = 0, t_first = -1, t_last = 0;

    void run();
    void step(std::vector<ServeRequest *> &batch);
};

size_t serve_request_bytes(uint32_t w, uint32_t h) { return sizeof(float4) * 4 * (size_t)w * h; }

// Device memory of a ReSTIR DI request: AOVs, two reservoir buffers, the
// G-buffer (material id, texture coordinates, intermediate, auxiliary), the shadow-ray list (rays
// and packed copy, S + 1 slots per pixel) with its occlusion bytes, 1/Z
// weights and flags, and the initial pass's wavefront state and closest-ray
// list (pixels rounded up to whole 256-lane workgroups of 8x8 tiles).
size_t serve_restir_bytes(uint32_t w, uint32_t h, const RestirOptions &o, int depth) {
    const size_t npix = (size_t)w * h;
    const size_t items = (size_t)((w + 7) / 8) * ((h + 7) / 8) * 64;
    const size_t slots = (items + 255) / 256 * 256;
    const size_t per_lane = (size_t)std::max(o.neighbours + 1, 2);
    const size_t levels = (size_t)depth + 1;
    size_t b = serve_request_bytes(w, h);
    b += 2 * 64 * npix + 4 * npix + 8 * npix + 2 * 16 * npix;
    b += 2 * 32 * per_lane * slots + per_lane * npix + 4 * (size_t)o.neighbours * npix + npix;
    b += 16 * npix + 16 * npix + 48 * levels * npix + levels * npix + 2 * 32 * slots;
    return b + (1u << 16);
}

void Server::step(std::vector<ServeRequest *> &all_batch) {
    std::vector<ServeRequest *> batch, laned;
    for (ServeRequest *r : all_batch) (r->on_lane ? laned : batch).push_back(r);
    KRT_HIP(hipEventRecord(ev[0], stream));
    // requests on lanes: the frames of this step, queued on their streams
    bool used[kServeLanes] = {};
    for (ServeRequest *r : laned) {
        Context &lane = *lanes[r->lane];
        if (!used[r->lane]) KRT_HIP(hipStreamWaitEvent(lane.stream, ev[0], 0));
        used[r->lane] = true;
        if (r->kind == KRT_REQUEST_PATH_TRACER) std::swap(lane.wave, r->wave);
        for (uint32_t f = 0; f < r->step_frames; f++) {
            LaunchInfo li{};
            li.width = r->w;
            li.height = r->h;
            std::memcpy(li.center, r->cam.center, sizeof(li.center));
            std::memcpy(li.U, r->cam.U, sizeof(li.U));
            std::memcpy(li.V, r->cam.V, sizeof(li.V));
            std::memcpy(li.W, r->cam.W, sizeof(li.W));
            std::memcpy(li.pv, r->cam.pv, sizeof(li.pv));
            li.max_depth = r->max_depth;
            li.samples = (int)(r->frames + f);
            li.accumulate = 1;
            li.user_seed = r->seed;
            li.jitter = 1;
            float4 *planes = r->aov.as<float4>();
            const size_t n = (size_t)r->w * r->h;
            li.color = planes;
            li.normal = planes + n;
            li.albedo = planes + 2 * n;
            li.position = planes + 3 * n;
            if (r->kind == KRT_REQUEST_RESTIR) {
                PassTimes times;
                restir_render(lane, *r->scene, li, r->ropt, r->rst, nullptr, times, nullptr, nullptr, false);
            } else {
                pt_wave(lane, *r->scene, li, nullptr, nullptr, nullptr, nullptr);
            }
        }
        if (r->kind == KRT_REQUEST_PATH_TRACER) std::swap(lane.wave, r->wave);
    }
    for (int l = 0; l < kServeLanes; l++)
        if (used[l]) {
            KRT_HIP(hipEventRecord(lane_done[l], lanes[l]->stream));
            KRT_HIP(hipStreamWaitEvent(stream, lane_done[l], 0));
        }
    uint32_t total = 0;
    table_host->count = (uint32_t)batch.size();
    for (size_t i = 0; i < batch.size(); i++) {
        ServeRequest &r = *batch[i];
        ServeDesc &d = table_host->d[i];
        std::memset(&d, 0, sizeof(d));
        LaunchInfo &li = d.li;
        li.width = r.w;
        li.height = r.h;
        std::memcpy(li.center, r.cam.center, sizeof(li.center));
        std::memcpy(li.U, r.cam.U, sizeof(li.U));
        std::memcpy(li.V, r.cam.V, sizeof(li.V));
        std::memcpy(li.W, r.cam.W, sizeof(li.W));
        std::memcpy(li.pv, r.cam.pv, sizeof(li.pv));
        li.max_depth = r.max_depth;
        li.samples = (int)r.frames;
        li.accumulate = 1;
        li.user_seed = r.seed;
        li.jitter = 1;
        float4 *planes = r.aov.as<float4>();
        const size_t n = (size_t)r.w * r.h;
        li.color = planes;
        li.normal = planes + n;
        li.albedo = planes + 2 * n;
        li.position = planes + 3 * n;
        d.sv = make_view(*r.scene);
        d.pix_begin = total;
        d.frames = r.step_frames;
        total += r.w * r.h;
    }
    table_host->total = total;
    if (!batch.empty()) {
        if (exec) {
            KRT_HIP(hipGraphLaunch(exec, stream));
        } else {
            KRT_HIP(hipMemcpyAsync(table_dev.p, table_host, sizeof(ServeTable), hipMemcpyHostToDevice, stream));
            serve_kernel<<<nblocks, kTraceBlock, 0, stream>>>(table_dev.as<ServeTable>(), hist.as<float4>(),
                                                             spill.as<int>());
            KRT_HIP(hipGetLastError());
        }
    }
    KRT_HIP(hipEventRecord(ev[1], stream));
    KRT_HIP(hipEventSynchronize(ev[1]));
    float ms = 0.0f;
    KRT_HIP(hipEventElapsedTime(&ms, ev[0], ev[1]));
    busy_ms += ms;
    step_ms.push_back(ms);
    steps++;
}

void Server::run() {
    try {
        KRT_HIP(hipSetDevice(ctx->device));
        std::unique_lock<std::mutex> lk(mu);
        while (true) {
            cv.wait(lk, [&] { return stop || !queue.empty() || !active.empty(); });
            if (stop) break;
            const double t = now_ms();
            // queueing deadline
            if (cfg.max_queue_ms > 0)
                for (auto it = queue.begin(); it != queue.end();) {
                    if (t - (*it)->t_submit > cfg.max_queue_ms) {
                        (*it)->finish(3);
                        (*it)->t_done = t;
                        rejected++;
                        it = queue.erase(it);
                    } else {
                        ++it;
                    }
                }
            // FIFO admission by slots and HBM budget
            while (!queue.empty() && active.size() < cfg.max_active) {
                ServeRequest *r = queue.front();
                if (scratch + in_use + r->bytes > cfg.hbm_budget) break;
                r->aov.alloc(serve_request_bytes(r->w, r->h));
                // frame 0 accumulates dst * 0 + sample: recycled memory holding
                // a non-finite value would poison it (stream-ordered, outside the
                // captured step)
                if (r->on_lane) {
                    r->lane = next_lane;
                    next_lane = (next_lane + 1) % kServeLanes;
                    Context &lane = *lanes[r->lane];
                    KRT_HIP(hipMemsetAsync(r->aov.p, 0, r->aov.bytes, lane.stream));
                    if (r->kind == KRT_REQUEST_RESTIR)
                        restir_alloc(lane, r->rst, r->w, r->h, r->seed);  // reservoirs resident from here on
                    else
                        wave_scratch_grow(r->wave, r->w, r->h, r->max_depth, r->scene->instanced, false);
                } else {
                    KRT_HIP(hipMemsetAsync(r->aov.p, 0, r->aov.bytes, stream));
                }
                in_use += r->bytes;
                peak = std::max(peak, in_use);
                r->state = 1;
                r->t_admit = t;
                active.push_back(r);
                queue.pop_front();
            }
            done_cv.notify_all();
            if (active.empty()) continue;
            max_concurrent = std::max<uint64_t>(max_concurrent, active.size());
            std::vector<ServeRequest *> batch = active;
            const double share = (double)cfg.step_pixels / (double)batch.size();
            for (ServeRequest *r : batch) {
                const double k = std::floor(share / ((double)r->w * r->h));
                const uint32_t left = r->spp - r->frames;
                r->step_frames = (uint32_t)std::max(1.0, std::min<double>({k, (double)left, (double)kServeMaxFrames}));
            }
            lk.unlock();
            step(batch);
            std::vector<ServeRequest *> finished;
            for (ServeRequest *r : batch) {
                r->frames += r->step_frames;
                if (r->frames >= r->spp) {
                    r->result.resize(4 * (size_t)r->w * r->h);
                    KRT_HIP(hipMemcpyAsync(r->result.data(), r->aov.p, sizeof(float) * r->result.size(),
                                           hipMemcpyDeviceToHost, stream));  // after the lanes (step waits)
                    finished.push_back(r);
                }
            }
            if (!finished.empty()) KRT_HIP(hipStreamSynchronize(stream));
            for (ServeRequest *r : finished) {
                r->aov.release();
                r->rst = RestirState();
                r->wave = WaveScratch();
            }
            // the lanes' traversal-stack slabs grow with the deepest scene
            // they traversed: counted as fixed scratch from then on
            size_t lane_bytes = 0;
            for (auto &l : lanes) lane_bytes += l->spill.bytes + l->counters.bytes;
            lk.lock();
            scratch = fixed_scratch + lane_bytes;
            const double td = now_ms();
            for (ServeRequest *r : finished) {
                r->finish(2);
                r->t_done = td;
                in_use -= r->bytes;
                latency.push_back(td - r->t_submit);
                samples += (double)r->w * r->h * r->spp;
                completed++;
                t_last = td;
                active.erase(std::find(active.begin(), active.end(), r));
            }
            done_cv.notify_all();
        }
    } catch (const std::exception &e) {
        std::lock_guard<std::mutex> g(mu);
        fault = e.what();
        stop = true;
        done_cv.notify_all();
    }
}

Server *serve_create(Context &ctx, const ServeConfig &cfg) {
    KRT_REQUIRE(cfg.max_active >= 1 && cfg.max_active <= (uint32_t)kServeMaxActive, KRT_ERR_INVALID,
                "max_active must be in [1, 64]");
    auto s = std::make_unique<Server>();
    s->ctx = &ctx;
    s->cfg = cfg;
    KRT_HIP(hipSetDevice(ctx.device));
    KRT_HIP(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
    KRT_HIP(hipHostMalloc((void **)&s->table_host, sizeof(ServeTable), hipHostMallocDefault));
    std::memset(s->table_host, 0, sizeof(ServeTable));
    s->table_dev.alloc(sizeof(ServeTable));
    s->nblocks = kServeBlocks;
    const size_t threads = (size_t)s->nblocks * kTraceBlock;
    s->hist.alloc(sizeof(float4) * 2 * (size_t)(kServeMaxDepth + 1) * threads);
    s->spill.alloc(sizeof(int) * (size_t)kServeSpillDepth * threads);
    s->scratch = s->fixed_scratch = s->table_dev.bytes + s->hist.bytes + s->spill.bytes;
    KRT_REQUIRE(s->scratch < cfg.hbm_budget, KRT_ERR_BUDGET, "HBM budget below the server's fixed scratch");
    for (auto &e : s->ev) KRT_HIP(hipEventCreate(&e));
    for (int l = 0; l < kServeLanes; l++) {
        s->lanes[l] = std::make_unique<Context>();
        s->lanes[l]->device = ctx.device;
        KRT_HIP(hipStreamCreateWithFlags(&s->lanes[l]->stream, hipStreamNonBlocking));
        s->lanes[l]->side = s->lanes[l]->stream;
        KRT_HIP(hipEventCreate(&s->lane_done[l]));
    }
    if (cfg.use_graph) {
        // one step = table upload + batched kernel, captured once
        KRT_HIP(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
        (void)hipMemcpyAsync(s->table_dev.p, s->table_host, sizeof(ServeTable), hipMemcpyHostToDevice, s->stream);
        serve_kernel<<<s->nblocks, kTraceBlock, 0, s->stream>>>(s->table_dev.as<ServeTable>(), s->hist.as<float4>(),
                                                               s->spill.as<int>());
        KRT_HIP(hipStreamEndCapture(s->stream, &s->graph));
        KRT_HIP(hipGraphInstantiate(&s->exec, s->graph, nullptr, nullptr, 0));
    }
    Server *raw = s.release();
    raw->th = std::thread([raw] { raw->run(); });
    return raw;
}

void serve_destroy(Server *s) {
    if (!s) return;
    {
        std::lock_guard<std::mutex> g(s->mu);
        s->stop = true;
        for (ServeRequest *r : s->queue) r->finish(3);
        s->queue.clear();
    }
    s->cv.notify_all();
    s->done_cv.notify_all();
    if (s->th.joinable()) s->th.join();
    (void)hipStreamSynchronize(s->stream);
    for (int l = 0; l < kServeLanes; l++)
        if (s->lanes[l]) (void)hipStreamSynchronize(s->lanes[l]->stream);
    if (s->exec) (void)hipGraphExecDestroy(s->exec);
    if (s->graph) (void)hipGraphDestroy(s->graph);
    for (auto &e : s->ev)
        if (e) (void)hipEventDestroy(e);
    s->all.clear();
    for (int l = 0; l < kServeLanes; l++) {
        if (!s->lanes[l]) continue;
        s->lanes[l]->spill.release();
        s->lanes[l]->counters.release();
        (void)hipStreamDestroy(s->lanes[l]->stream);
        if (s->lane_done[l]) (void)hipEventDestroy(s->lane_done[l]);
        s->lanes[l].reset();
    }
    s->table_dev.release();
    s->hist.release();
    s->spill.release();
    if (s->table_host) (void)hipHostFree(s->table_host);
    (void)hipStreamDestroy(s->stream);
    delete s;
}

uint64_t serve_submit(Server *s, Scene &scene, const krt_camera &cam, uint32_t w, uint32_t h, uint32_t spp,
                      int max_depth, uint32_t seed, int kind, const RestirOptions &ropt) {
    KRT_REQUIRE(kind == KRT_REQUEST_PATH_TRACER || kind == KRT_REQUEST_RESTIR, KRT_ERR_INVALID,
                "unknown request kind");
    if (kind == KRT_REQUEST_RESTIR)
        KRT_REQUIRE(ropt.neighbours >= 0 && ropt.neighbours <= 8 && ropt.candidates >= 0 && ropt.m_cap >= 0 &&
                        ropt.radius >= 0 && ropt.spatial_passes >= 1,
                    KRT_ERR_INVALID, "restir options out of range");
    KRT_REQUIRE(scene.accel.built, KRT_ERR_STATE, "scene has no acceleration structure");
    KRT_REQUIRE(!scene.stale, KRT_ERR_STATE, "vertices moved: call krt_accel_refit or krt_accel_build first");
    KRT_REQUIRE(w > 0 && h > 0 && (uint64_t)w * h <= (1ull << 26), KRT_ERR_INVALID, "resolution out of range");
    KRT_REQUIRE(spp >= 1, KRT_ERR_INVALID, "spp must be >= 1");
    KRT_REQUIRE(max_depth >= 0 && max_depth <= kServeMaxDepth, KRT_ERR_INVALID, "max_depth must be in [0, 8]");
    KRT_REQUIRE(scene.accel.depth_bound < (uint32_t)(kLdsDepth + kServeSpillDepth), KRT_ERR_INVALID,
                "scene BVH deeper than the server's traversal stack");
    for (int k = 0; k < 3; k++)
        KRT_REQUIRE(std::isfinite(cam.center[k]) && std::isfinite(cam.U[k]) && std::isfinite(cam.V[k]) &&
                        std::isfinite(cam.W[k]),
                    KRT_ERR_INVALID, "camera has non-finite components");
    auto r = std::make_unique<ServeRequest>();
    r->scene = &scene;
    r->cam = cam;
    r->w = w;
    r->h = h;
    r->spp = spp;
    r->max_depth = max_depth;
    r->seed = seed;
    r->kind = kind;
    r->ropt = ropt;
    r->on_lane = kind == KRT_REQUEST_RESTIR || (uint64_t)w * h >= s->cfg.wave_pixels;
    if (kind == KRT_REQUEST_RESTIR) {
        r->bytes = serve_restir_bytes(w, h, ropt, max_depth);
    } else {
        r->bytes = serve_request_bytes(w, h);
        if (r->on_lane) r->bytes += wave_scratch_grow(r->wave, w, h, max_depth, scene.instanced, false, true);
    }
    std::lock_guard<std::mutex> g(s->mu);
    KRT_REQUIRE(s->fault.empty(), KRT_ERR_HIP, "server stopped: " + s->fault);
    KRT_REQUIRE(s->scratch + r->bytes <= s->cfg.hbm_budget, KRT_ERR_BUDGET,
                "request needs more HBM than the server's budget");
    r->id = s->next_id++;
    scene.serve_refs++;  // held until the request is done or rejected
    r->t_submit = now_ms();
    if (s->t_first < 0) s->t_first = r->t_submit;
    s->submitted++;
    s->queue.push_back(r.get());
    const uint64_t id = r->id;
    s->all[id] = std::move(r);
    s->cv.notify_all();
    return id;
}

int serve_wait(Server *s, uint64_t id, float *color, double *latency_ms, uint32_t *w, uint32_t *h) {
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->all.find(id);
    KRT_REQUIRE(it != s->all.end(), KRT_ERR_INVALID, "unknown request id");
    ServeRequest *r = it->second.get();
    s->done_cv.wait(lk, [&] { return r->state >= 2 || !s->fault.empty(); });
    KRT_REQUIRE(s->fault.empty() || r->state >= 2, KRT_ERR_HIP, "server stopped: " + s->fault);
    if (latency_ms) *latency_ms = r->t_done - r->t_submit;
    if (w) *w = r->w;
    if (h) *h = r->h;
    const int st = r->state;
    if (st == 2 && color) std::memcpy(color, r->result.data(), sizeof(float) * r->result.size());
    s->all.erase(it);
    return st == 2 ? KRT_OK : KRT_ERR_BUDGET;
}

void serve_drain(Server *s) {
    std::unique_lock<std::mutex> lk(s->mu);
    s->done_cv.wait(lk, [&] { return (s->queue.empty() && s->active.empty()) || !s->fault.empty(); });
    KRT_REQUIRE(s->fault.empty(), KRT_ERR_HIP, "server stopped: " + s->fault);
}

static double percentile(std::vector<double> v, double q) {
    if (v.empty()) return 0.0;
    std::sort(v.begin(), v.end());
    const double pos = q * (double)(v.size() - 1);
    const size_t i = (size_t)pos;
    const double f = pos - (double)i;
    return i + 1 < v.size() ? v[i] * (1.0 - f) + v[i + 1] * f : v[i];
}

void serve_stats(Server *s, krt_server_stats &out) {
    std::lock_guard<std::mutex> g(s->mu);
    std::memset(&out, 0, sizeof(out));
    out.submitted = s->submitted;
    out.completed = s->completed;
    out.rejected = s->rejected;
    out.active = (uint32_t)s->active.size();
    out.queued = (uint32_t)s->queue.size();
    out.max_concurrent = (uint32_t)s->max_concurrent;
    out.steps = s->steps;
    out.p50_ms = percentile(s->latency, 0.5);
    out.p99_ms = percentile(s->latency, 0.99);
    out.step_p50_ms = percentile(s->step_ms, 0.5);
    out.step_p99_ms = percentile(s->step_ms, 0.99);
    out.samples = s->samples;
    out.busy_ms = s->busy_ms;
    out.wall_ms = s->t_first < 0 ? 0.0 : s->t_last - s->t_first;
    out.hbm_scratch = s->scratch;
    out.hbm_in_use = s->in_use;
    out.hbm_peak = s->peak;
}

}  // namespace krt
