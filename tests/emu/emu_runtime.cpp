// Runtime of the CPU HIP emulation (hip/hip_runtime.h): device memory is host
// memory from posix_memalign (so AddressSanitizer knows each allocation's exact
// bounds), every call completes before it returns, and a kernel's workgroups
// run one after another, each as one fiber per lane on the calling thread.
// A lane runs until it reaches a rendezvous (__syncthreads, a wave operation)
// or returns; when every live lane of the workgroup (or of the wave) waits at
// the same rendezvous, the scheduler computes its results and lets them go on.
#include <hip/hip_runtime.h>
#include <sanitizer/common_interface_defs.h>
#include <sys/mman.h>

#include <chrono>
#include <cstdlib>
#include <mutex>
#include <string>
#include <vector>

krt_emu_idx threadIdx, blockIdx, blockDim, gridDim;

struct krt_emu_stream {
    std::vector<std::function<void()>> *capture = nullptr;
};
struct krt_emu_event {
    std::chrono::steady_clock::time_point t;
};
struct krt_emu_graph {
    std::vector<std::function<void()>> ops;
};

// x86-64 context switch: saves the callee-saved registers, MXCSR and the x87
// control word of the current context on its stack, stores its stack pointer
// in *save and resumes the context whose stack pointer is `load`.
extern "C" void krt_emu_swap(void **save, void *load);
extern "C" void krt_emu_tramp();
asm(R"(
    .text
    .globl krt_emu_swap
    .type krt_emu_swap, @function
krt_emu_swap:
    pushq %rbp
    pushq %rbx
    pushq %r12
    pushq %r13
    pushq %r14
    pushq %r15
    subq $8, %rsp
    stmxcsr (%rsp)
    fnstcw 4(%rsp)
    movq %rsp, (%rdi)
    movq %rsi, %rsp
    ldmxcsr (%rsp)
    fldcw 4(%rsp)
    addq $8, %rsp
    popq %r15
    popq %r14
    popq %r13
    popq %r12
    popq %rbx
    popq %rbp
    ret
    .size krt_emu_swap, .-krt_emu_swap

    .globl krt_emu_tramp
    .type krt_emu_tramp, @function
krt_emu_tramp:
    movq %r12, %rdi
    callq *%r13
    ud2
    .size krt_emu_tramp, .-krt_emu_tramp
)");

namespace krt_emu {
namespace {

constexpr int kMaxLanes = 1024;
constexpr size_t kStack = 512 * 1024;

enum class Wait { Run, Block, Wave, Done };
enum class Op { Sync, Count, Ballot, Exchange, Mfma };

struct Lane {
    void *sp = nullptr;
    char *stack = nullptr;  // lowest address (a guard page), kStack bytes
    void *fake = nullptr;
    Wait wait = Wait::Done;
    Op op = Op::Sync;
    const char *what = "";
    uint64_t value = 0, result = 0;
    int src = 0;
    krt_emu_bf16x8 a, b;
    krt_emu_f32x4 c, d;
};

struct Sched {
    Lane lanes[kMaxLanes];
    int created = 0;
    void *main_sp = nullptr;
    void *main_fake = nullptr;
    const void *main_bottom = nullptr;
    size_t main_size = 0;
    int cur = -1, n = 0;
    const std::function<void()> *body = nullptr;
    const char *name = "";
};

Sched S;

// KRT_EMU_STATS=1: time per kernel, printed at exit
struct Stats {
    struct Row {
        double s = 0;
        long long launches = 0, blocks = 0;
    };
    std::vector<std::pair<std::string, Row>> rows;
    bool on = std::getenv("KRT_EMU_STATS") != nullptr;
    Row &row(const char *name) {
        for (auto &r : rows)
            if (r.first == name) return r.second;
        rows.push_back({name, Row{}});
        return rows.back().second;
    }
    ~Stats() {
        if (!on) return;
        for (auto &r : rows)
            std::fprintf(stderr, "emu stats: %-40s %8lld launches %10lld workgroups %9.3f s\n", r.first.c_str(),
                         r.second.launches, r.second.blocks, r.second.s);
    }
};
Stats g_stats;
std::recursive_mutex g_kernel_mutex;  // one kernel at a time over all host threads
hipError_t g_last = hipSuccess;

void set_index(int i) {
    threadIdx = {(unsigned)i % blockDim.x, ((unsigned)i / blockDim.x) % blockDim.y,
                 (unsigned)i / (blockDim.x * blockDim.y)};
}

// on a lane: back to the scheduler until the lane is resumed
void yield() {
    Lane &L = S.lanes[S.cur];
    __sanitizer_start_switch_fiber(&L.fake, S.main_bottom, S.main_size);
    krt_emu_swap(&L.sp, S.main_sp);
    __sanitizer_finish_switch_fiber(L.fake, &S.main_bottom, &S.main_size);
}

extern "C" void lane_main(void *) {
    __sanitizer_finish_switch_fiber(nullptr, &S.main_bottom, &S.main_size);
    for (;;) {
        (*S.body)();
        S.lanes[S.cur].wait = Wait::Done;
        yield();
    }
}

void create_lane(int i) {
    Lane &L = S.lanes[i];
    void *m = mmap(nullptr, kStack, PROT_READ | PROT_WRITE, MAP_PRIVATE | MAP_ANONYMOUS | MAP_NORESERVE, -1, 0);
    if (m == MAP_FAILED) {
        std::fprintf(stderr, "emu: cannot map a lane stack\n");
        std::abort();
    }
    L.stack = static_cast<char *>(m);
    mprotect(L.stack, 4096, PROT_NONE);  // guard page
    uintptr_t top = ((uintptr_t)L.stack + kStack) & ~(uintptr_t)15;
    uint64_t *sp = reinterpret_cast<uint64_t *>(top - 16 - 64);
    uint32_t mxcsr;
    uint16_t fpcw;
    asm volatile("stmxcsr %0" : "=m"(mxcsr));
    asm volatile("fnstcw %0" : "=m"(fpcw));
    std::memset(sp, 0, 64);
    std::memcpy(reinterpret_cast<char *>(sp), &mxcsr, 4);
    std::memcpy(reinterpret_cast<char *>(sp) + 4, &fpcw, 2);
    sp[1] = 0;                                  // r15
    sp[2] = 0;                                  // r14
    sp[3] = (uint64_t)(uintptr_t)&lane_main;    // r13: entry
    sp[4] = 0;                                  // r12: argument
    sp[5] = 0;                                  // rbx
    sp[6] = 0;                                  // rbp
    sp[7] = (uint64_t)(uintptr_t)&krt_emu_tramp;  // return address
    L.sp = sp;
}

void resume(int i) {
    S.cur = i;
    set_index(i);
    Lane &L = S.lanes[i];
    __sanitizer_start_switch_fiber(&S.main_fake, L.stack + 4096, kStack - 4096);
    krt_emu_swap(&S.main_sp, L.sp);
    __sanitizer_finish_switch_fiber(S.main_fake, nullptr, nullptr);
    S.cur = -1;
}

// Completes the rendezvous that every live lane of wave w waits at, if any.
bool complete_wave(int w) {
    const int lo = 64 * w, hi = std::min(S.n, lo + 64);
    int waiting = 0, live = 0;
    for (int i = lo; i < hi; i++) {
        if (S.lanes[i].wait == Wait::Done) continue;
        live++;
        waiting += S.lanes[i].wait == Wait::Wave;
    }
    if (live == 0 || waiting != live) return false;
    const Op op = S.lanes[lo].wait == Wait::Wave ? S.lanes[lo].op : [&] {
        for (int i = lo; i < hi; i++)
            if (S.lanes[i].wait == Wait::Wave) return S.lanes[i].op;
        return Op::Sync;
    }();
    for (int i = lo; i < hi; i++)
        if (S.lanes[i].wait == Wait::Wave && S.lanes[i].op != op) {
            std::fprintf(stderr, "emu: lanes of one wave in kernel %s meet at different operations (%s, %s)\n",
                         S.name, S.lanes[lo].what, S.lanes[i].what);
            std::abort();
        }
    auto live_lane = [&](int l) { return l >= 0 && l < 64 && lo + l < hi && S.lanes[lo + l].wait == Wait::Wave; };
    if (op == Op::Ballot) {
        uint64_t mask = 0;
        for (int l = 0; l < 64; l++)
            if (live_lane(l) && S.lanes[lo + l].value) mask |= 1ull << l;
        for (int i = lo; i < hi; i++) S.lanes[i].result = mask;
    } else if (op == Op::Exchange) {
        for (int i = lo; i < hi; i++) {
            Lane &L = S.lanes[i];
            if (L.wait != Wait::Wave) continue;
            L.result = live_lane(L.src) ? S.lanes[lo + L.src].value : 0;
        }
    } else if (op == Op::Mfma) {
        for (int i = lo; i < hi; i++) {
            Lane &L = S.lanes[i];
            if (L.wait != Wait::Wave) continue;
            const int lane = i - lo, col = lane % 16;
            for (int r = 0; r < 4; r++) {
                const int row = 4 * (lane / 16) + r;
                float s = 0.0f;
                for (int k = 0; k < 32; k++) {
                    const int la = row + 16 * (k / 8), lb = col + 16 * (k / 8);
                    const float av = live_lane(la) ? (float)S.lanes[lo + la].a[k % 8] : 0.0f;
                    const float bv = live_lane(lb) ? (float)S.lanes[lo + lb].b[k % 8] : 0.0f;
                    s += av * bv;
                }
                L.d[r] = L.c[r] + s;
            }
        }
    }
    for (int i = lo; i < hi; i++)
        if (S.lanes[i].wait == Wait::Wave) S.lanes[i].wait = Wait::Run;
    return true;
}

// Completes the workgroup barrier when every live lane waits at it.
bool complete_block() {
    int live = 0, waiting = 0;
    uint64_t count = 0;
    for (int i = 0; i < S.n; i++) {
        if (S.lanes[i].wait == Wait::Done) continue;
        live++;
        if (S.lanes[i].wait == Wait::Block) {
            waiting++;
            count += S.lanes[i].value != 0;
        }
    }
    if (live == 0 || waiting != live) return false;
    for (int i = 0; i < S.n; i++)
        if (S.lanes[i].wait == Wait::Block) {
            S.lanes[i].result = count;
            S.lanes[i].wait = Wait::Run;
        }
    return true;
}

void run_block() {
    for (int i = 0; i < S.n; i++) S.lanes[i].wait = Wait::Run;
    for (;;) {
        bool any = false;
        for (int i = 0; i < S.n; i++)
            if (S.lanes[i].wait == Wait::Run) {
                resume(i);
                any = true;
            }
        bool done = true;
        for (int i = 0; i < S.n; i++) done = done && S.lanes[i].wait == Wait::Done;
        if (done) return;
        bool moved = false;
        for (int w = 0; 64 * w < S.n; w++) moved = complete_wave(w) || moved;
        if (!moved) moved = complete_block();
        if (!moved && !any) {
            std::fprintf(stderr, "emu: kernel %s, workgroup %u: no lane can go on\n", S.name, blockIdx.x);
            for (int i = 0; i < S.n; i++)
                if (S.lanes[i].wait != Wait::Done)
                    std::fprintf(stderr, "  lane %d waits at %s\n", i, S.lanes[i].what);
            std::abort();
        }
    }
}

void run_kernel(const char *name, dim3 grid, dim3 block, const std::function<void()> &body) {
    const long long lanes = (long long)block.x * block.y * block.z;
    if (lanes < 1 || lanes > kMaxLanes || grid.x == 0 || grid.y == 0 || grid.z == 0) {
        std::fprintf(stderr, "emu: kernel %s launched with grid (%u,%u,%u) block (%u,%u,%u)\n", name, grid.x, grid.y,
                     grid.z, block.x, block.y, block.z);
        g_last = hipErrorInvalidValue;
        return;
    }
    std::lock_guard<std::recursive_mutex> lk(g_kernel_mutex);
    while (S.created < lanes) create_lane(S.created++);
    const krt_emu_idx save_t = threadIdx, save_b = blockIdx, save_d = blockDim, save_g = gridDim;
    S.n = (int)lanes;
    S.body = &body;
    S.name = name;
    blockDim = {block.x, block.y, block.z};
    gridDim = {grid.x, grid.y, grid.z};
    const auto t0 = std::chrono::steady_clock::now();
    for (unsigned z = 0; z < grid.z; z++)
        for (unsigned y = 0; y < grid.y; y++)
            for (unsigned x = 0; x < grid.x; x++) {
                blockIdx = {x, y, z};
                run_block();
            }
    if (g_stats.on) {
        auto &r = g_stats.row(name);
        r.launches++;
        r.blocks += (long long)grid.x * grid.y * grid.z;
        r.s += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
    threadIdx = save_t, blockIdx = save_b, blockDim = save_d, gridDim = save_g;
}

Lane &me() {
    if (S.cur < 0) {
        std::fprintf(stderr, "emu: a device rendezvous outside a kernel\n");
        std::abort();
    }
    return S.lanes[S.cur];
}

uint64_t meet(Wait where, Op op, const char *what, uint64_t value, int src = 0) {
    Lane &L = me();
    L.wait = where;
    L.op = op;
    L.what = what;
    L.value = value;
    L.src = src;
    yield();
    return S.lanes[S.cur].result;
}

}  // namespace

void syncthreads() { meet(Wait::Block, Op::Sync, "__syncthreads", 0); }
int syncthreads_count(int pred) { return (int)meet(Wait::Block, Op::Count, "__syncthreads_count", pred != 0); }
unsigned long long wave_ballot(int pred) { return meet(Wait::Wave, Op::Ballot, "__ballot", pred != 0); }
uint64_t wave_exchange(uint64_t mine, int src) { return meet(Wait::Wave, Op::Exchange, "__shfl", mine, src); }

krt_emu_f32x4 mfma_16x16x32_bf16(krt_emu_bf16x8 a, krt_emu_bf16x8 b, krt_emu_f32x4 c) {
    Lane &L = me();
    L.a = a;
    L.b = b;
    L.c = c;
    meet(Wait::Wave, Op::Mfma, "mfma", 0);
    return S.lanes[S.cur].d;
}

void launch(const char *name, dim3 grid, dim3 block, size_t, hipStream_t s, std::function<void()> body) {
    if (s && s->capture) {
        s->capture->push_back([name, grid, block, body] { run_kernel(name, grid, block, body); });
        return;
    }
    run_kernel(name, grid, block, body);
}

}  // namespace krt_emu

namespace {
template <class F>
hipError_t on_stream(hipStream_t s, F f) {
    if (s && s->capture) s->capture->push_back(f);
    else f();
    return hipSuccess;
}
}  // namespace

// KRT_EMU_FILL: fill every new device allocation with this byte (0-255), or
// with pseudo-random bytes ("random"), as memory that a GPU allocator hands
// out again holds whatever its last owner left; a kernel that reads memory
// nothing has written then gives different answers from run to run.
static void fill_new(void *p, size_t n) {
    static const char *mode = std::getenv("KRT_EMU_FILL");
    if (!mode || !n) return;
    if (std::strcmp(mode, "random") == 0) {
        static uint64_t x = 0x9E3779B97F4A7C15ull;
        unsigned char *b = static_cast<unsigned char *>(p);
        for (size_t i = 0; i < n; i++) {
            x ^= x << 13, x ^= x >> 7, x ^= x << 17;
            b[i] = (unsigned char)x;
        }
    } else {
        std::memset(p, std::atoi(mode) & 255, n);
    }
}

hipError_t hipMalloc(void **p, size_t n) {
    *p = nullptr;
    if (posix_memalign(p, 256, n ? n : 1) != 0) return hipErrorOutOfMemory;
    fill_new(*p, n);
    return hipSuccess;
}
hipError_t hipFree(void *p) {
    std::free(p);
    return hipSuccess;
}
hipError_t hipHostMalloc(void **p, size_t n, unsigned) { return hipMalloc(p, n); }
hipError_t hipHostFree(void *p) { return hipFree(p); }
hipError_t hipMemcpy(void *dst, const void *src, size_t n, hipMemcpyKind) {
    if (n) std::memmove(dst, src, n);
    return hipSuccess;
}
hipError_t hipMemcpyAsync(void *dst, const void *src, size_t n, hipMemcpyKind, hipStream_t s) {
    return on_stream(s, [=] {
        if (n) std::memmove(dst, src, n);
    });
}
hipError_t hipMemset(void *dst, int v, size_t n) {
    if (n) std::memset(dst, v, n);
    return hipSuccess;
}
hipError_t hipMemsetAsync(void *dst, int v, size_t n, hipStream_t s) {
    return on_stream(s, [=] {
        if (n) std::memset(dst, v, n);
    });
}
hipError_t hipStreamCreate(hipStream_t *s) {
    *s = new krt_emu_stream();
    return hipSuccess;
}
hipError_t hipStreamCreateWithFlags(hipStream_t *s, unsigned) { return hipStreamCreate(s); }
hipError_t hipStreamDestroy(hipStream_t s) {
    delete s;
    return hipSuccess;
}
hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }
hipError_t hipStreamWaitEvent(hipStream_t, hipEvent_t, unsigned) { return hipSuccess; }
hipError_t hipStreamBeginCapture(hipStream_t s, hipStreamCaptureMode) {
    if (!s || s->capture) return hipErrorInvalidValue;
    s->capture = new std::vector<std::function<void()>>();
    return hipSuccess;
}
hipError_t hipStreamEndCapture(hipStream_t s, hipGraph_t *g) {
    if (!s || !s->capture) return hipErrorInvalidValue;
    *g = new krt_emu_graph{std::move(*s->capture)};
    delete s->capture;
    s->capture = nullptr;
    return hipSuccess;
}
hipError_t hipGraphInstantiate(hipGraphExec_t *exec, hipGraph_t g, hipGraphNode_t *, char *, size_t) {
    *exec = new krt_emu_graph{g->ops};
    return hipSuccess;
}
hipError_t hipGraphLaunch(hipGraphExec_t exec, hipStream_t) {
    for (auto &op : exec->ops) op();
    return hipSuccess;
}
hipError_t hipGraphExecDestroy(hipGraphExec_t exec) {
    delete exec;
    return hipSuccess;
}
hipError_t hipGraphDestroy(hipGraph_t g) {
    delete g;
    return hipSuccess;
}
hipError_t hipEventCreate(hipEvent_t *e) {
    *e = new krt_emu_event();
    return hipSuccess;
}
hipError_t hipEventDestroy(hipEvent_t e) {
    delete e;
    return hipSuccess;
}
hipError_t hipEventRecord(hipEvent_t e, hipStream_t s) {
    return on_stream(s, [e] { e->t = std::chrono::steady_clock::now(); });
}
hipError_t hipEventSynchronize(hipEvent_t) { return hipSuccess; }
hipError_t hipEventElapsedTime(float *ms, hipEvent_t a, hipEvent_t b) {
    *ms = std::chrono::duration<float, std::milli>(b->t - a->t).count();
    return hipSuccess;
}
hipError_t hipDeviceSynchronize() { return hipSuccess; }
hipError_t hipGetLastError() {
    const hipError_t e = krt_emu::g_last;
    krt_emu::g_last = hipSuccess;
    return e;
}
hipError_t hipSetDevice(int d) { return d == 0 ? hipSuccess : hipErrorInvalidValue; }
hipError_t hipGetDevice(int *d) {
    *d = 0;
    return hipSuccess;
}
hipError_t hipGetDeviceCount(int *n) {
    *n = 1;
    return hipSuccess;
}
hipError_t hipGetSymbolAddress(void **p, const void *symbol) {
    *p = const_cast<void *>(symbol);
    return hipSuccess;
}
hipError_t hipMemGetInfo(size_t *free_bytes, size_t *total_bytes) {
    *free_bytes = (size_t)64 << 30;
    *total_bytes = (size_t)64 << 30;
    return hipSuccess;
}
const char *hipGetErrorString(hipError_t e) {
    switch (e) {
        case hipSuccess: return "no error";
        case hipErrorInvalidValue: return "invalid argument (emulated)";
        case hipErrorOutOfMemory: return "out of memory (emulated)";
        default: return "emulated HIP error";
    }
}
