// CPU simulation of the HIP runtime subset the library uses (test
// infrastructure; see include/hip/hip_runtime.h).
//
// Allocations: every hipMalloc / hipHostMalloc gets its own mapping with an
// inaccessible page after the data (SIM_GUARD=end, the default: the data
// ends at the guard, rounded up to 256 bytes) or before it (SIM_GUARD=start),
// so a kernel or a copy that leaves its buffer traps on the first byte out.
// The SIGSEGV handler reports the faulting address, the allocation it lies
// next to, the kernel, workgroup and thread, and exits with status 86.
//
// Launches run synchronously: workgroup by workgroup, each thread a fiber on
// the calling OS thread.  __syncthreads() parks a fiber until every live
// fiber of the block has arrived; a wave collective or wave barrier parks it until every
// live lane of its 64-lane wave has.  A barrier no fiber can pass (threads
// of a block waiting at different barriers) is reported as a deadlock
// (exit status 87).  Streams and events are trivial; graph capture records
// launches and replays them.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <csignal>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <mutex>
#include <string>
#include <sys/mman.h>
#include <unistd.h>
#include <vector>

// ---------------------------------------------------------------------------
// guarded allocations
// ---------------------------------------------------------------------------
namespace {

struct Alloc {
    char *map;       // whole mapping (guard pages included)
    size_t map_len;
    char *ptr;       // user pointer
    size_t bytes;    // requested bytes
    bool host;
    uint64_t id;
};

std::mutex g_alloc_mtx;
std::map<uintptr_t, Alloc> g_allocs;  // by map start
uint64_t g_next_id = 1;
size_t g_page = 4096;

// SIM_FILL=<byte>: fresh allocations hold that byte (a GPU's hipMalloc and
// hipHostMalloc memory is not cleared); unset: zeros
int fill_byte()
{
    static int v = -2;
    if (v == -2) {
        const char *e = getenv("SIM_FILL");
        v = e ? (int)(strtol(e, nullptr, 0) & 0xff) : -1;
    }
    return v;
}

bool guard_at_start()
{
    static int v = -1;
    if (v < 0) {
        const char *e = getenv("SIM_GUARD");
        v = (e && std::strcmp(e, "start") == 0) ? 1 : 0;
    }
    return v == 1;
}

void *guarded_alloc(size_t bytes, bool host)
{
    const size_t data = ((std::max<size_t>(bytes, 1) + g_page - 1) / g_page) * g_page;
    const size_t len = data + 2 * g_page;
    char *m = static_cast<char *>(mmap(nullptr, len, PROT_NONE, MAP_PRIVATE | MAP_ANONYMOUS | MAP_NORESERVE, -1, 0));
    if (m == MAP_FAILED) return nullptr;
    if (mprotect(m + g_page, data, PROT_READ | PROT_WRITE) != 0) {
        munmap(m, len);
        return nullptr;
    }
    if (fill_byte() >= 0) std::memset(m + g_page, fill_byte(), data);
    char *p;
    if (guard_at_start()) {
        p = m + g_page;
    } else {
        const size_t r = (bytes + 255) & ~size_t(255);
        p = m + g_page + data - r;
    }
    std::lock_guard<std::mutex> lk(g_alloc_mtx);
    g_allocs[(uintptr_t)m] = Alloc{m, len, p, bytes, host, g_next_id++};
    return p;
}

hipError_t guarded_free(void *p)
{
    if (!p) return hipSuccess;
    std::lock_guard<std::mutex> lk(g_alloc_mtx);
    for (auto it = g_allocs.begin(); it != g_allocs.end(); ++it) {
        if (it->second.ptr == p) {
            munmap(it->second.map, it->second.map_len);
            g_allocs.erase(it);
            return hipSuccess;
        }
    }
    return hipErrorInvalidValue;
}

// ---------------------------------------------------------------------------
// fibers
// ---------------------------------------------------------------------------
extern "C" void sim_switch(void **save_sp, void *to_sp);
asm(R"(
.text
.globl sim_switch
.type sim_switch,@function
sim_switch:
    pushq %rbp
    pushq %rbx
    pushq %r12
    pushq %r13
    pushq %r14
    pushq %r15
    movq %rsp, (%rdi)
    movq %rsi, %rsp
    popq %r15
    popq %r14
    popq %r13
    popq %r12
    popq %rbx
    popq %rbp
    ret
.size sim_switch, .-sim_switch
)");

enum FState { RUN = 0, BAR_BLOCK = 1, BAR_WAVE = 2, DONE = 3 };

constexpr int kMaxThreads = 1024;
constexpr size_t kStack = 512 * 1024;
constexpr int kWaveBytes = 64;

struct Fiber {
    void *sp = nullptr;
    char *stack = nullptr;  // mapping (lowest page is a guard)
    int state = DONE;
};

Fiber g_fibers[kMaxThreads];
void *g_sched_sp = nullptr;
int g_cur = -1;
int g_nthreads = 0;
const std::function<void()> *g_body = nullptr;
const char *g_kernel = "(none)";
alignas(16) unsigned char g_wave_buf[kMaxThreads / 64][2][64 * kWaveBytes];
unsigned g_wave_gen[kMaxThreads / 64];
size_t g_wave_bytes[kMaxThreads / 64];
alignas(256) unsigned char g_dyn_smem[160 * 1024];

void fiber_main()
{
    (*g_body)();
    g_fibers[g_cur].state = DONE;
    sim_switch(&g_fibers[g_cur].sp, g_sched_sp);
    __builtin_trap();  // never resumed
}

void fiber_init(int i)
{
    Fiber &f = g_fibers[i];
    if (!f.stack) {
        char *m = static_cast<char *>(mmap(nullptr, kStack, PROT_READ | PROT_WRITE, MAP_PRIVATE | MAP_ANONYMOUS, -1, 0));
        if (m == MAP_FAILED) {
            fprintf(stderr, "sim: cannot map a fiber stack\n");
            _exit(88);
        }
        mprotect(m, g_page, PROT_NONE);  // stack overflow traps
        f.stack = m;
    }
    uintptr_t top = (uintptr_t)(f.stack + kStack) & ~uintptr_t(15);
    void **sp = reinterpret_cast<void **>(top);
    *--sp = nullptr;                              // return address slot of fiber_main
    *--sp = reinterpret_cast<void *>(&fiber_main);  // sim_switch returns here
    for (int r = 0; r < 6; r++) *--sp = nullptr;  // rbp rbx r12 r13 r14 r15
    f.sp = sp;
    f.state = RUN;
}

void set_thread(int i)
{
    const unsigned bx = sim::g_bdim.x, by = sim::g_bdim.y;
    sim::g_thread.x = (unsigned)i % bx;
    sim::g_thread.y = ((unsigned)i / bx) % by;
    sim::g_thread.z = (unsigned)i / (bx * by);
}

void park(int state)
{
    g_fibers[g_cur].state = state;
    sim_switch(&g_fibers[g_cur].sp, g_sched_sp);
}

void report_deadlock()
{
    int counts[4] = {0, 0, 0, 0};
    for (int i = 0; i < g_nthreads; i++) counts[g_fibers[i].state]++;
    fprintf(stderr,
            "sim: DEADLOCK in kernel %s, block (%u,%u,%u): %d threads at __syncthreads, %d at a wave "
            "collective, %d exited\n",
            g_kernel, sim::g_block.x, sim::g_block.y, sim::g_block.z, counts[BAR_BLOCK], counts[BAR_WAVE],
            counts[DONE]);
    _exit(87);
}

void run_workgroup(int n)
{
    g_nthreads = n;
    const int waves = (n + 63) / 64;
    for (int w = 0; w < waves; w++) g_wave_gen[w] = 0;
    for (int i = 0; i < n; i++) fiber_init(i);
    for (;;) {
        for (int i = 0; i < n; i++) {
            if (g_fibers[i].state != RUN) continue;
            g_cur = i;
            set_thread(i);
            sim_switch(&g_sched_sp, g_fibers[i].sp);
        }
        g_cur = -1;
        int done = 0, at_block = 0;
        for (int i = 0; i < n; i++) {
            done += g_fibers[i].state == DONE;
            at_block += g_fibers[i].state == BAR_BLOCK;
        }
        if (done == n) return;
        bool released = false;
        for (int w = 0; w < waves; w++) {
            const int l0 = 64 * w, l1 = std::min(n, l0 + 64);
            int waiting = 0, live = 0;
            for (int i = l0; i < l1; i++) {
                live += g_fibers[i].state != DONE;
                waiting += g_fibers[i].state == BAR_WAVE;
            }
            if (waiting > 0 && waiting == live) {
                for (int i = l0; i < l1; i++) g_fibers[i].state = g_fibers[i].state == BAR_WAVE ? RUN : DONE;
                g_wave_gen[w]++;
                released = true;
            }
        }
        if (!released && at_block == n - done) {
            for (int i = 0; i < n; i++)
                if (g_fibers[i].state == BAR_BLOCK) g_fibers[i].state = RUN;
            released = true;
        }
        if (!released) report_deadlock();
    }
}

// ---------------------------------------------------------------------------
// trap handler
// ---------------------------------------------------------------------------
void on_segv(int sig, siginfo_t *si, void *)
{
    const uintptr_t a = (uintptr_t)si->si_addr;
    fprintf(stderr, "sim: FAULT (signal %d) at address %#lx in kernel %s, block (%u,%u,%u), thread %d\n", sig,
            (unsigned long)a, g_kernel, sim::g_block.x, sim::g_block.y, sim::g_block.z, g_cur);
    for (auto &kv : g_allocs) {
        const Alloc &al = kv.second;
        if (a >= (uintptr_t)al.map && a < (uintptr_t)al.map + al.map_len) {
            fprintf(stderr, "sim:   in the guard of %s allocation #%lu (%zu bytes at %p): offset %+ld from its start\n",
                    al.host ? "pinned host" : "device", (unsigned long)al.id, al.bytes, (void *)al.ptr,
                    (long)(a - (uintptr_t)al.ptr));
        }
    }
    for (int i = 0; i < kMaxThreads; i++) {
        const Fiber &f = g_fibers[i];
        if (f.stack && a >= (uintptr_t)f.stack && a < (uintptr_t)f.stack + g_page)
            fprintf(stderr, "sim:   stack overflow of thread %d\n", i);
    }
    fflush(stderr);
    _exit(86);
}

struct Init {
    Init()
    {
        g_page = (size_t)sysconf(_SC_PAGESIZE);
        static char altstack[1 << 16];
        stack_t ss{};
        ss.ss_sp = altstack;
        ss.ss_size = sizeof(altstack);
        sigaltstack(&ss, nullptr);
        struct sigaction sa{};
        sa.sa_sigaction = on_segv;
        sa.sa_flags = SA_SIGINFO | SA_ONSTACK;
        sigaction(SIGSEGV, &sa, nullptr);
        sigaction(SIGBUS, &sa, nullptr);
    }
} g_init;

// ---------------------------------------------------------------------------
// streams, events, graphs
// ---------------------------------------------------------------------------
struct Launch {
    std::string name;
    dim3 grid, block;
    size_t smem;
    std::function<void()> body;
};

}  // namespace

struct ihipStream_t {
    int device = 0;
    bool capturing = false;
    sim_graph *graph = nullptr;
};
struct sim_event {
    double t = 0.0;
};
struct sim_graph {
    std::vector<Launch> launches;
};

namespace {
double now_ms()
{
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}
}  // namespace

namespace sim {
Idx g_thread{0, 0, 0}, g_block{0, 0, 0}, g_bdim{1, 1, 1}, g_gdim{1, 1, 1};

static void run_launch(const char *name, dim3 grid, dim3 block, const std::function<void()> &body)
{
    const int n = (int)(block.x * block.y * block.z);
    if (n <= 0 || n > kMaxThreads) {
        fprintf(stderr, "sim: kernel %s launched with %d threads per block\n", name, n);
        _exit(88);
    }
    if (g_cur >= 0) {
        fprintf(stderr, "sim: launch from device code\n");
        _exit(88);
    }
    const char *saved_name = g_kernel;
    const std::function<void()> *saved_body = g_body;
    g_kernel = name;
    g_body = &body;
    g_bdim = Idx{block.x, block.y, block.z};
    g_gdim = Idx{grid.x, grid.y, grid.z};
    for (unsigned z = 0; z < grid.z; z++)
        for (unsigned y = 0; y < grid.y; y++)
            for (unsigned x = 0; x < grid.x; x++) {
                g_block = Idx{x, y, z};
                run_workgroup(n);
            }
    g_kernel = saved_name;
    g_body = saved_body;
}

void launch(const char *name, dim3 grid, dim3 block, size_t smem, hipStream_t s, hipEvent_t e0, hipEvent_t e1,
            const std::function<void()> &body)
{
    if (smem > sizeof(g_dyn_smem)) {
        fprintf(stderr, "sim: kernel %s asks for %zu bytes of dynamic LDS\n", name, smem);
        _exit(88);
    }
    if (s && s->capturing) {
        s->graph->launches.push_back(Launch{name, grid, block, smem, body});
        return;
    }
    if (e0) e0->t = now_ms();
    run_launch(name, grid, block, body);
    if (e1) e1->t = now_ms();
}

void syncthreads()
{
    if (g_cur < 0) {
        fprintf(stderr, "sim: __syncthreads outside a kernel\n");
        _exit(88);
    }
    park(BAR_BLOCK);
}

void *dynamic_smem() { return g_dyn_smem; }

int lane_id() { return g_cur & 63; }

void wave_barrier()
{
    if (g_cur < 0) {
        fprintf(stderr, "sim: wave barrier outside a kernel\n");
        _exit(88);
    }
    park(BAR_WAVE);
}

const void *wave_exchange(const void *in, size_t bytes)
{
    if (bytes > (size_t)kWaveBytes) {
        fprintf(stderr, "sim: wave collective of %zu bytes\n", bytes);
        _exit(88);
    }
    const int w = g_cur / 64, lane = g_cur & 63;
    const int par = g_wave_gen[w] & 1;
    unsigned char *buf = g_wave_buf[w][par];
    g_wave_bytes[w] = bytes;
    std::memcpy(buf + (size_t)lane * bytes, in, bytes);
    // lanes that have exited contribute zeros
    const int l0 = 64 * w;
    for (int l = 0; l < 64; l++) {
        const int i = l0 + l;
        if (i >= g_nthreads || g_fibers[i].state == DONE) std::memset(buf + (size_t)l * bytes, 0, bytes);
    }
    park(BAR_WAVE);
    return buf;
}
}  // namespace sim

// ---------------------------------------------------------------------------
// runtime API
// ---------------------------------------------------------------------------
hipError_t hipGetDeviceCount(int *n)
{
    *n = 1;
    return hipSuccess;
}
hipError_t hipGetDevice(int *d)
{
    *d = 0;
    return hipSuccess;
}
hipError_t hipSetDevice(int d) { return d == 0 ? hipSuccess : hipErrorInvalidDevice; }
hipError_t hipStreamGetDevice(hipStream_t, hipDevice_t *d)
{
    *d = 0;
    return hipSuccess;
}
struct ihipCtx_t {
    int device = 0;
};
static ihipCtx_t g_ctx;
hipError_t hipCtxGetCurrent(hipCtx_t *c)
{
    *c = &g_ctx;
    return hipSuccess;
}
hipError_t hipCtxPushCurrent(hipCtx_t c) { return c == &g_ctx ? hipSuccess : hipErrorInvalidValue; }
hipError_t hipCtxPopCurrent(hipCtx_t *c)
{
    if (c) *c = &g_ctx;
    return hipSuccess;
}
hipError_t hipCtxGetDevice(hipDevice_t *d)
{
    *d = 0;
    return hipSuccess;
}
hipError_t hipMalloc(void **p, size_t bytes)
{
    *p = guarded_alloc(bytes, false);
    return *p ? hipSuccess : hipErrorOutOfMemory;
}
hipError_t hipFree(void *p) { return guarded_free(p); }
hipError_t hipHostMalloc(void **p, size_t bytes, unsigned)
{
    *p = guarded_alloc(bytes, true);
    return *p ? hipSuccess : hipErrorOutOfMemory;
}
hipError_t hipHostFree(void *p) { return guarded_free(p); }
hipError_t hipMemcpy(void *dst, const void *src, size_t bytes, hipMemcpyKind)
{
    if (bytes) std::memmove(dst, src, bytes);
    return hipSuccess;
}
hipError_t hipMemcpyAsync(void *dst, const void *src, size_t bytes, hipMemcpyKind k, hipStream_t s)
{
    if (s && s->capturing) return hipErrorNotSupported;
    return hipMemcpy(dst, src, bytes, k);
}
hipError_t hipMemsetAsync(void *dst, int v, size_t bytes, hipStream_t)
{
    if (bytes) std::memset(dst, v, bytes);
    return hipSuccess;
}
hipError_t hipStreamCreateWithFlags(hipStream_t *s, unsigned)
{
    *s = new ihipStream_t();
    return hipSuccess;
}
hipError_t hipStreamDestroy(hipStream_t s)
{
    delete s;
    return hipSuccess;
}
hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }
hipError_t hipStreamWaitEvent(hipStream_t, hipEvent_t, unsigned) { return hipSuccess; }
hipError_t hipDeviceGetAttribute(int *v, hipDeviceAttribute_t attr, int)
{
    if (attr != hipDeviceAttributeMultiprocessorCount) return hipErrorInvalidValue;
    *v = 4;
    return hipSuccess;
}
hipError_t hipOccupancyMaxActiveBlocksPerMultiprocessor(int *n, const void *, int, size_t)
{
    *n = 2;
    return hipSuccess;
}
hipError_t hipEventCreate(hipEvent_t *e)
{
    *e = new sim_event();
    return hipSuccess;
}
hipError_t hipEventCreateWithFlags(hipEvent_t *e, unsigned) { return hipEventCreate(e); }
hipError_t hipEventDestroy(hipEvent_t e)
{
    delete e;
    return hipSuccess;
}
hipError_t hipEventRecord(hipEvent_t e, hipStream_t)
{
    if (e) e->t = now_ms();
    return hipSuccess;
}
hipError_t hipEventSynchronize(hipEvent_t) { return hipSuccess; }
hipError_t hipEventQuery(hipEvent_t) { return hipSuccess; }  // launches run to completion
hipError_t hipEventElapsedTime(float *ms, hipEvent_t a, hipEvent_t b)
{
    *ms = (float)(b->t - a->t);
    return hipSuccess;
}
hipError_t hipGetLastError() { return hipSuccess; }
const char *hipGetErrorString(hipError_t) { return "sim error"; }
hipError_t hipDeviceSynchronize() { return hipSuccess; }
hipError_t hipStreamBeginCapture(hipStream_t s, hipStreamCaptureMode)
{
    s->capturing = true;
    s->graph = new sim_graph();
    return hipSuccess;
}
hipError_t hipStreamEndCapture(hipStream_t s, hipGraph_t *g)
{
    s->capturing = false;
    *g = s->graph;
    s->graph = nullptr;
    return hipSuccess;
}
hipError_t hipGraphInstantiate(hipGraphExec_t *e, hipGraph_t g, void *, void *, size_t)
{
    *e = new sim_graph(*g);
    return hipSuccess;
}
hipError_t hipGraphLaunch(hipGraphExec_t e, hipStream_t)
{
    for (auto &l : e->launches) sim::run_launch(l.name.c_str(), l.grid, l.block, l.body);
    return hipSuccess;
}
hipError_t hipGraphDestroy(hipGraph_t g)
{
    delete g;
    return hipSuccess;
}
hipError_t hipGraphExecDestroy(hipGraphExec_t e)
{
    delete e;
    return hipSuccess;
}
hipError_t hipGetSymbolAddress(void **p, const void *symbol)
{
    *p = const_cast<void *>(symbol);
    return hipSuccess;
}

// test helpers: guarded "device" buffers for inputs and outputs
extern "C" void *sim_device_alloc(size_t bytes) { return guarded_alloc(bytes, false); }
extern "C" int sim_device_free(void *p) { return guarded_free(p); }
extern "C" size_t sim_live_allocations()
{
    std::lock_guard<std::mutex> lk(g_alloc_mtx);
    return g_allocs.size();
}
